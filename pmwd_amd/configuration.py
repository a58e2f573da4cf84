"""Configuration: static parameters of a simulation (pmwd/configuration.py:21-325).

Same field names, defaults, validation and derived properties as pmwd's
frozen ``Configuration``.  Differences: no JAX pytree plumbing (a frozen
dataclass with ``replace``), and the top-hat variance grid ``varlin_R`` is the
reciprocal of the transfer wavenumber grid because the FFTLog provider
(mcfit) is not used.  Floats are Python floats (FP64 on the host).
"""
from dataclasses import dataclass, replace as _dc_replace
from functools import cached_property
import math
from typing import ClassVar, Optional, Tuple, Union

import numpy as np
import torch


_DT = {
    'float16': torch.float16, 'float32': torch.float32, 'float64': torch.float64,
    'int8': torch.int8, 'int16': torch.int16, 'int32': torch.int32, 'int64': torch.int64,
}


def _as_torch_dtype(dt):
    if isinstance(dt, torch.dtype):
        return dt
    name = np.dtype(dt).name
    if name not in _DT:
        raise ValueError(f'unsupported dtype {dt}')
    return _DT[name]


def _frozen(a):
    """A read-only array: derived tables are computed once per
    configuration and shared by every caller."""
    a.flags.writeable = False
    return a


@dataclass(frozen=True)
class Configuration:
    """Configuration parameters, immutable (configuration.py:21-196).

    Parameters follow pmwd: ``ptcl_spacing``, ``ptcl_grid_shape``,
    ``mesh_shape`` (int/float ratio or tuple), dtypes, units, transfer and
    growth settings, ``lpt_order``, ``a_start``, ``a_stop``,
    ``a_lpt_maxstep``, ``a_nbody_maxstep``, ``symp_splits``, ``chunk_size``.

    Raises
    ------
    ValueError
        Incorrect or inconsistent parameter values (configuration.py:155-191).
    """

    ptcl_spacing: float
    ptcl_grid_shape: Tuple[int, ...]

    mesh_shape: Union[float, Tuple[int, ...]] = 1

    cosmo_dtype: object = torch.float64
    pmid_dtype: object = torch.int16
    float_dtype: object = torch.float32

    k_pivot_Mpc: float = 0.05
    T_cmb: float = 2.7255

    M_sun_SI: ClassVar[float] = 1.98847e30
    Mpc_SI: ClassVar[float] = 3.0856775815e22
    H_0_SI: ClassVar[float] = 1e5 / Mpc_SI
    c_SI: ClassVar[int] = 299792458
    G_SI: ClassVar[float] = 6.67430e-11

    M: float = 1e10 * M_sun_SI
    L: float = Mpc_SI
    T: float = 1 / H_0_SI

    transfer_fit: bool = True
    transfer_fit_nowiggle: bool = False
    transfer_lgk_min: float = -4
    transfer_lgk_max: float = 3
    transfer_lgk_maxstep: float = 1 / 128

    growth_rtol: Optional[float] = None
    growth_atol: Optional[float] = None
    growth_inistep: Union[float, None, Tuple[Optional[float], Optional[float]]] = (1, None)

    lpt_order: int = 2

    a_start: float = 1 / 64
    a_stop: float = 1
    a_lpt_maxstep: float = 1 / 128
    a_nbody_maxstep: float = 1 / 64

    symp_splits: Tuple[Tuple[float, float], ...] = ((0, 0.5), (1, 0.5))

    chunk_size: int = 2**24

    def __post_init__(self):
        set_ = lambda k, v: object.__setattr__(self, k, v)
        set_('ptcl_grid_shape', tuple(int(s) for s in self.ptcl_grid_shape))
        if isinstance(self.mesh_shape, (int, float)):
            set_('mesh_shape', tuple(round(s * self.mesh_shape) for s in self.ptcl_grid_shape))
        set_('mesh_shape', tuple(int(s) for s in self.mesh_shape))
        if len(self.ptcl_grid_shape) != len(self.mesh_shape):
            raise ValueError('particle and mesh grid dimensions differ')
        if any(sm < sp for sp, sm in zip(self.ptcl_grid_shape, self.mesh_shape)):
            raise ValueError('mesh grid cannot be smaller than particle grid')
        if any(self.ptcl_grid_shape[0] * sm != self.mesh_shape[0] * sp
               for sp, sm in zip(self.ptcl_grid_shape[1:], self.mesh_shape[1:])):
            raise ValueError('particle and mesh grid aspect ratios differ')

        set_('cosmo_dtype', _as_torch_dtype(self.cosmo_dtype))
        set_('pmid_dtype', _as_torch_dtype(self.pmid_dtype))
        set_('float_dtype', _as_torch_dtype(self.float_dtype))
        if not self.cosmo_dtype.is_floating_point:
            raise ValueError('cosmo_dtype must be floating point numbers')
        if self.pmid_dtype.is_floating_point or not self.pmid_dtype.is_signed:
            raise ValueError('pmid_dtype must be signed integers')
        if not self.float_dtype.is_floating_point:
            raise ValueError('float_dtype must be floating point numbers')

        growth_tol = math.sqrt(torch.finfo(self.cosmo_dtype).eps)
        if self.growth_rtol is None:
            set_('growth_rtol', growth_tol)
        if self.growth_atol is None:
            set_('growth_atol', growth_tol)

        if any(len(s) != 2 for s in self.symp_splits):
            raise ValueError(f'symp_splits={self.symp_splits} not supported')
        symp_splits_sum = tuple(sum(s) for s in zip(*self.symp_splits))
        if symp_splits_sum != (1, 1):
            raise ValueError(f'sum of symplectic splits = {symp_splits_sum} != (1, 1)')

    def replace(self, **changes):
        """Copy with some fields replaced (re-runs validation)."""
        return _dc_replace(self, **changes)

    def __hash__(self):
        return hash((self.ptcl_spacing, self.ptcl_grid_shape, self.mesh_shape, self.a_start,
                     self.a_stop, self.a_nbody_maxstep, self.a_lpt_maxstep, self.symp_splits,
                     self.lpt_order, str(self.float_dtype), str(self.cosmo_dtype)))

    @property
    def dim(self):
        return len(self.ptcl_grid_shape)

    @property
    def ptcl_cell_vol(self):
        return self.ptcl_spacing ** self.dim

    @property
    def ptcl_num(self):
        return int(np.prod(self.ptcl_grid_shape))

    @property
    def box_size(self):
        return tuple(self.ptcl_spacing * s for s in self.ptcl_grid_shape)

    @property
    def box_vol(self):
        return float(np.prod(self.box_size))

    @property
    def cell_size(self):
        return self.ptcl_spacing * self.ptcl_grid_shape[0] / self.mesh_shape[0]

    @property
    def cell_vol(self):
        return self.cell_size ** self.dim

    @property
    def mesh_size(self):
        return int(np.prod(self.mesh_shape))

    @property
    def V(self):
        return self.L / self.T

    @property
    def H_0(self):
        return self.H_0_SI * self.T

    @property
    def c(self):
        return self.c_SI / self.V

    @property
    def G(self):
        return self.G_SI * self.M / (self.L * self.V**2)

    @property
    def rho_crit(self):
        return 3 * self.H_0**2 / (8 * math.pi * self.G)

    @property
    def transfer_k_num(self):
        return 1 + math.ceil((self.transfer_lgk_max - self.transfer_lgk_min)
                             / self.transfer_lgk_maxstep) + 1

    @property
    def transfer_lgk_step(self):
        return (self.transfer_lgk_max - self.transfer_lgk_min) / (self.transfer_k_num - 2)

    @cached_property
    def transfer_k(self):
        k = np.logspace(self.transfer_lgk_min, self.transfer_lgk_max,
                        num=self.transfer_k_num - 1, dtype=np.float64)
        return _frozen(np.concatenate((np.array([0.0]), k)))

    @property
    def a_lpt_num(self):
        return math.ceil(self.a_start / self.a_lpt_maxstep)

    @property
    def a_lpt_step(self):
        return self.a_start / self.a_lpt_num

    @property
    def a_nbody_num(self):
        return math.ceil((self.a_stop - self.a_start) / self.a_nbody_maxstep)

    @property
    def a_nbody_step(self):
        return (self.a_stop - self.a_start) / self.a_nbody_num

    @cached_property
    def a_lpt(self):
        return _frozen(np.linspace(0, self.a_start, num=self.a_lpt_num + 1, dtype=np.float64))

    @cached_property
    def a_nbody(self):
        return _frozen(np.linspace(self.a_start, self.a_stop, num=1 + self.a_nbody_num,
                                   dtype=np.float64))

    @cached_property
    def growth_a(self):
        return _frozen(np.concatenate((self.a_lpt, self.a_nbody[1:])))

    @cached_property
    def varlin_R(self):
        """Top-hat radii of the variance table: reciprocal transfer wavenumbers."""
        return _frozen(1.0 / self.transfer_k[1:][::-1])
