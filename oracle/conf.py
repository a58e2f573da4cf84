"""Restatement of pmwd ``Configuration`` (pmwd/configuration.py:21-325).

Oracle-side only (test infrastructure).  Holds the fields and derived
properties the hot path and the cosmology tables need, computed with the same
formulas as the reference.
"""
import math
from dataclasses import dataclass, field, replace
from typing import Tuple

import numpy as np


@dataclass(frozen=True)
class OConf:
    """Configuration restated (configuration.py:101-146 defaults)."""
    ptcl_spacing: float
    ptcl_grid_shape: Tuple[int, ...]
    mesh_shape: object = 1
    cosmo_dtype: object = np.float64
    pmid_dtype: object = np.int16
    float_dtype: object = np.float32
    k_pivot_Mpc: float = 0.05
    T_cmb: float = 2.7255
    transfer_fit: bool = True
    transfer_fit_nowiggle: bool = False
    transfer_lgk_min: float = -4
    transfer_lgk_max: float = 3
    transfer_lgk_maxstep: float = 1 / 128
    growth_rtol: float = None
    growth_atol: float = None
    growth_inistep: object = (1, None)
    lpt_order: int = 2
    a_start: float = 1 / 64
    a_stop: float = 1
    a_lpt_maxstep: float = 1 / 128
    a_nbody_maxstep: float = 1 / 64
    symp_splits: tuple = ((0, 0.5), (1, 0.5))
    chunk_size: int = 2**24

    # constants (configuration.py:110-122)
    M_sun_SI = 1.98847e30
    Mpc_SI = 3.0856775815e22
    H_0_SI = 1e5 / 3.0856775815e22
    c_SI = 299792458
    G_SI = 6.67430e-11

    def __post_init__(self):
        # configuration.py:151-153
        if isinstance(self.mesh_shape, (int, float)):
            ms = tuple(round(s * self.mesh_shape) for s in self.ptcl_grid_shape)
            object.__setattr__(self, 'mesh_shape', ms)
        object.__setattr__(self, 'mesh_shape', tuple(int(s) for s in self.mesh_shape))
        tol = math.sqrt(np.finfo(self.cosmo_dtype).eps)  # configuration.py:180-185
        if self.growth_rtol is None:
            object.__setattr__(self, 'growth_rtol', tol)
        if self.growth_atol is None:
            object.__setattr__(self, 'growth_atol', tol)

    def replace(self, **kw):
        return replace(self, **kw)

    # units (configuration.py:119-122, 241-263)
    @property
    def M(self):
        return 1e10 * self.M_sun_SI

    @property
    def L(self):
        return self.Mpc_SI

    @property
    def T(self):
        return 1 / self.H_0_SI

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
        # configuration.py:227-229
        return self.ptcl_spacing * self.ptcl_grid_shape[0] / self.mesh_shape[0]

    @property
    def mesh_size(self):
        return int(np.prod(self.mesh_shape))

    # tables (configuration.py:265-325)
    @property
    def transfer_k_num(self):
        return 1 + math.ceil((self.transfer_lgk_max - self.transfer_lgk_min)
                             / self.transfer_lgk_maxstep) + 1

    @property
    def transfer_k(self):
        k = np.logspace(self.transfer_lgk_min, self.transfer_lgk_max,
                        num=self.transfer_k_num - 1, dtype=self.cosmo_dtype)
        return np.concatenate((np.array([0.0]), k))

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
    def a_lpt(self):
        return np.linspace(0, self.a_start, num=self.a_lpt_num + 1, dtype=self.cosmo_dtype)

    @property
    def a_nbody(self):
        return np.linspace(self.a_start, self.a_stop, num=1 + self.a_nbody_num,
                           dtype=self.cosmo_dtype)

    @property
    def growth_a(self):
        return np.concatenate((self.a_lpt, self.a_nbody[1:]))

    @property
    def varlin_R(self):
        """Radii of the top-hat variance table.

        The reference takes them from ``mcfit.TophatVar(transfer_k[1:]).y``
        (configuration.py:173-178, 322-325), an FFTLog grid reciprocal to the
        wavenumbers.  mcfit is absent here, so the oracle uses the reciprocal
        of the same log grid (parity unpinned against mcfit).
        """
        return 1.0 / self.transfer_k[1:][::-1]
