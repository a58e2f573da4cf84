"""Cosmology parameters and background functions (pmwd/cosmology.py:15-269).

Parameters are float64 host scalars.  Gradients with respect to them are
returned by the adjoint entry points as plain floats (``CosmoCot``), since
the tables that depend on them are built on the host in FP64.
"""
from dataclasses import dataclass, field, replace as _dc_replace, fields
import math
from typing import ClassVar, Optional

import numpy as np

from .configuration import Configuration


@dataclass(frozen=True)
class Cosmology:
    """Cosmological parameters plus the transfer/growth/varlin tables.

    Parameters with trailing underscores may be None and then take the
    ``*_fixed`` values (cosmology.py:15-58).  Linear operators act on the
    parameter fields, as for pmwd's tangent/cotangent vectors.
    """

    conf: Configuration = field(repr=False)

    A_s_1e9: float
    n_s: float
    Omega_m: float
    Omega_b: float
    h: float

    Omega_k_: Optional[float] = None
    Omega_k_fixed: ClassVar[float] = 0
    w_0_: Optional[float] = None
    w_0_fixed: ClassVar[float] = -1
    w_a_: Optional[float] = None
    w_a_fixed: ClassVar[float] = 0

    transfer: Optional[np.ndarray] = field(default=None, compare=False, repr=False)
    growth: Optional[np.ndarray] = field(default=None, compare=False, repr=False)
    growth_dOm: Optional[np.ndarray] = field(default=None, compare=False, repr=False)
    varlin: Optional[np.ndarray] = field(default=None, compare=False, repr=False)

    _params: ClassVar[tuple] = ('A_s_1e9', 'n_s', 'Omega_m', 'Omega_b', 'h',
                                'Omega_k_', 'w_0_', 'w_a_')

    def __post_init__(self):
        for name in self._params:
            v = getattr(self, name)
            if v is not None:
                object.__setattr__(self, name, float(v))

    def replace(self, **changes):
        return _dc_replace(self, **changes)

    def _map(self, fn, other=None):
        kw = {}
        for name in self._params:
            a = getattr(self, name)
            b = getattr(other, name) if other is not None else None
            if a is None:
                kw[name] = None
            else:
                kw[name] = fn(a, b)
        return self.replace(transfer=None, growth=None, growth_dOm=None, varlin=None, **kw)

    def __add__(self, other):
        return self._map(lambda a, b: a + (b or 0.0), other)

    def __sub__(self, other):
        return self._map(lambda a, b: a - (b or 0.0), other)

    def __mul__(self, s):
        return self._map(lambda a, b: a * s)

    __rmul__ = __mul__

    @classmethod
    def from_sigma8(cls, conf, sigma8, *args, **kwargs):
        """Construct with sigma8 instead of A_s (cosmology.py:95-105)."""
        from .boltzmann import boltzmann
        cosmo = cls(conf, 1, *args, **kwargs)
        # sigma8 reads the transfer and growth tables and two nodes of the
        # variance table, not the whole variance table
        cosmo = boltzmann(cosmo, conf, varlin=False)
        A_s_1e9 = (sigma8 / cosmo.sigma8)**2
        return cls(conf, A_s_1e9, *args, **kwargs)

    def astype(self, dtype):
        conf = self.conf.replace(cosmo_dtype=dtype)
        return self.replace(conf=conf)

    @property
    def k_pivot(self):
        return self.conf.k_pivot_Mpc / (self.h * self.conf.Mpc_SI) * self.conf.L

    @property
    def A_s(self):
        return self.A_s_1e9 * 1e-9

    @property
    def Omega_c(self):
        return self.Omega_m - self.Omega_b

    @property
    def Omega_k(self):
        return self.Omega_k_fixed if self.Omega_k_ is None else self.Omega_k_

    @property
    def Omega_de(self):
        return 1 - (self.Omega_m + self.Omega_k)

    @property
    def w_0(self):
        return self.w_0_fixed if self.w_0_ is None else self.w_0_

    @property
    def w_a(self):
        return self.w_a_fixed if self.w_a_ is None else self.w_a_

    @property
    def sigma8(self):
        """sqrt(varlin(8 Mpc/h, a=1)): the variance table's interpolation at
        8 Mpc/h, from its two nodes around it (the table itself need not be
        built)."""
        from .boltzmann import growth, sigma2_at
        if self.transfer is None or self.growth is None:
            raise ValueError('sigma8 needs the transfer and growth tables: call boltzmann first.')
        R = 8 * self.conf.Mpc_SI / self.conf.L
        s2 = sigma2_at(R, self, self.conf) * float(growth(1.0, self, self.conf))**2
        return float(np.sqrt(s2))

    @property
    def ptcl_mass(self):
        return self.conf.rho_crit * self.Omega_m * self.conf.ptcl_cell_vol


def SimpleLCDM(conf, **kw):
    """Simple LCDM (cosmology.py:164-172)."""
    p = dict(A_s_1e9=2.0, n_s=0.96, Omega_m=0.3, Omega_b=0.05, h=0.7)
    p.update(kw)
    return Cosmology(conf, **p)


def Planck18(conf, **kw):
    """Planck 2018 (cosmology.py:174-182)."""
    p = dict(A_s_1e9=2.105, n_s=0.9665, Omega_m=0.3111, Omega_b=0.04897, h=0.6766)
    p.update(kw)
    return Cosmology(conf, **p)


def _de(a, cosmo):
    p = -3 * (1 + cosmo.w_0 + cosmo.w_a)
    return a**p * np.exp(-3 * cosmo.w_a * (1 - a)), p


def E2(a, cosmo):
    """Squared Hubble scaling E^2(a) (cosmology.py:185-219)."""
    a = np.asarray(a, dtype=np.float64)
    de_a, _ = _de(a, cosmo)
    return cosmo.Omega_m * a**-3 + cosmo.Omega_k * a**-2 + cosmo.Omega_de * de_a


def E2_da(a, cosmo):
    """dE^2/da."""
    a = np.asarray(a, dtype=np.float64)
    de_a, p = _de(a, cosmo)
    dde = (p / a + 3 * cosmo.w_a) * de_a
    return -3 * cosmo.Omega_m * a**-4 - 2 * cosmo.Omega_k * a**-3 + cosmo.Omega_de * dde


def H_deriv(a, cosmo):
    """dlnH/dlna = a E2'/(2 E2) (cosmology.py:222-242)."""
    a = np.asarray(a, dtype=np.float64)
    return 0.5 * a * E2_da(a, cosmo) / E2(a, cosmo)


def Omega_m_a(a, cosmo):
    """Omega_m(a) = Omega_m a^-3 / E2 (cosmology.py:245-269)."""
    a = np.asarray(a, dtype=np.float64)
    return cosmo.Omega_m / (a**3 * E2(a, cosmo))


# derivatives w.r.t. Omega_m at fixed Omega_k, w_0, w_a (Omega_de = 1-Om-Ok)
def E2_dOm(a, cosmo):
    a = np.asarray(a, dtype=np.float64)
    de_a, _ = _de(a, cosmo)
    return a**-3 - de_a


def E2_da_dOm(a, cosmo):
    a = np.asarray(a, dtype=np.float64)
    de_a, p = _de(a, cosmo)
    return -3 * a**-4 - (p / a + 3 * cosmo.w_a) * de_a


def H_deriv_dOm(a, cosmo):
    a = np.asarray(a, dtype=np.float64)
    e2, e2a = E2(a, cosmo), E2_da(a, cosmo)
    return 0.5 * a * (E2_da_dOm(a, cosmo) * e2 - e2a * E2_dOm(a, cosmo)) / e2**2


def Omega_m_a_dOm(a, cosmo):
    a = np.asarray(a, dtype=np.float64)
    e2 = E2(a, cosmo)
    return a**-3 * (e2 - cosmo.Omega_m * E2_dOm(a, cosmo)) / e2**2
