"""Transfer function, growth and variance tables (pmwd/boltzmann.py), FP64 host.

These tables are computed once per cosmology on the host in float64, as the
reference does (its ``cosmo_dtype``).  ``growth_integ`` additionally
integrates the forward sensitivity of the growth ODE to Omega_m on the same
steps, which the N-body adjoint uses for the cosmology cotangent.  The
top-hat variance is a Simpson quadrature in ln k (pmwd uses mcfit's FFTLog).
The growth ODE and the variance integral run in the library's host code
(``pmx_growth_ode``, ``pmx_tophat_var``), a few milliseconds per cosmology.
"""
import ctypes

import numpy as np



def transfer_fit(k, cosmo, conf):
    """Eisenstein & Hu (1998) matter transfer function (boltzmann.py:32-123)."""
    k = np.asarray(k, dtype=np.float64) * cosmo.h / conf.L * conf.Mpc_SI  # [1/Mpc]
    theta2 = (conf.T_cmb / 2.7)**2
    wm = cosmo.Omega_m * cosmo.h**2
    wb = cosmo.Omega_b * cosmo.h**2
    fb = cosmo.Omega_b / cosmo.Omega_m
    fc = cosmo.Omega_c / cosmo.Omega_m

    z_eq = 2.50e4 * wm / theta2**2
    k_eq = 7.46e-2 * wm / theta2
    b1 = 0.313 * wm**-0.419 * (1 + 0.607 * wm**0.674)
    b2 = 0.238 * wm**0.223
    z_d = 1291 * wm**0.251 / (1 + 0.659 * wm**0.828) * (1 + b1 * wb**b2)
    R_d = 31.5 * wb / theta2**2 * (1e3 / z_d)
    R_eq = 31.5 * wb / theta2**2 * (1e3 / z_eq)
    s = (2 / (3 * k_eq) * np.sqrt(6 / R_eq)
         * np.log((np.sqrt(1 + R_d) + np.sqrt(R_eq + R_d)) / (1 + np.sqrt(R_eq))))
    k_silk = 1.6 * wb**0.52 * wm**0.73 * (1 + (10.4 * wm)**-0.95)
    ks = k * s

    if conf.transfer_fit_nowiggle:
        ag = 1 - 0.328 * np.log(431 * wm) * fb + 0.38 * np.log(22.3 * wm) * fb**2
        gamma_eff = ag + (1 - ag) / (1 + (0.43 * ks)**4)
        q = k / (13.41 * k_eq * gamma_eff)
        L0 = np.log(2 * np.e + 1.8 * q)
        C0 = 14.2 + 731 / (1 + 62.5 * q)
        return L0 / (L0 + C0 * q**2)

    a1 = (46.9 * wm)**0.670 * (1 + (32.1 * wm)**-0.532)
    a2 = (12.0 * wm)**0.424 * (1 + (45.0 * wm)**-0.582)
    alpha_c = a1**-fb * a2**-fb**3
    bb1 = 0.944 / (1 + (458 * wm)**-0.708)
    bb2 = (0.395 * wm)**-0.0266
    beta_c = 1 / (1 + bb1 * (fc**bb2 - 1))

    def t0(alpha, beta):
        q = k / (13.41 * k_eq)
        L = np.log(np.e + 1.8 * beta * q)
        C = 14.2 / alpha + 386 / (1 + 69.9 * q**1.08)
        return L / (L + C * q**2)

    f = 1 / (1 + (ks / 5.4)**4)
    T_c = f * t0(1, beta_c) + (1 - f) * t0(alpha_c, beta_c)

    y = (1 + z_eq) / (1 + z_d)
    x = np.sqrt(1 + y)
    G = y * (-6 * x + (2 + 3 * y) * np.log((x + 1) / (x - 1)))
    alpha_b = 2.07 * k_eq * s * (1 + R_d)**-0.75 * G
    beta_node = 8.41 * wm**0.435
    beta_b = 0.5 + fb + (3 - 2 * fb) * np.sqrt(1 + (17.2 * wm)**2)
    T_b = ((t0(1, 1) / (1 + (ks / 5.2)**2)
            + alpha_b * ks**3 / (beta_b**3 + ks**3) * np.exp(-(k / k_silk)**1.4))
           * np.sinc(ks**2 / (np.pi * np.cbrt(beta_node**3 + ks**3))))
    return fc * T_c + fb * T_b


def transfer_integ(cosmo, conf):
    """Tabulate the transfer function at ``conf.transfer_k`` (boltzmann.py:8-28)."""
    if not conf.transfer_fit:
        raise NotImplementedError('only the Eisenstein & Hu fit is implemented, as in pmwd')
    return cosmo.replace(transfer=transfer_fit(conf.transfer_k, cosmo, conf))


def transfer(k, cosmo, conf):
    """Linear interpolation of the transfer table (boltzmann.py:126-160)."""
    if cosmo.transfer is None:
        raise ValueError('Transfer table is empty. Call transfer_integ or boltzmann first.')
    return np.interp(np.asarray(k, dtype=np.float64), conf.transfer_k, cosmo.transfer)


def growth_integ(cosmo, conf):
    """Tabulate growth functions at ``conf.growth_a`` (boltzmann.py:163-229).

    ``cosmo.growth[order-1][deriv]`` holds D_m/a^m, D_m'/a^m, D_m''/a^m
    (primes are d/dlna); ``cosmo.growth_dOm`` their Omega_m derivatives.
    """
    eps = np.finfo(np.float64).eps
    a_ic = 0.5 * np.cbrt(eps)
    if a_ic >= conf.a_lpt_step:
        a_ic = 0.1 * conf.a_lpt_step
    a = conf.growth_a.copy()
    a[0] = a_ic
    lna = np.ascontiguousarray(np.log(a))
    y0 = np.array([1, 0, 3 / 7, 0, 0, 0, 0, 0], dtype=np.float64)
    Y = np.empty((len(lna), 8))
    Yd = np.empty((len(lna), 8))
    dt0 = conf.growth_inistep
    dt0 = dt0[0] if isinstance(dt0, tuple) else dt0
    _host_call('pmx_growth_ode', float(cosmo.Omega_m), float(cosmo.Omega_k), float(cosmo.w_0),
               float(cosmo.w_a), _dptr(lna), len(lna), _dptr(y0), float(conf.growth_rtol),
               float(conf.growth_atol), float(dt0) if dt0 is not None else 0.0, 4, _dptr(Y),
               _dptr(Yd))
    m = np.array([1.0, 2.0])[:, None]

    def table(G, Gp, Gpp):
        return np.stack((G, m * G + Gp, m**2 * G + 2 * m * Gp + Gpp), axis=1)

    G = np.stack([Y[:, 0], Y[:, 2]])
    Gp = np.stack([Y[:, 1], Y[:, 3]])
    Gpp = np.stack([Yd[:, 1], Yd[:, 3]])
    S = np.stack([Y[:, 4], Y[:, 6]])
    Sp = np.stack([Y[:, 5], Y[:, 7]])
    Spp = np.stack([Yd[:, 5], Yd[:, 7]])
    return cosmo.replace(growth=table(G, Gp, Gpp), growth_dOm=table(S, Sp, Spp))


def growth(a, cosmo, conf, order=1, deriv=0):
    """D_m^(n)(a) by interpolation of the table (boltzmann.py:233-269)."""
    if cosmo.growth is None:
        raise ValueError('Growth table is empty. Call growth_integ or boltzmann first.')
    a = np.asarray(a, dtype=np.float64)
    return a**order * np.interp(a, conf.growth_a, cosmo.growth[order - 1][deriv])


def growth_dOm(a, cosmo, conf, order=1, deriv=0):
    """d growth / d Omega_m at fixed a."""
    if cosmo.growth_dOm is None:
        raise ValueError('Growth sensitivity table is empty. Call growth_integ first.')
    a = np.asarray(a, dtype=np.float64)
    return a**order * np.interp(a, conf.growth_a, cosmo.growth_dOm[order - 1][deriv])


def linear_power(k, a, cosmo, conf):
    """Linear matter power spectrum in [L^3] (boltzmann.py:399-455)."""
    if conf.dim != 3:
        raise ValueError(f'dim={conf.dim} not supported')
    k = np.asarray(k, dtype=np.float64)
    T = transfer(k, cosmo, conf)
    with np.errstate(divide='ignore'):
        P = (0.32 * cosmo.A_s * cosmo.k_pivot * (k / cosmo.k_pivot)**cosmo.n_s
             * (np.pi * (conf.c / conf.H_0)**2 / cosmo.Omega_m * T)**2)
    if a is not None:
        P = P * growth(a, cosmo, conf)**2
    return P


def _dptr(x):
    return x.ctypes.data_as(ctypes.c_void_p)


def _host_call(name, *args):
    """A host-only function of the library (float64 tables, no device)."""
    from . import _lib
    _lib.check(getattr(_lib.lib(), name)(*args))


def _tophat_var(R, cosmo, conf):
    """sigma^2(R) of the linear power at a=1: Simpson in ln k over
    ``conf.transfer_k[1:]`` (trapezoid for an even node count)."""
    k = np.ascontiguousarray(conf.transfer_k[1:])
    P = np.ascontiguousarray(linear_power(k, None, cosmo, conf))
    R = np.ascontiguousarray(np.atleast_1d(np.asarray(R, dtype=np.float64)))
    out = np.empty(len(R))
    _host_call('pmx_tophat_var', _dptr(k), _dptr(P), len(k), _dptr(R), len(R), _dptr(out))
    return out


def sigma2_at(R, cosmo, conf):
    """np.interp(R, conf.varlin_R, varlin table) computed from the two table
    nodes around R only (the interpolation reads no other node)."""
    grid = conf.varlin_R
    i = int(np.searchsorted(grid, float(R), side='right'))
    i = min(max(i, 1), len(grid) - 1)
    v = _tophat_var(grid[i - 1:i + 1], cosmo, conf)
    return float(np.interp(float(R), grid[i - 1:i + 1], v))


def varlin_integ(cosmo, conf):
    """Tabulate sigma^2(R) at ``conf.varlin_R`` (boltzmann.py:272-291)."""
    return cosmo.replace(varlin=_tophat_var(conf.varlin_R, cosmo, conf))


def varlin(R, a, cosmo, conf):
    """sigma^2(R, a) by interpolation (boltzmann.py:294-335)."""
    if cosmo.varlin is None:
        raise ValueError('Linear matter overdensity variance table is empty. '
                         'Call varlin_integ or boltzmann first.')
    s2 = np.interp(np.asarray(R, dtype=np.float64), conf.varlin_R, cosmo.varlin)
    if a is not None:
        s2 = s2 * growth(a, cosmo, conf)**2
    return s2


def boltzmann(cosmo, conf, transfer=True, growth=True, varlin=True):
    """Precompute the tables (boltzmann.py:338-374)."""
    cosmo = transfer_integ(cosmo, conf) if transfer else cosmo.replace(transfer=None)
    cosmo = growth_integ(cosmo, conf) if growth else cosmo.replace(growth=None, growth_dOm=None)
    cosmo = varlin_integ(cosmo, conf) if varlin else cosmo.replace(varlin=None)
    return cosmo
