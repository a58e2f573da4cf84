"""Transfer function, growth and variance tables (pmwd/boltzmann.py), FP64 host.

These tables are computed once per cosmology on the host in float64, as the
reference does (its ``cosmo_dtype``).  ``growth_integ`` additionally
integrates the forward sensitivity of the growth ODE to Omega_m on the same
steps, which the N-body adjoint uses for the cosmology cotangent.  The
top-hat variance is a Simpson quadrature in ln k (pmwd uses mcfit's FFTLog).
"""
import numpy as np

from .cosmology import (E2, E2_da, H_deriv, Omega_m_a, H_deriv_dOm, Omega_m_a_dOm,
                        E2_dOm)
from .ode_util import odeint


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


def _growth_rhs(cosmo):
    """RHS of the 2LPT growth ODE in ln a plus its Omega_m sensitivity.

    State (G1, G1', G2, G2', S1, S1', S2, S2') with S = dG/dOmega_m
    (boltzmann.py:197-205 for the first four).
    """
    def f(y, lna):
        a = np.exp(lna)
        h = H_deriv(a, cosmo)
        om = 1.5 * Omega_m_a(a, cosmo)
        dh = H_deriv_dOm(a, cosmo)
        dom = 1.5 * Omega_m_a_dOm(a, cosmo)
        G1, G1p, G2, G2p, S1, S1p, S2, S2p = y
        G1pp = -(3 + h - om) * G1 - (4 + h) * G1p
        G2pp = om * G1**2 - (8 + 2 * h - om) * G2 - (6 + h) * G2p
        S1pp = (-(3 + h - om) * S1 - (4 + h) * S1p
                - (dh - dom) * G1 - dh * G1p)
        S2pp = (2 * om * G1 * S1 - (8 + 2 * h - om) * S2 - (6 + h) * S2p
                + dom * G1**2 - (2 * dh - dom) * G2 - dh * G2p)
        return np.array([G1p, G1pp, G2p, G2pp, S1p, S1pp, S2p, S2pp])
    return f


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
    lna = np.log(a)
    f = _growth_rhs(cosmo)
    y0 = np.array([1, 0, 3 / 7, 0, 0, 0, 0, 0], dtype=np.float64)
    Y = odeint(f, y0, lna, rtol=conf.growth_rtol, atol=conf.growth_atol,
               dt0=conf.growth_inistep, nerr=4)
    Yd = np.stack([f(y, t) for y, t in zip(Y, lna)])
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


def _tophat_var(R, cosmo, conf):
    k = conf.transfer_k[1:]
    P = linear_power(k, None, cosmo, conf)
    lnk = np.log(k)
    R = np.atleast_1d(np.asarray(R, dtype=np.float64))
    x = np.outer(R, k)
    small = x < 1e-3
    xs = np.where(small, 1.0, x)
    W = np.where(small, 1 - x**2 / 10, 3 * (np.sin(xs) - xs * np.cos(xs)) / xs**3)
    integrand = k**3 * P / (2 * np.pi**2) * W**2
    n = integrand.shape[1]
    h = lnk[1] - lnk[0]
    if n % 2 == 1:
        w = np.full(n, 2.0)
        w[1::2] = 4.0
        w[0] = w[-1] = 1.0
        return integrand @ w * (h / 3)
    return np.trapezoid(integrand, lnk, axis=1)


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
