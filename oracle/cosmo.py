"""FP64 cosmology restated from pmwd (oracle side, test infrastructure).

Follows pmwd/cosmology.py (E2 185-219, H_deriv 222-242, Omega_m_a 245-269,
from_sigma8 95-105), pmwd/boltzmann.py (transfer_fit 32-123, growth_integ
163-229, growth 233-269, varlin 294-335, linear_power 399-455) and the vendored
Dormand-Prince integrator pmwd/ode_util.py (60-186).

The top-hat variance replaces the third-party ``mcfit.TophatVar`` FFTLog
(absent here) by a Simpson quadrature in ln k on the same wavenumber grid;
sigma8 and every ``from_sigma8`` normalisation is therefore *parity unpinned*
against mcfit.
"""
from dataclasses import dataclass, replace
import math

import numpy as np

from .conf import OConf


@dataclass(frozen=True)
class OCosmo:
    """Cosmology parameters (cosmology.py:15-161) plus the tables."""
    conf: OConf
    A_s_1e9: float
    n_s: float
    Omega_m: float
    Omega_b: float
    h: float
    Omega_k_: float = None
    w_0_: float = None
    w_a_: float = None
    transfer: np.ndarray = None
    growth: np.ndarray = None
    varlin: np.ndarray = None

    def replace(self, **kw):
        return replace(self, **kw)

    @property
    def A_s(self):
        return self.A_s_1e9 * 1e-9

    @property
    def k_pivot(self):
        c = self.conf
        return c.k_pivot_Mpc / (self.h * c.Mpc_SI) * c.L

    @property
    def Omega_c(self):
        return self.Omega_m - self.Omega_b

    @property
    def Omega_k(self):
        return 0.0 if self.Omega_k_ is None else self.Omega_k_

    @property
    def Omega_de(self):
        return 1 - (self.Omega_m + self.Omega_k)

    @property
    def w_0(self):
        return -1.0 if self.w_0_ is None else self.w_0_

    @property
    def w_a(self):
        return 0.0 if self.w_a_ is None else self.w_a_

    @property
    def sigma8(self):
        R = 8 * self.conf.Mpc_SI / self.conf.L
        return float(np.sqrt(varlin(R, 1.0, self, self.conf)))


def SimpleLCDM(conf, **kw):
    p = dict(A_s_1e9=2.0, n_s=0.96, Omega_m=0.3, Omega_b=0.05, h=0.7)
    p.update(kw)
    return OCosmo(conf, **p)


def E2(a, cosmo):
    """cosmology.py:185-219."""
    a = np.asarray(a, dtype=np.float64)
    de_a = a**(-3 * (1 + cosmo.w_0 + cosmo.w_a)) * np.exp(-3 * cosmo.w_a * (1 - a))
    return cosmo.Omega_m * a**-3 + cosmo.Omega_k * a**-2 + cosmo.Omega_de * de_a


def _dE2_da(a, cosmo):
    a = np.asarray(a, dtype=np.float64)
    p = -3 * (1 + cosmo.w_0 + cosmo.w_a)
    de_a = a**p * np.exp(-3 * cosmo.w_a * (1 - a))
    dde = (p / a + 3 * cosmo.w_a) * de_a
    return -3 * cosmo.Omega_m * a**-4 - 2 * cosmo.Omega_k * a**-3 + cosmo.Omega_de * dde


def H_deriv(a, cosmo):
    """dlnH/dlna = a E2'/(2 E2) (cosmology.py:222-242, by autodiff there)."""
    a = np.asarray(a, dtype=np.float64)
    return 0.5 * a * _dE2_da(a, cosmo) / E2(a, cosmo)


def Omega_m_a(a, cosmo):
    """cosmology.py:245-269."""
    a = np.asarray(a, dtype=np.float64)
    return cosmo.Omega_m / (a**3 * E2(a, cosmo))


def eh98_scales(cosmo, conf):
    """Scales of the Eisenstein & Hu 1998 fit (boltzmann.py:40-66): matter-
    radiation equality z_eq and k_eq [1/Mpc], drag epoch z_d, baryon-photon
    ratios R_d and R_eq, sound horizon s [Mpc] (EH98 eq. 6) and Silk scale
    k_silk [1/Mpc]."""
    T2_cmb_norm = (conf.T_cmb / 2.7)**2
    h2 = cosmo.h**2
    w_m = cosmo.Omega_m * h2
    w_b = cosmo.Omega_b * h2
    z_eq = 2.50e4 * w_m / T2_cmb_norm**2
    k_eq = 7.46e-2 * w_m / T2_cmb_norm
    b1 = 0.313 * w_m**-0.419 * (1 + 0.607 * w_m**0.674)
    b2 = 0.238 * w_m**0.223
    z_d = 1291 * w_m**0.251 / (1 + 0.659 * w_m**0.828) * (1 + b1 * w_b**b2)
    R_d = 31.5 * w_b / T2_cmb_norm**2 * (1e3 / z_d)
    R_eq = 31.5 * w_b / T2_cmb_norm**2 * (1e3 / z_eq)
    s = (2 / (3 * k_eq) * np.sqrt(6 / R_eq)
         * np.log((np.sqrt(1 + R_d) + np.sqrt(R_eq + R_d)) / (1 + np.sqrt(R_eq))))
    k_silk = 1.6 * w_b**0.52 * w_m**0.73 * (1 + (10.4 * w_m)**-0.95)
    return dict(w_m=w_m, w_b=w_b, z_eq=z_eq, k_eq=k_eq, z_d=z_d, R_d=R_d, R_eq=R_eq, s=s,
                k_silk=k_silk)


def transfer_fit(k, cosmo, conf):
    """Eisenstein & Hu 1998 fit, restated from boltzmann.py:32-123."""
    k = np.asarray(k, dtype=np.float64)
    k = k * cosmo.h / conf.L * conf.Mpc_SI
    sc = eh98_scales(cosmo, conf)
    w_m, w_b, z_eq, k_eq, z_d = sc['w_m'], sc['w_b'], sc['z_eq'], sc['k_eq'], sc['z_d']
    R_d, s, k_silk = sc['R_d'], sc['s'], sc['k_silk']
    f_b = cosmo.Omega_b / cosmo.Omega_m
    f_c = cosmo.Omega_c / cosmo.Omega_m

    if conf.transfer_fit_nowiggle:
        alpha_gamma = (1 - 0.328 * np.log(431 * w_m) * f_b
                       + 0.38 * np.log(22.3 * w_m) * f_b**2)
        gamma_eff_ratio = alpha_gamma + (1 - alpha_gamma) / (1 + (0.43 * k * s)**4)
        q_eff = k / (13.41 * k_eq * gamma_eff_ratio)
        L0 = np.log(2 * np.e + 1.8 * q_eff)
        C0 = 14.2 + 731 / (1 + 62.5 * q_eff)
        return L0 / (L0 + C0 * q_eff**2)

    a1 = (46.9 * w_m)**0.670 * (1 + (32.1 * w_m)**-0.532)
    a2 = (12.0 * w_m)**0.424 * (1 + (45.0 * w_m)**-0.582)
    alpha_c = a1**-f_b * a2**-f_b**3
    b1 = 0.944 / (1 + (458 * w_m)**-0.708)
    b2 = (0.395 * w_m)**-0.0266
    beta_c = 1 / (1 + b1 * (f_c**b2 - 1))

    def T0_tilde(k, alpha_c, beta_c):
        q = k / (13.41 * k_eq)
        L = np.log(np.e + 1.8 * beta_c * q)
        C = 14.2 / alpha_c + 386 / (1 + 69.9 * q**1.08)
        return L / (L + C * q**2)

    f = 1 / (1 + (k * s / 5.4)**4)
    T_c = f * T0_tilde(k, 1, beta_c) + (1 - f) * T0_tilde(k, alpha_c, beta_c)
    y = (1 + z_eq) / (1 + z_d)
    x = np.sqrt(1 + y)
    G = y * (-6 * x + (2 + 3 * y) * np.log((x + 1) / (x - 1)))
    alpha_b = 2.07 * k_eq * s * (1 + R_d)**-0.75 * G
    beta_node = 8.41 * w_m**0.435
    beta_b = 0.5 + f_b + (3 - 2 * f_b) * np.sqrt(1 + (17.2 * w_m)**2)
    T_b = ((T0_tilde(k, 1, 1) / (1 + (k * s / 5.2)**2)
            + alpha_b * (k * s)**3 / (beta_b**3 + (k * s)**3) * np.exp(-(k / k_silk)**1.4))
           * np.sinc((k * s)**2 / (np.pi * np.cbrt(beta_node**3 + (k * s)**3))))
    return f_c * T_c + f_b * T_b


def transfer(k, cosmo, conf):
    """boltzmann.py:126-160: linear interpolation of the table."""
    if cosmo.transfer is None:
        raise ValueError('Transfer table is empty.')
    return np.interp(np.asarray(k, dtype=np.float64), conf.transfer_k, cosmo.transfer)


# ---- Dormand-Prince, restated from ode_util.py:60-186 (forward only) ----

_ALPHA = np.array([1 / 5, 3 / 10, 4 / 5, 8 / 9, 1., 1., 0])
_BETA = np.array([
    [1 / 5, 0, 0, 0, 0, 0, 0],
    [3 / 40, 9 / 40, 0, 0, 0, 0, 0],
    [44 / 45, -56 / 15, 32 / 9, 0, 0, 0, 0],
    [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729, 0, 0, 0],
    [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656, 0, 0],
    [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0]])
_C_SOL = np.array([35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0])
_C_ERR = np.array([35 / 384 - 1951 / 21600, 0, 500 / 1113 - 22642 / 50085,
                   125 / 192 - 451 / 720, -2187 / 6784 - -12231 / 42400,
                   11 / 84 - 649 / 6300, -1. / 60.])
_C_MID = np.array([6025192743 / 30085553152 / 2, 0, 51252292925 / 65400821598 / 2,
                   -2691868925 / 45128329728 / 2, 187940372067 / 1594534317056 / 2,
                   -1776094331 / 19743644256 / 2, 11237099 / 235043384 / 2])


def _rk_step(func, y0, f0, t0, dt):
    k = np.zeros((7, y0.shape[0]))
    k[0] = f0
    for i in range(1, 7):
        ti = t0 + dt * _ALPHA[i - 1]
        yi = y0 + dt * (_BETA[i - 1] @ k)
        k[i] = func(yi, ti)
    y1 = dt * (_C_SOL @ k) + y0
    err = dt * (_C_ERR @ k)
    return y1, k[-1], err, k


def _interp_fit(y0, y1, k, dt):
    y_mid = y0 + dt * (_C_MID @ k)
    dy0, dy1 = k[0], k[-1]
    a = -2. * dt * dy0 + 2. * dt * dy1 - 8. * y0 - 8. * y1 + 16. * y_mid
    b = 5. * dt * dy0 - 3. * dt * dy1 + 18. * y0 + 14. * y1 - 32. * y_mid
    c = -4. * dt * dy0 + dt * dy1 - 11. * y0 - 5. * y1 + 16. * y_mid
    return np.stack([a, b, c, dt * dy0, y0])


def _initial_step(fun, t0, y0, order, rtol, atol, f0):
    scale = atol + np.abs(y0) * rtol
    d0 = np.linalg.norm(y0 / scale)
    d1 = np.linalg.norm(f0 / scale)
    h0 = 1e-6 if (d0 < 1e-5 or d1 < 1e-5) else 0.01 * d0 / d1
    y1 = y0 + h0 * f0
    f1 = fun(y1, t0 + h0)
    d2 = np.linalg.norm((f1 - f0) / scale) / h0
    if d1 <= 1e-15 and d2 <= 1e-15:
        h1 = max(1e-6, h0 * 1e-3)
    else:
        h1 = (0.01 / max(d1, d2)) ** (1. / (order + 1.))
    return min(100. * h0, h1)


def odeint(func, y0, t, rtol, atol, mxstep=np.inf, hmax=np.inf, dt0=None):
    """Adaptive Dopri5 as in ode_util.py:188-232 (values at ``t``)."""
    y = np.asarray(y0, dtype=np.float64)
    f = func(y, t[0])
    dt = dt0[0] if isinstance(dt0, tuple) else dt0
    if dt is None:
        dt = _initial_step(func, t[0], y, 4, rtol, atol, f)
    dt = min(max(dt, 0.), hmax)
    tc = t[0]
    last_t = t[0]
    coeff = np.stack([y] * 5)
    ys = [y.copy()]
    for target in t[1:]:
        i = 0
        while tc < target and i < mxstep and dt > 0:
            ny, nf, nerr, k = _rk_step(func, y, f, tc, dt)
            nt = tc + dt
            tol = atol + rtol * np.maximum(np.abs(y), np.abs(ny))
            ratio = np.sqrt(np.mean((nerr / tol)**2))
            ncoeff = _interp_fit(y, ny, k, dt)
            dfac = 1.0 if ratio < 1 else 0.2
            fac = min(10.0, max(ratio**(-1.0 / 5.0) * 0.9, dfac)) if ratio != 0 else 10.0
            ndt = min(max(dt * fac, 0.), hmax)
            if ratio <= 1.:
                y, f, last_t, tc, coeff = ny, nf, tc, nt, ncoeff
            dt = ndt
            i += 1
        rel = (target - last_t) / (tc - last_t)
        yt = coeff[0]
        for c in coeff[1:]:
            yt = yt * rel + c
        ys.append(yt)
    return np.stack(ys)


def growth_ode(G, lna, cosmo):
    """Right-hand side of the 2LPT growth ODE (boltzmann.py:197-205)."""
    a = np.exp(lna)
    dlnH = H_deriv(a, cosmo)
    Omf = 1.5 * Omega_m_a(a, cosmo)
    G1, G1p, G2, G2p = np.split(np.asarray(G), 4, axis=-1)
    G1pp = -(3 + dlnH - Omf) * G1 - (4 + dlnH) * G1p
    G2pp = Omf * G1**2 - (8 + 2 * dlnH - Omf) * G2 - (6 + dlnH) * G2p
    return np.concatenate((G1p, G1pp, G2p, G2pp), axis=-1)


def growth_integ(cosmo, conf):
    """boltzmann.py:163-229: table growth[order][deriv][a] of D_m/a^m."""
    eps = np.finfo(np.float64).eps
    a_ic = 0.5 * np.cbrt(eps)
    if a_ic >= conf.a_lpt_step:
        a_ic = 0.1 * conf.a_lpt_step
    a = conf.growth_a.copy()
    a[0] = a_ic
    lna = np.log(a)
    G_ic = np.array((1, 0, 3 / 7, 0), dtype=np.float64)
    G = odeint(lambda y, t: growth_ode(y, t, cosmo), G_ic, lna,
               rtol=conf.growth_rtol, atol=conf.growth_atol, dt0=conf.growth_inistep)
    Gd = growth_ode(G, lna[:, None], cosmo)
    n = len(a)
    G = G.reshape(n, 2, 2)
    Gd = Gd.reshape(n, 2, 2)
    G = np.concatenate((G, Gd[..., -1:]), axis=2)   # (n, order, [G, G', G''])
    G = np.moveaxis(G, 0, 2)                        # (order, 3, n)
    m = np.array((1., 2.))[:, None]
    table = np.stack((G[:, 0], m * G[:, 0] + G[:, 1],
                      m**2 * G[:, 0] + 2 * m * G[:, 1] + G[:, 2]), axis=1)
    return cosmo.replace(growth=table)


def growth(a, cosmo, conf, order=1, deriv=0):
    """boltzmann.py:233-269."""
    if cosmo.growth is None:
        raise ValueError('Growth table is empty.')
    a = np.asarray(a, dtype=np.float64)
    return a**order * np.interp(a, conf.growth_a, cosmo.growth[order - 1][deriv])


def linear_power(k, a, cosmo, conf):
    """boltzmann.py:399-455 (FP64)."""
    k = np.asarray(k, dtype=np.float64)
    T = transfer(k, cosmo, conf)
    with np.errstate(divide='ignore'):
        P = (0.32 * cosmo.A_s * cosmo.k_pivot * (k / cosmo.k_pivot)**cosmo.n_s
             * (np.pi * (conf.c / conf.H_0)**2 / cosmo.Omega_m * T)**2)
    if a is not None:
        P = P * growth(a, cosmo, conf)**2
    return P


def tophat_var(R, cosmo, conf):
    """sigma^2(R) = int dlnk k^3 P(k) W(kR)^2 / (2 pi^2), Simpson in ln k.

    Stand-in for ``mcfit.TophatVar`` (boltzmann.py:272-291); parity unpinned.
    """
    k = conf.transfer_k[1:]
    P = linear_power(k, None, cosmo, conf)
    lnk = np.log(k)
    R = np.atleast_1d(np.asarray(R, dtype=np.float64))
    x = k[None, :] * R[:, None]
    with np.errstate(invalid='ignore', divide='ignore'):
        W = np.where(x > 1e-3, 3 * (np.sin(x) - x * np.cos(x)) / x**3, 1 - x**2 / 10)
    f = k**3 * P / (2 * np.pi**2) * W**2
    h = lnk[1] - lnk[0]
    n = f.shape[1]
    if (n - 1) % 2 == 0:
        w = np.ones(n); w[1:-1:2] = 4; w[2:-1:2] = 2
        return (f * w).sum(axis=1) * h / 3
    return np.trapezoid(f, lnk, axis=1)


def varlin_integ(cosmo, conf):
    return cosmo.replace(varlin=tophat_var(conf.varlin_R, cosmo, conf))


def varlin(R, a, cosmo, conf):
    """boltzmann.py:294-335."""
    if cosmo.varlin is None:
        raise ValueError('varlin table is empty.')
    s2 = np.interp(np.asarray(R, dtype=np.float64), conf.varlin_R, cosmo.varlin)
    if a is not None:
        s2 = s2 * growth(a, cosmo, conf)**2
    return s2


def boltzmann(cosmo, conf):
    """boltzmann.py:338-374."""
    cosmo = cosmo.replace(transfer=transfer_fit(conf.transfer_k, cosmo, conf))
    cosmo = growth_integ(cosmo, conf)
    cosmo = varlin_integ(cosmo, conf)
    return cosmo


def from_sigma8(conf, sigma8, **kw):
    """cosmology.py:95-105."""
    c1 = boltzmann(OCosmo(conf, 1.0, **kw), conf)
    return OCosmo(conf, (sigma8 / c1.sigma8)**2, **kw)
