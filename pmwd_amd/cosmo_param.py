"""Differentiable (Omega_m, sigma8) through the forward model.

pmwd takes gradients with respect to the cosmology by differentiating the
whole model with jax: boltzmann tables (boltzmann.py:163-374), linear modes
(modes.py:67-117), LPT (lpt.py:136-212) and the N-body factors
(nbody.py:12-36), with A_s set from sigma8 (cosmology.py:95-105).  Here the
tables are float64 host arrays, so every host quantity that enters a device
tensor op is wrapped in one autograd node that carries its derivatives with
respect to the parameters:

* the linear-mode factor sqrt(V P(k)): exactly proportional to sigma8 (A_s
  goes as sigma8^2); its Omega_m derivative at fixed sigma8, with A_s
  renormalised, by a float64 central difference of the host tables (the
  transfer fit and the sigma8 integral are smooth closed forms);
* the LPT growth scalars D_n(a) and a^2 H D_n'(a): Omega_m derivatives from
  the growth ODE's forward sensitivities;
* the N-body step factors: handled by ``nbody(..., omega_m=...)``.

``theta`` arguments are dicts with optional keys ``'Omega_m'`` and
``'sigma8'`` holding 0-d float64 tensors (usually ``requires_grad=True``)
whose values equal the cosmology's.
"""
import math

import numpy as np
import torch

from .boltzmann import boltzmann, growth, growth_dOm
from .cosmology import E2, E2_dOm

# central-difference step in Omega_m for the host tables
OM_STEP = 1e-5


def check_theta(theta, cosmo):
    """Validates a parameter dict against the cosmology it stands for."""
    if theta is None:
        return None
    for key in theta:
        if key not in ('Omega_m', 'sigma8'):
            raise ValueError(f'unsupported differentiable parameter {key!r}')
    if 'Omega_m' in theta and abs(float(theta['Omega_m'].detach()) - cosmo.Omega_m) > 1e-12:
        raise ValueError('theta["Omega_m"] does not equal cosmo.Omega_m')
    if 'sigma8' in theta and abs(float(theta['sigma8'].detach()) - cosmo.sigma8) > 1e-9 * cosmo.sigma8:
        raise ValueError('theta["sigma8"] does not equal cosmo.sigma8')
    return theta


def at_fixed_sigma8(cosmo, conf, Omega_m):
    """The cosmology at another Omega_m with A_s renormalised so that sigma8
    is unchanged (the from_sigma8 parametrisation), with its transfer and
    growth tables (no variance table)."""
    s8 = cosmo.sigma8
    c = boltzmann(cosmo.replace(Omega_m=float(Omega_m), transfer=None, growth=None,
                                growth_dOm=None, varlin=None), conf, varlin=False)
    scale = (s8 / c.sigma8)**2
    return c.replace(A_s_1e9=c.A_s_1e9 * scale)


class _HostValue(torch.autograd.Function):
    """A host-computed value v(Omega_m, sigma8) entering the device graph:
    forward returns v, backward sum(g * dv/dtheta) for each parameter."""

    @staticmethod
    def forward(ctx, value, t_om, d_om, t_s8, d_s8):
        ctx.save_for_backward(d_om if d_om is not None else torch.empty(0),
                              d_s8 if d_s8 is not None else torch.empty(0))
        ctx.has = (t_om is not None, t_s8 is not None)
        ctx.meta = [(t.dtype, t.device) if t is not None else None for t in (t_om, t_s8)]
        return value.clone()

    @staticmethod
    def backward(ctx, g):
        d_om, d_s8 = ctx.saved_tensors
        out = []
        for has, d, meta in zip(ctx.has, (d_om, d_s8), ctx.meta):
            if not has:
                out.append(None)
                continue
            s = (g.to(torch.float64) * d.to(device=g.device, dtype=torch.float64)).sum()
            out.append(s.to(dtype=meta[0], device=meta[1]))
        return None, out[0], None, out[1], None


def host_value(value, theta, d_om=None, d_s8=None):
    """``value`` (a device tensor) tied to theta through the given derivative
    tensors (same shape, any float dtype); returns value itself when theta
    holds no parameter that requires grad."""
    if theta is None:
        return value
    t_om = theta.get('Omega_m')
    t_s8 = theta.get('sigma8')
    t_om = t_om if (t_om is not None and d_om is not None) else None
    t_s8 = t_s8 if (t_s8 is not None and d_s8 is not None) else None
    if t_om is None and t_s8 is None:
        return value
    return _HostValue.apply(value, t_om, d_om if t_om is not None else None,
                            t_s8, d_s8 if t_s8 is not None else None)


def modes_factor_derivs(k, a, cosmo, conf, power_fn):
    """d sqrt(V P(k, a)) / d Omega_m at fixed sigma8 and d / d sigma8 (= the
    factor over sigma8).  The a-independent part is a float64 central
    difference of the host tables; the growth D(a), when a is given, enters
    through its forward sensitivity (no difference over the ODE solution)."""
    V = conf.box_vol
    f0 = torch.sqrt(power_fn(k, None, cosmo, conf) * V)
    cp = at_fixed_sigma8(cosmo, conf, cosmo.Omega_m + OM_STEP)
    cm = at_fixed_sigma8(cosmo, conf, cosmo.Omega_m - OM_STEP)
    fp = torch.sqrt(power_fn(k, None, cp, conf) * V)
    fm = torch.sqrt(power_fn(k, None, cm, conf) * V)
    d0 = (fp - fm) / (2 * OM_STEP)
    if a is None:
        return d0, f0 / cosmo.sigma8
    D = float(growth(a, cosmo, conf))
    D_om = float(growth_dOm(a, cosmo, conf))
    return d0 * D + f0 * D_om, f0 * (D / cosmo.sigma8)


def lpt_growth_derivs(a, cosmo, conf, order):
    """D_n(a), a^2 sqrt(E2(a)) D_n'(a) and their Omega_m derivatives."""
    D = float(growth(a, cosmo, conf, order=order))
    dD = float(growth(a, cosmo, conf, order=order, deriv=1))
    D_om = float(growth_dOm(a, cosmo, conf, order=order))
    dD_om = float(growth_dOm(a, cosmo, conf, order=order, deriv=1))
    e2 = float(E2(a, cosmo))
    e = math.sqrt(e2)
    e_om = float(E2_dOm(a, cosmo)) / (2 * e)
    v = a**2 * e * dD
    v_om = a**2 * (e_om * dD + e * dD_om)
    return D, D_om, v, v_om
