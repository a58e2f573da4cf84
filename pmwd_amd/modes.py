"""White noise and linear modes (pmwd/modes.py:12-117).

``white_noise`` draws float32 normals on the device (``pmx_white_noise``:
Philox4x32-10 keyed by the seed and a float64 Box-Muller transform; pmwd's
jax.random threefry stream is not reproducible without jax); Fourier modes
use the HIP r2c with 'ortho' norm.
``linear_modes`` multiplies by sqrt(V Plin(|k|)) with |k| formed in float32
from the float32 wavevectors and Plin evaluated in float64, as pmwd.
"""
import ctypes
import math

import numpy as np
import torch

from .boltzmann import growth
from .cosmo_param import check_theta, host_value, modes_factor_derivs
from .particles import default_device
from .pm_util import fftfreq, fftfwd, fftinv


def white_noise(seed, conf, real=False, unit_abs=False, device=None):
    """White-noise real or Fourier modes, zero mean, unit variance
    (modes.py:12-49); the real noise is drawn on the device."""
    from . import _lib
    device = torch.device(device) if device is not None else default_device()
    _lib.require_device(torch.empty(0, device=device), 'white_noise device')
    modes = torch.empty(conf.ptcl_grid_shape, dtype=torch.float32, device=device)
    c = _lib.context_any_cell(conf.ptcl_grid_shape, device)
    _lib.check(_lib.lib().pmx_white_noise(c.bind(), int(seed) % 2**64, modes.numel(),
                                          _lib.ptr(modes)))
    if real and not unit_abs:
        return modes
    modes = fftfwd(modes, norm='ortho')
    if unit_abs:
        modes = modes / torch.abs(modes)
    if real:
        modes = fftinv(modes, shape=conf.ptcl_grid_shape, norm='ortho')
    return modes


def _interp(x, xp, fp):
    """numpy.interp on device tensors (float64), clamped at the ends."""
    i = torch.searchsorted(xp, x, right=True).clamp(1, len(xp) - 1)
    x0, x1 = xp[i - 1], xp[i]
    f0, f1 = fp[i - 1], fp[i]
    t = (x - x0) / (x1 - x0)
    out = f0 + t * (f1 - f0)
    out = torch.where(x <= xp[0], fp[0], out)
    return torch.where(x >= xp[-1], fp[-1], out)


def linear_power_k(k, a, cosmo, conf):
    """linear_power (boltzmann.py:399-455) on a float64 device tensor of k."""
    dev = k.device
    xp = torch.tensor(conf.transfer_k, dtype=torch.float64, device=dev)
    fp = torch.as_tensor(cosmo.transfer, dtype=torch.float64, device=dev)
    T = _interp(k, xp, fp)
    P = (0.32 * cosmo.A_s * cosmo.k_pivot * (k / cosmo.k_pivot)**cosmo.n_s
         * (math.pi * (conf.c / conf.H_0)**2 / cosmo.Omega_m * T)**2)
    if a is not None:
        P = P * float(growth(a, cosmo, conf))**2
    return P


def _needs_graph(modes, theta):
    if not torch.is_grad_enabled():
        return False
    if modes.requires_grad:
        return True
    return theta is not None and any(t.requires_grad for t in theta.values())


def _linear_modes_fused(modes, cosmo, conf, a, real):
    """linear_modes without gradients: the r2c of real white noise, then one
    library pass (pmx_linear_modes) for the 'ortho' scale and sqrt(V P)."""
    from . import _lib
    dev = modes.device
    c = _lib.context(conf.ptcl_grid_shape, conf.ptcl_spacing, dev)
    if torch.is_complex(modes):
        out = modes.detach().to(torch.complex64).clone().contiguous()
        scale = 1.0
    else:
        out = fftfwd(modes.detach().to(conf.float_dtype))
        scale = 1 / math.sqrt(math.prod(conf.ptcl_grid_shape))
    tk = np.ascontiguousarray(conf.transfer_k, dtype=np.float64)
    tT = np.ascontiguousarray(cosmo.transfer, dtype=np.float64)
    par = np.array([0.32 * cosmo.A_s * cosmo.k_pivot, cosmo.k_pivot, cosmo.n_s,
                    math.pi * (conf.c / conf.H_0)**2 / cosmo.Omega_m, conf.box_vol,
                    float(growth(a, cosmo, conf))**2 if a is not None else 1.0],
                   dtype=np.float64)
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    _lib.check(_lib.lib().pmx_linear_modes(c.bind(), _lib.ptr(out), float(scale), P(tk), P(tT),
                                           len(tk), P(par)))
    if real:
        out = fftinv(out, shape=conf.ptcl_grid_shape, norm=conf.ptcl_spacing)
    return out


def linear_modes(modes, cosmo, conf, a=None, real=False, theta=None):
    """sqrt(V Plin(k)) * white-noise modes (modes.py:67-117).

    ``theta`` ({'Omega_m': t, 'sigma8': t} of 0-d float64 tensors) makes the
    result differentiable with respect to those parameters (A_s follows
    sigma8 as in Cosmology.from_sigma8); see ``cosmo_param``.
    """
    if cosmo.transfer is None:
        raise ValueError('Transfer table is empty. Call transfer_integ or boltzmann first.')
    theta = check_theta(theta, cosmo)
    if not _needs_graph(modes, theta):
        return _linear_modes_fused(modes, cosmo, conf, a, real)
    if not torch.is_complex(modes):
        modes = fftfwd(modes.to(conf.float_dtype), norm='ortho')
    kvec = fftfreq(conf.ptcl_grid_shape, conf.ptcl_spacing, dtype=conf.float_dtype,
                   device=modes.device)
    k2 = 0
    for k in kvec:
        k2 = k2 + k * k
    k = torch.sqrt(k2).to(torch.float64)
    P = linear_power_k(k, a, cosmo, conf)
    fac = torch.sqrt(P * conf.box_vol).to(conf.float_dtype)
    if theta is not None:
        d_om, d_s8 = modes_factor_derivs(k, a, cosmo, conf, linear_power_k)
        fac = host_value(fac, theta, d_om, d_s8)
    modes = modes * fac
    if real:
        modes = fftinv(modes, shape=conf.ptcl_grid_shape, norm=conf.ptcl_spacing)
    return modes
