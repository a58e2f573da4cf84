"""Lagrangian perturbation theory initial conditions (pmwd/lpt.py:13-212).

2LPT on the particle grid with the HIP real FFTs: first-order potential,
the second-order source from six strain fields, then ``-ik`` gradients for
the displacements and canonical velocities.  Without gradients the whole
of it is one library call; the differentiable path runs the k-space
multiplies and real-space combinations as device tensor ops around the FFT
kernels.
"""
import ctypes

import numpy as np
import torch

from .boltzmann import growth
from .cosmo_param import check_theta, host_value, lpt_growth_derivs
from .cosmology import E2
from .gravity import laplace, neg_grad, nyquist_mask
from .particles import Particles
from .pm_util import fftfreq, fftfwd, fftinv


def _strain(kvec, i, j, pot, conf):
    """Strain component (lpt.py:13-37); Nyquist zeroed only when i != j."""
    k_i, k_j = kvec[i], kvec[j]
    if i != j:
        k_i = torch.where(nyquist_mask(k_i, conf.ptcl_spacing), torch.zeros_like(k_i), k_i)
        k_j = torch.where(nyquist_mask(k_j, conf.ptcl_spacing), torch.zeros_like(k_j), k_j)
    s = (-k_i * k_j) * pot
    return fftinv(s, shape=conf.ptcl_grid_shape).to(conf.float_dtype)


def _L(kvec, pot, conf):
    """Second-order source (lpt.py:40-76) with pot_n = pot_m."""
    L = torch.zeros(conf.ptcl_grid_shape, dtype=conf.float_dtype, device=pot.device)
    dim = conf.dim
    diag = [_strain(kvec, i, i, pot, conf) for i in range(dim)]
    for i in range(dim):
        for j in range(dim - 1, i, -1):
            L = L + diag[i] * diag[j]
    for i in range(dim - 1):
        for j in range(i + 1, dim):
            s = _strain(kvec, i, j, pot, conf)
            L = L - s * s
    return L


def _growth_f32(a, cosmo, conf, order):
    """float32 D_n(a) and a^2 H(a) D_n'(a) (lpt.py:196-201)."""
    D = float(np.float32(growth(a, cosmo, conf, order=order)))
    dD = growth(a, cosmo, conf, order=order, deriv=1)
    return D, float(np.float32(a**2 * np.sqrt(E2(a, cosmo)) * dD))


def _needs_graph(modes, theta):
    if not torch.is_grad_enabled():
        return False
    if modes.requires_grad:
        return True
    return theta is not None and any(t.requires_grad for t in theta.values())


def _lpt_fused(modes, cosmo, conf):
    """The whole of lpt() in the library (pmx_lpt): first-order gradient,
    the six strains and their source, the second-order gradient, each added
    to the particles in one pass."""
    from . import _lib
    dev = modes.device
    # gen_grid's arrays are fresh: the library adds to them in place
    ptcl = Particles.gen_grid(conf, vel=True, device=dev)
    disp, vel = ptcl.disp, ptcl.vel
    a = conf.a_start
    fac = np.zeros(4, dtype=np.float32)
    for order in range(1, 1 + conf.lpt_order):
        fac[2 * order - 2:2 * order] = _growth_f32(a, cosmo, conf, order)
    m = modes.detach().to(torch.complex64).contiguous()
    c = _lib.context(conf.ptcl_grid_shape, conf.ptcl_spacing, dev)
    _lib.check(_lib.lib().pmx_lpt(c.bind(), _lib.ptr(m), float(conf.ptcl_cell_vol),
                                  int(conf.lpt_order), fac.ctypes.data_as(ctypes.c_void_p),
                                  _lib.ptr(disp), _lib.ptr(vel)))
    return ptcl.replace(disp=disp, vel=vel), None


def lpt(modes, cosmo, conf, theta=None):
    """LPT at ``conf.lpt_order`` (lpt.py:136-212); returns (ptcl, obsvbl).

    Without gradients this is one library call (``pmx_lpt``); with them the
    same steps run as device tensor ops around the library FFTs, which
    autograd differentiates.  ``theta`` ({'Omega_m': t}) makes the growth
    factors differentiable with respect to Omega_m; see ``cosmo_param``.
    """
    if conf.dim != 3:
        raise ValueError(f'dim={conf.dim} not supported')
    if conf.lpt_order not in (0, 1, 2, 3):
        raise ValueError(f'lpt_order={conf.lpt_order} not supported')
    if conf.lpt_order > 2:
        raise NotImplementedError('lpt_order=3 is not implemented (as in pmwd)')
    theta = check_theta(theta, cosmo)
    if not _needs_graph(modes, theta):
        return _lpt_fused(modes, cosmo, conf)
    dev = modes.device
    modes = modes / conf.ptcl_cell_vol
    kvec = fftfreq(conf.ptcl_grid_shape, conf.ptcl_spacing, dtype=conf.float_dtype, device=dev)
    pots = []
    if conf.lpt_order > 0:
        pots.append(laplace(kvec, modes))
    if conf.lpt_order > 1:
        pots.append(laplace(kvec, fftfwd(_L(kvec, pots[0], conf))))
    ptcl = Particles.gen_grid(conf, vel=True, device=dev)
    disp, vel = ptcl.disp.clone(), ptcl.vel.clone()
    a = conf.a_start
    for order in range(1, 1 + conf.lpt_order):
        D, a2HDp = _growth_f32(a, cosmo, conf, order)
        if theta is not None and 'Omega_m' in theta:
            _, D_om, _, v_om = lpt_growth_derivs(a, cosmo, conf, order)
            f32 = lambda x: torch.tensor(x, dtype=conf.float_dtype, device=dev)
            D = host_value(f32(D), theta, torch.tensor(D_om, dtype=torch.float64))
            a2HDp = host_value(f32(a2HDp), theta, torch.tensor(v_om, dtype=torch.float64))
        for i, k in enumerate(kvec):
            g = neg_grad(k, pots[order - 1], conf.ptcl_spacing)
            g = fftinv(g, shape=conf.ptcl_grid_shape).to(conf.float_dtype).reshape(-1)
            disp[:, i] = disp[:, i] + D * g
            vel[:, i] = vel[:, i] + a2HDp * g
    return ptcl.replace(disp=disp.contiguous(), vel=vel.contiguous()), None
