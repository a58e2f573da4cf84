"""Lagrangian perturbation theory initial conditions (pmwd/lpt.py:13-212).

2LPT on the particle grid with the HIP real FFTs: first-order potential,
the second-order source from six strain fields, then ``-ik`` gradients for
the displacements and canonical velocities.  k-space multiplies and the
real-space combinations are device tensor ops around the FFT kernels.
"""
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


def lpt(modes, cosmo, conf, theta=None):
    """LPT at ``conf.lpt_order`` (lpt.py:136-212); returns (ptcl, obsvbl).

    ``theta`` ({'Omega_m': t}) makes the growth factors differentiable with
    respect to Omega_m; see ``cosmo_param``.
    """
    if conf.dim != 3:
        raise ValueError(f'dim={conf.dim} not supported')
    if conf.lpt_order not in (0, 1, 2, 3):
        raise ValueError(f'lpt_order={conf.lpt_order} not supported')
    if conf.lpt_order > 2:
        raise NotImplementedError('lpt_order=3 is not implemented (as in pmwd)')
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
    theta = check_theta(theta, cosmo)
    for order in range(1, 1 + conf.lpt_order):
        D = float(np.float32(growth(a, cosmo, conf, order=order)))
        dD = growth(a, cosmo, conf, order=order, deriv=1)
        a2HDp = float(np.float32(a**2 * np.sqrt(E2(a, cosmo)) * dD))
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
