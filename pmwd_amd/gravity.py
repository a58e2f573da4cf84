"""Poisson solve and particle accelerations (pmwd/gravity.py:9-72).

``gravity`` runs the fused HIP pipeline: binned CIC scatter, the z-r2c pass
with ``(rho-1)*1.5*Omega_m`` fused on load, the y pass, the x pass fused with
Green's function, ``-ik`` (Nyquist zeroed) and the inverse x FFT for all
three components, the inverse y pass, the z-c2r pass, and a 3-component CIC
gather.  ``laplace``/``neg_grad`` are the standalone k-space operators on
complex device tensors.
"""
import math

import torch

from . import _lib


def laplace(kvec, src, cosmo=None):
    """-src/k^2 with k=0 -> 0 (gravity.py:9-16; self-adjoint)."""
    k2 = 0
    for k in kvec:
        k2 = k2 + k * k
    nz = k2 != 0
    safe = torch.where(nz, k2, torch.ones_like(k2))
    re = torch.where(nz, (-src.real) / safe, torch.zeros_like(safe))
    im = torch.where(nz, (-src.imag) / safe, torch.zeros_like(safe))
    return torch.complex(re, im)


def nyquist_mask(k, spacing):
    nyq = math.pi / spacing
    eps = nyq * torch.finfo(k.dtype).eps
    return torch.abs(torch.abs(k) - nyq) <= eps


def neg_grad(k, pot, spacing):
    """(-i k) pot with the Nyquist planes zeroed (gravity.py:37-44)."""
    km = torch.where(nyquist_mask(k, spacing), torch.zeros_like(k), k)
    return torch.complex(km * pot.imag, -(km * pot.real))


def omega_coef(cosmo, conf):
    """1.5 * Omega_m in float32 as gravity.py:54 forms it."""
    om = torch.tensor(cosmo.Omega_m, dtype=torch.float32)
    return float(torch.tensor(1.5, dtype=torch.float32) * om)


class _GravityFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, disp, pmid, conf, val, coef):
        n = pmid.shape[0]
        disp_c = disp.detach().contiguous()
        acc = torch.empty_like(disp_c)
        c = _lib.context(conf.mesh_shape, conf.cell_size, disp.device)
        _lib.check(_lib.lib().pmx_gravity(c.bind(), _lib.ptr(pmid), _lib.ptr(disp_c), n, val, coef,
                                          _lib.ptr(acc)))
        ctx.save_for_backward(disp_c)
        ctx.pmid, ctx.conf, ctx.val, ctx.coef = pmid, conf, val, coef
        # the backward uses this context too: autograd may run it on another
        # thread, where context() would not see the caller's context provider
        ctx.pmx = c
        return acc

    @staticmethod
    def backward(ctx, acc_cot):
        (disp,) = ctx.saved_tensors
        conf = ctx.conf
        n = ctx.pmid.shape[0]
        acc = torch.empty_like(disp)
        disp_cot = torch.empty_like(disp)
        c = ctx.pmx
        _lib.check(_lib.lib().pmx_gravity_vjp(c.bind(), _lib.ptr(ctx.pmid), _lib.ptr(disp), n,
                                              ctx.val, ctx.coef, _lib.ptr(acc_cot.contiguous()),
                                              _lib.ptr(acc), _lib.ptr(disp_cot), None))
        return disp_cot, None, None, None, None


def gravity(a, ptcl, cosmo, conf):
    """Gravitational accelerations in [H_0^2 L] (gravity.py:47-72).

    ``a`` is unused, as in pmwd.  Differentiable w.r.t. ``ptcl.disp``.
    """
    if conf.dim != 3:
        raise NotImplementedError('gravity is implemented for 3D only')
    _lib.require_device(ptcl.disp, 'ptcl.disp')
    val = float(torch.tensor(conf.mesh_size / conf.ptcl_num, dtype=torch.float32))
    return _GravityFn.apply(ptcl.disp, ptcl.pmid.contiguous(), conf, val, omega_coef(cosmo, conf))
