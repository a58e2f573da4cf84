"""FFT wrappers and wavevectors (pmwd/pm_util.py:159-344) on the HIP FFT.

``fftfwd``/``fftinv`` run the hand-written 3D real FFT of libpmx for float32
3D fields on the GPU; other ranks/dtypes raise (pmwd's hot path only needs
3D float32).  Normalisation conventions are pmwd's: backward (unscaled
forward, 1/N inverse), 'ortho', 'forward', or a grid spacing ``norm`` that
multiplies by ``norm**d`` (forward) or ``norm**-d`` (inverse).
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib


def fftfreq(shape, spacing, dtype=torch.float64, sparse=True, device=None):
    """Angular wavevectors ``2 pi fftfreq(n)/spacing`` (pm_util.py:159-199).

    Computed in float64 as ``k / n`` then cast, with the last axis from
    rfftfreq; returns broadcastable sparse grids (or dense when
    ``sparse=False``).
    """
    from .particles import default_device
    device = device or default_device()
    period = 1.0 if spacing is None else 2 * math.pi / spacing
    kvec = []
    for axis, s in enumerate(shape):
        last = axis == len(shape) - 1
        n = s // 2 + 1 if last else s
        j = np.arange(n, dtype=np.float64)
        if not last:
            j[(s + 1) // 2:] -= s
        k = torch.as_tensor((j / s) * period, dtype=torch.float64).to(dtype).to(device)
        view = [1] * len(shape)
        view[axis] = n
        kvec.append(k.reshape(view))
    if not sparse:
        kvec = list(torch.broadcast_tensors(*kvec))
    return kvec


def _ctx_for(shape, device):
    # the FFT only needs the shape: any context of this shape serves
    return _lib.context_any_cell(shape, device)


def _check_field(f, complex_):
    if not isinstance(f, torch.Tensor):
        raise TypeError('expected a torch tensor')
    _lib.require_device(f, 'field')
    if f.dim() != 3:
        raise NotImplementedError('the HIP FFT implements 3D fields only')
    if complex_ and f.dtype != torch.complex64:
        raise NotImplementedError('the HIP FFT implements complex64 (float32) fields only')
    if not complex_ and f.dtype != torch.float32:
        raise NotImplementedError('the HIP FFT implements float32 fields only')


def _rfftn_raw(f, ctx=None):
    _check_field(f, False)
    f = f.contiguous()
    n0, n1, n2 = f.shape
    out = torch.empty((n0, n1, n2 // 2 + 1), dtype=torch.complex64, device=f.device)
    ctx = ctx or _ctx_for(f.shape, f.device)
    _lib.check(_lib.lib().pmx_rfftn(ctx.bind(), _lib.ptr(f), _lib.ptr(out)))
    return out


def _irfftn_raw(f, shape, ctx=None):
    _check_field(f, True)
    f = f.contiguous()
    n0, n1, n2 = shape
    if tuple(f.shape) != (n0, n1, n2 // 2 + 1):
        raise ValueError(f'complex field shape {tuple(f.shape)} does not match s={shape}')
    out = torch.empty((n0, n1, n2), dtype=torch.float32, device=f.device)
    ctx = ctx or _ctx_for(shape, f.device)
    _lib.check(_lib.lib().pmx_irfftn(ctx.bind(), _lib.ptr(f), _lib.ptr(out)))
    return out


def _interior(n2):
    """Last-axis bins of the r2c layout that stand for two full-spectrum bins
    (all but k=0 and, for even n2, k=n2/2)."""
    return slice(1, (n2 + 1) // 2)


def r2c_adjoint(g, shape, irfftn):
    """Adjoint of the unnormalised real-to-complex transform: N irfftn(g) with
    the interior last-axis bins halved.  With irfftn dropping the imaginary
    parts of the self-conjugate bins, this is Re(A^H g) for real inputs."""
    g = g.clone()
    g[..., _interior(shape[-1])] *= 0.5
    return irfftn(g.contiguous(), shape) * math.prod(shape)


def c2r_adjoint(y, shape, rfftn):
    """Adjoint of the 1/N-scaled complex-to-real transform: rfftn(y)/N with the
    interior last-axis bins doubled (the conjugate-Wirtinger gradient torch
    expects for complex inputs)."""
    g = rfftn(y.contiguous()) * (1 / math.prod(shape))
    g[..., _interior(shape[-1])] *= 2
    return g


# The backward of each transform runs on the context its forward used:
# autograd may call it on another thread, where the caller's context
# provider (serving's per-request pool) is not in scope.

class _RFFTn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f):
        ctx.shape = tuple(f.shape)
        ctx.pmx = _ctx_for(f.shape, f.device)
        return _rfftn_raw(f, ctx.pmx)

    @staticmethod
    def backward(ctx, g):
        return r2c_adjoint(g, ctx.shape, lambda x, s: _irfftn_raw(x, s, ctx.pmx))


class _IRFFTn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f, shape):
        ctx.shape = shape
        ctx.pmx = _ctx_for(shape, f.device)
        return _irfftn_raw(f, shape, ctx.pmx)

    @staticmethod
    def backward(ctx, y):
        return c2r_adjoint(y, ctx.shape, lambda x: _rfftn_raw(x, ctx.pmx)), None


def fftfwd(f, shape=None, axes=None, norm=None):
    """rfftn of a real field (pm_util.py:236-289)."""
    if torch.is_complex(f):
        raise ValueError('input field must be real')
    if axes is not None and tuple(axes) not in ((0, 1, 2), (-3, -2, -1)):
        raise NotImplementedError('only the full 3D transform is implemented')
    if shape is not None and tuple(shape) != tuple(f.shape):
        raise NotImplementedError('zero-padding/cropping (s != shape) is not implemented')
    out = _RFFTn.apply(f)
    N = math.prod(f.shape)
    if norm in (None, 'backward'):
        return out
    if norm == 'ortho':
        return out * (1 / math.sqrt(N))
    if norm == 'forward':
        return out * (1 / N)
    return out * float(norm)**3


def fftinv(f, shape=None, axes=None, norm=None):
    """irfftn of a Hermitian complex field (pm_util.py:292-344)."""
    if not torch.is_complex(f):
        raise ValueError('input field must be Hermitian complex')
    if axes is not None and tuple(axes) not in ((0, 1, 2), (-3, -2, -1)):
        raise NotImplementedError('only the full 3D transform is implemented')
    if shape is None:
        shape = (f.shape[0], f.shape[1], 2 * (f.shape[2] - 1))
    shape = tuple(int(s) for s in shape)
    out = _IRFFTn.apply(f, shape)
    N = math.prod(shape)
    if norm in (None, 'backward'):
        return out
    if norm == 'ortho':
        return out * math.sqrt(N)
    if norm == 'forward':
        return out * N
    return out * float(norm)**-3


def fft(f, shape=None, axes=None, norm=None):
    """fftfwd for real input, fftinv for complex input (pm_util.py:202-233)."""
    if torch.is_complex(f):
        return fftinv(f, shape, axes, norm)
    return fftfwd(f, shape, axes, norm)


def cic_geom(conf, spatial_shape, offset=0, cell_size=None):
    """pmx_cic_geom for enmesh(pmid, disp, conf.cell_size, conf.mesh_shape,
    offset, cell_size, spatial_shape) (pm_util.py:33-156)."""
    dim = conf.dim
    if dim > 3:
        raise NotImplementedError('dim > 3')
    g = _lib.CicGeom()
    off = np.broadcast_to(np.asarray(offset, dtype=np.float64), (dim,))
    for d in range(3):
        g.mesh_shape[d] = conf.mesh_shape[d] if d < dim else 0
        g.out_shape[d] = int(spatial_shape[d]) if d < dim else 1
        g.offset[d] = float(off[d]) if d < dim else 0.0
    g.cell_size = float(conf.cell_size)
    g.out_cell_size = -1.0 if cell_size is None else float(cell_size)
    return g


def cell_keys(ptcl, conf):
    """uint32 C-order key of each particle's lower CIC corner (as int64 tensor).

    The index of ``enmesh(...)[0][:, 0]`` (pm_util.py:119-137) raveled with
    ``Particles.raveled_id`` strides (particles.py:156-182).
    """
    n = len(ptcl)
    key = torch.empty(n, dtype=torch.int32, device=ptcl.disp.device)
    c = _lib.context(conf.mesh_shape, conf.cell_size, ptcl.disp.device)
    _lib.check(_lib.lib().pmx_cell_keys(c.bind(), _lib.ptr(ptcl.pmid.contiguous()),
                                        _lib.ptr(ptcl.disp.contiguous()), n, _lib.ptr(key)))
    return key.to(torch.int64) & 0xFFFFFFFF


def tile_keys(ptcl, conf):
    """Scatter-tile index of each particle's lower corner, and the tile shape."""
    n = len(ptcl)
    key = torch.empty(n, dtype=torch.int32, device=ptcl.disp.device)
    c = _lib.context(conf.mesh_shape, conf.cell_size, ptcl.disp.device)
    _lib.check(_lib.lib().pmx_tile_keys(c.bind(), _lib.ptr(ptcl.pmid.contiguous()),
                                        _lib.ptr(ptcl.disp.contiguous()), n, _lib.ptr(key)))
    ts = (ctypes.c_int * 3)()
    _lib.check(_lib.lib().pmx_tile_shape(c.h, ctypes.byref(ts)))
    return key.to(torch.int64) & 0xFFFFFFFF, tuple(ts)


def sort_perm(keys, device=None):
    """Stable ascending argsort of non-negative keys < 2^31 (HIP radix sort)."""
    keys = torch.as_tensor(keys)
    if device is not None:
        keys = keys.to(device)
    _lib.require_device(keys, 'keys')
    n = keys.numel()
    k32 = keys.to(torch.int32).contiguous()
    perm = torch.empty(n, dtype=torch.int32, device=keys.device)
    c = _lib.context((1, 1, 1), 1.0, keys.device)
    _lib.check(_lib.lib().pmx_sort_perm(c.bind(), _lib.ptr(k32), n, _lib.ptr(perm)))
    return perm.to(torch.int64)


def is_default_cic(conf, spatial_shape, offset, cell_size):
    """True when enmesh takes the hot default path (3D, no offset/cell change)."""
    return (conf.dim == 3 and cell_size is None and tuple(spatial_shape) == tuple(conf.mesh_shape)
            and np.all(np.asarray(offset) == 0))
