"""CIC readout (pmwd/gather.py:8-144) on the HIP kernels.

``val + sum_j mesh[c_j] w_j`` with out-of-bounds neighbours reading 0.  The
VJP follows ``_gather_bwd``: a disp cotangent from the mesh values and the
fraction gradients, a mesh cotangent scattered from the value cotangent,
and the value cotangent passed through.
"""
import numpy as np
import torch

from . import _lib
from .pm_util import cic_geom


class _GatherFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, disp, mesh, val, pmid, conf, offset, cell_size, val_scalar):
        dim = conf.dim
        n = pmid.shape[0]
        sshape = tuple(mesh.shape[:dim])
        chan = tuple(mesh.shape[dim:])
        nchan = int(np.prod(chan)) if chan else 1
        disp_c = disp.detach().contiguous()
        mesh_c = mesh.detach().contiguous()
        val_c = val.detach().contiguous() if val is not None else None
        out = torch.empty((n,) + chan, dtype=conf.float_dtype, device=disp.device)
        c = _lib.context(conf.mesh_shape, conf.cell_size, disp.device)
        g = cic_geom(conf, sshape, offset, cell_size)
        _lib.check(_lib.lib().pmx_gather_cic(c.bind(), g, _lib.ptr(pmid), _lib.ptr(disp_c), n,
                                             _lib.ptr(mesh_c), nchan, _lib.ptr(val_c),
                                             float(val_scalar), _lib.ptr(out)))
        ctx.save_for_backward(disp_c, mesh_c)
        ctx.pmid, ctx.conf, ctx.offset, ctx.cell_size = pmid, conf, offset, cell_size
        ctx.nchan, ctx.sshape, ctx.has_val = nchan, sshape, val is not None
        # the backward uses this context too: autograd may run it on another
        # thread, where context() would not see the caller's context provider
        ctx.pmx = c
        return out

    @staticmethod
    def backward(ctx, val_cot):
        disp, mesh = ctx.saved_tensors
        conf = ctx.conf
        n = ctx.pmid.shape[0]
        val_cot = val_cot.contiguous()
        disp_cot = torch.empty_like(disp) if ctx.needs_input_grad[0] else None
        mesh_cot = torch.zeros_like(mesh) if ctx.needs_input_grad[1] else None
        if disp_cot is not None or mesh_cot is not None:
            c = ctx.pmx
            g = cic_geom(conf, ctx.sshape, ctx.offset, ctx.cell_size)
            _lib.check(_lib.lib().pmx_gather_cic_adj(
                c.bind(), g, _lib.ptr(ctx.pmid), _lib.ptr(disp), n, _lib.ptr(mesh),
                _lib.ptr(val_cot), ctx.nchan, _lib.ptr(disp_cot), _lib.ptr(mesh_cot)))
        vgrad = val_cot if (ctx.has_val and ctx.needs_input_grad[2]) else None
        return disp_cot, mesh_cot, vgrad, None, None, None, None, None


def gather(ptcl, conf, mesh, val=0, offset=0, cell_size=None):
    """Gather particle values from the mesh multilinearly (gather.py:8-30)."""
    pmid, disp = ptcl.pmid, ptcl.disp
    _lib.require_device(disp, 'ptcl.disp')
    dev = disp.device
    mesh = torch.as_tensor(mesh).to(device=dev, dtype=conf.float_dtype)
    chan = tuple(mesh.shape[conf.dim:])
    val_t = None
    val_scalar = 0.0
    if isinstance(val, torch.Tensor) and val.dim() > 0:
        val_t = val.to(device=dev, dtype=conf.float_dtype)
        if tuple(val_t.shape[1:]) != chan:
            raise ValueError(f'channel shape mismatch: {chan} != {tuple(val_t.shape[1:])}')
    else:
        val_scalar = float(val)
        if chan != ():
            raise ValueError(f'channel shape mismatch: {chan} != ()')
    return _GatherFn.apply(disp, mesh, val_t, pmid.contiguous(), conf, offset, cell_size,
                           val_scalar)
