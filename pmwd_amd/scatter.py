"""CIC mass assignment (pmwd/scatter.py:8-150) on the HIP kernels.

Default arguments run the tile-binned, LDS-privatised scatter
(``pmx_scatter_mass``); other arguments (offset, output cell size, channels,
per-particle values, dims 1-2) run the generic kernel.  The VJP follows
``_scatter_bwd``: disp and val cotangents from a gather of the mesh
cotangent, which passes through unchanged to the input mesh.
"""
import numpy as np
import torch

from . import _lib
from .pm_util import cic_geom, is_default_cic


def _val_default(conf):
    return conf.mesh_size / conf.ptcl_num


class _ScatterFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, disp, mesh, val, pmid, conf, offset, cell_size, val_scalar):
        L = _lib.lib()
        dim = conf.dim
        n = pmid.shape[0]
        chan = tuple(mesh.shape[dim:])
        nchan = int(np.prod(chan)) if chan else 1
        sshape = tuple(mesh.shape[:dim])
        out = mesh.detach().clone().contiguous()
        disp_c = disp.detach().contiguous()
        val_c = val.detach().contiguous() if val is not None else None
        # the backward uses this context too: autograd may run it on another
        # thread, where context() would not see the caller's context provider
        c = _lib.context(conf.mesh_shape, conf.cell_size, disp.device)
        if val_c is None and nchan == 1 and is_default_cic(conf, sshape, offset, cell_size):
            _lib.check(L.pmx_scatter_mass(c.bind(), _lib.ptr(pmid), _lib.ptr(disp_c), n,
                                          float(val_scalar), _lib.ptr(out), 1))
        else:
            g = cic_geom(conf, sshape, offset, cell_size)
            _lib.check(L.pmx_scatter_cic(c.bind(), g, _lib.ptr(pmid), _lib.ptr(disp_c), n,
                                         _lib.ptr(val_c), float(val_scalar), nchan, _lib.ptr(out)))
        ctx.save_for_backward(disp_c, val_c if val_c is not None else None)
        ctx.pmid, ctx.conf, ctx.offset, ctx.cell_size = pmid, conf, offset, cell_size
        ctx.val_scalar, ctx.nchan, ctx.sshape = val_scalar, nchan, sshape
        ctx.pmx = c
        return out

    @staticmethod
    def backward(ctx, mesh_cot):
        disp, val = ctx.saved_tensors
        conf = ctx.conf
        n = ctx.pmid.shape[0]
        mesh_cot = mesh_cot.contiguous()
        disp_cot = torch.empty_like(disp) if ctx.needs_input_grad[0] else None
        val_cot = None
        if val is not None and ctx.needs_input_grad[2]:
            val_cot = torch.empty_like(val)
        if disp_cot is not None or val_cot is not None:
            c = ctx.pmx
            g = cic_geom(conf, ctx.sshape, ctx.offset, ctx.cell_size)
            _lib.check(_lib.lib().pmx_scatter_cic_adj(
                c.bind(), g, _lib.ptr(ctx.pmid), _lib.ptr(disp), n, _lib.ptr(mesh_cot),
                _lib.ptr(val), float(ctx.val_scalar), ctx.nchan, _lib.ptr(disp_cot),
                _lib.ptr(val_cot)))
        mesh_grad = mesh_cot if ctx.needs_input_grad[1] else None
        return disp_cot, mesh_grad, val_cot, None, None, None, None, None


def scatter(ptcl, conf, mesh=None, val=None, offset=0, cell_size=None):
    """Scatter particle values to the mesh multilinearly (scatter.py:8-30).

    Parameters follow pmwd: ``mesh`` defaults to zeros of
    ``conf.mesh_shape + val.shape[1:]``; ``val`` to ``conf.mesh_size /
    conf.ptcl_num``; ``offset`` is the mesh offset (0D broadcasts);
    ``cell_size`` the output mesh cell size (default ``conf.cell_size``).
    """
    pmid, disp = ptcl.pmid, ptcl.disp
    _lib.require_device(disp, 'ptcl.disp')
    dev = disp.device
    val_scalar = 0.0
    val_t = None
    if val is None:
        val_scalar = _val_default(conf)
    elif isinstance(val, torch.Tensor) and val.dim() > 0:
        val_t = val.to(device=dev, dtype=conf.float_dtype)
    else:
        val_scalar = float(val)
    chan = tuple(val_t.shape[1:]) if val_t is not None else ()
    if mesh is None:
        mesh = torch.zeros(tuple(conf.mesh_shape) + chan, dtype=conf.float_dtype, device=dev)
    else:
        mesh = torch.as_tensor(mesh).to(device=dev, dtype=conf.float_dtype)
    if tuple(mesh.shape[conf.dim:]) != chan:
        raise ValueError(f'channel shape mismatch: {tuple(mesh.shape[conf.dim:])} != {chan}')
    return _ScatterFn.apply(disp, mesh, val_t, pmid.contiguous(), conf, offset, cell_size,
                            val_scalar)
