"""Particle state (pmwd/particles.py:18-346) held as device tensors.

``pmid`` is (N, dim) int16, ``disp``/``vel``/``acc`` are (N, dim) float32,
row-major as in pmwd, so the HIP kernels read them in place.
"""
import ctypes
from dataclasses import dataclass, field, replace as _dc_replace
from typing import Any, Optional

import numpy as np
import torch

from .configuration import Configuration
from .cosmology import E2


def default_device():
    return torch.device('cuda') if torch.cuda.is_available() else torch.device('cpu')


@dataclass(frozen=True)
class Particles:
    """Particle state: ``pmid``, ``disp``, optional ``vel``, ``acc``, ``attr``.

    Tensors are cast to ``conf.pmid_dtype`` / ``conf.float_dtype`` on
    construction (particles.py:61-73).
    """

    conf: Configuration = field(repr=False)
    pmid: Any
    disp: Any
    vel: Optional[Any] = None
    acc: Optional[Any] = None
    attr: Any = None

    def __post_init__(self):
        conf = self.conf
        dev = None
        for name in ('pmid', 'disp', 'vel', 'acc'):
            v = getattr(self, name)
            if v is None:
                continue
            dt = conf.pmid_dtype if name == 'pmid' else conf.float_dtype
            if not isinstance(v, torch.Tensor):
                v = torch.as_tensor(np.asarray(v), device=dev or default_device())
            if v.dtype != dt:
                v = v.to(dt)
            dev = dev or v.device
            object.__setattr__(self, name, v)

    def replace(self, **changes):
        return _dc_replace(self, **changes)

    def __len__(self):
        return len(self.pmid)

    def __getitem__(self, key):
        kw = {n: (getattr(self, n)[key] if getattr(self, n) is not None else None)
              for n in ('pmid', 'disp', 'vel', 'acc')}
        return Particles(self.conf, attr=self.attr, **kw)

    @property
    def device(self):
        return self.pmid.device

    @classmethod
    def from_pos(cls, conf, pos, wrap=True):
        """pmid and disp from positions (particles.py:81-107)."""
        pos = torch.as_tensor(pos)
        if pos.device.type == 'cpu' and torch.cuda.is_available():
            pos = pos.to(default_device())
        pmid = torch.round(pos / conf.cell_size)
        disp = pos - pmid * conf.cell_size
        pmid = pmid.to(conf.pmid_dtype)
        disp = disp.to(conf.float_dtype)
        if wrap:
            shape = torch.tensor(conf.mesh_shape, dtype=conf.pmid_dtype, device=pmid.device)
            pmid = torch.remainder(pmid, shape)
        return cls(conf, pmid, disp)

    @classmethod
    def gen_grid(cls, conf, vel=False, acc=False, device=None):
        """Particles on the Lagrangian grid (particles.py:109-144)."""
        device = torch.device(device) if device is not None else default_device()
        pmid, disp = [], []
        for sp, sm in zip(conf.ptcl_grid_shape, conf.mesh_shape):
            p1 = np.rint(np.linspace(0, sm, num=sp, endpoint=False)).astype(np.int16)
            d1 = (np.arange(sp) * sm - p1.astype(np.int64) * sp) * (conf.cell_size / sp)
            pmid.append(torch.as_tensor(p1, device=device).to(conf.pmid_dtype))
            disp.append(torch.as_tensor(d1.astype(np.float32), device=device).to(conf.float_dtype))
        pmid = torch.stack(torch.meshgrid(*pmid, indexing='ij'), dim=-1).reshape(-1, conf.dim)
        disp = torch.stack(torch.meshgrid(*disp, indexing='ij'), dim=-1).reshape(-1, conf.dim)
        v = torch.zeros_like(disp) if vel else None
        a = torch.zeros_like(disp) if acc else None
        return cls(conf, pmid.contiguous(), disp.contiguous(), vel=v, acc=a)

    def raveled_id(self, dtype=torch.int64, wrap=False):
        """C-order raveled ids from pmid (particles.py:156-182)."""
        conf = self.conf
        pmid = self.pmid.to(torch.int64)
        shape = conf.mesh_shape
        if wrap:
            pmid = torch.remainder(pmid, torch.tensor(shape, device=pmid.device))
        out = torch.zeros(len(pmid), dtype=torch.int64, device=pmid.device)
        for d, s in enumerate(shape):
            out = out * s + pmid[:, d]
        return out.to(dtype)

    def pos(self, dtype=torch.float64, wrap=True):
        """Positions pmid*cell + disp (particles.py:184-209)."""
        conf = self.conf
        pos = self.pmid.to(dtype) * conf.cell_size + self.disp.to(dtype)
        if wrap:
            pos = torch.remainder(pos, torch.tensor(conf.box_size, dtype=dtype, device=pos.device))
        return pos


def ptcl_enmesh(ptcl, conf, offset=0, cell_size=None, mesh_shape=None, wrap=True, drop=True,
                grad=False):
    """Multilinear mesh indices and fractions of particles (particles.py:212-254).

    Returns ``ind`` (ptcl_num, 2**dim, dim) of ``conf.pmid_dtype``, ``frac``
    (ptcl_num, 2**dim) and, with ``grad``, ``frac_grad`` (ptcl_num, 2**dim,
    dim), computed on the device by ``pmx_enmesh`` with pmwd's enmesh
    arithmetic (pm_util.py:33-156).  ``wrap`` wraps around the periodic
    ``conf.mesh_shape``; without it, ``drop`` replaces negative indices by
    the ``mesh_shape`` extent so that they fall outside the mesh.
    """
    from . import _lib
    from .pm_util import cic_geom
    _lib.require_device(ptcl.disp, 'ptcl.disp')
    if mesh_shape is None:
        mesh_shape = conf.mesh_shape
    dim = conf.dim
    n = len(ptcl)
    dev = ptcl.disp.device
    ind = torch.empty((n, 2**dim, dim), dtype=torch.int16, device=dev)
    frac = torch.empty((n, 2**dim), dtype=torch.float32, device=dev)
    fg = torch.empty((n, 2**dim, dim), dtype=torch.float32, device=dev) if grad else None
    g = cic_geom(conf, mesh_shape, offset, cell_size)
    c = _lib.context(conf.mesh_shape, conf.cell_size, dev)
    pmid = ptcl.pmid.to(torch.int16).contiguous()
    disp = ptcl.disp.to(torch.float32).contiguous()
    _lib.check(_lib.lib().pmx_enmesh(c.bind(), ctypes.byref(g), int(bool(wrap)), int(bool(drop)),
                                     _lib.ptr(pmid), _lib.ptr(disp), n, _lib.ptr(ind),
                                     _lib.ptr(frac), _lib.ptr(fg)))
    ind = ind.to(conf.pmid_dtype)
    if grad:
        return ind, frac, fg
    return ind, frac


def ptcl_pos(ptcl, conf, dtype=float, wrap=True):
    """Removed from pmwd in favour of ``Particles.pos`` (particles.py:257-258)."""
    raise RuntimeError('Deprecated and replaced by ptcl.pos')


def ptcl_rpos(ptcl, ref, conf, wrap=True):
    """Positions relative to reference points/particles (particles.py:261-292)."""
    if not isinstance(ref, Particles):
        ref = Particles.from_pos(conf, ref, wrap=False)
    rpos = (ptcl.pmid - ref.pmid).to(conf.float_dtype) * conf.cell_size
    rpos = rpos + (ptcl.disp - ref.disp)
    if wrap:
        box = torch.tensor(conf.box_size, dtype=conf.float_dtype, device=rpos.device)
        rpos = rpos - torch.round(rpos / box) * box
    return rpos


def ptcl_rsd(ptcl, los, a, cosmo):
    """Redshift-space displacement (particles.py:295-324)."""
    conf = cosmo.conf
    los = torch.as_tensor(los, dtype=conf.float_dtype, device=ptcl.vel.device)
    a = torch.as_tensor(a, dtype=torch.float64)
    E = torch.as_tensor(np.sqrt(E2(a.cpu().numpy(), cosmo)), dtype=conf.float_dtype,
                        device=ptcl.vel.device)
    a = a.to(conf.float_dtype).to(ptcl.vel.device)
    rsd = (ptcl.vel * los).sum(dim=1, keepdim=True)
    return rsd * los / (a**2 * E)


def ptcl_los(ptcl, obs, conf):
    """Line-of-sight unit vectors (particles.py:327-346)."""
    los = ptcl_rpos(ptcl, obs, conf, wrap=False)
    return los / torch.linalg.norm(los, dim=1, keepdim=True)
