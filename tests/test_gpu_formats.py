"""GPU parity of the particle formats and the standalone CIC/Poisson exports:
ptcl_enmesh (pmx_enmesh), Particles.pos/from_pos/raveled_id, pmx_gather_cic3
and pmx_poisson_grad_adj."""
import ctypes

import numpy as np
import pytest
import scipy.fft as sfft
import torch

from oracle import pm as opm

pytestmark = pytest.mark.gpu


def _ptcl(shape, ms, cell, std, seed):
    pmid, disp = opm.gen_grid(shape, ms, cell)
    rng = np.random.default_rng(seed)
    disp = (disp + std * rng.standard_normal(disp.shape)).astype(np.float32)
    disp[::11] = 0.0
    disp[1::13] = -0.0
    return pmid, disp


@pytest.mark.parametrize('shape,ratio,spacing', [((8, 8, 8), 2, 1.0), ((7, 9, 5), 1, 1.7)])
@pytest.mark.parametrize('kw', [dict(), dict(wrap=False), dict(wrap=False, drop=False),
                                dict(offset=0.3), dict(offset=(0.25, -0.5, 1.5), grad=True),
                                dict(cell_size=0.5, mesh_shape=None, grad=True),
                                dict(cell_size=0.75, wrap=False, grad=True)])
def test_ptcl_enmesh_matches_enmesh(dev, shape, ratio, spacing, kw):
    """ptcl_enmesh (particles.py:212-254) equals the oracle's enmesh
    (pm_util.py:33-156) bit for bit: indices, fractions, gradients."""
    import pmwd_amd as pa
    conf = pa.Configuration(spacing, shape, mesh_shape=ratio)
    pmid, disp = _ptcl(shape, conf.mesh_shape, conf.cell_size, 3.0 * conf.cell_size, 3)
    ptcl = pa.Particles(conf, torch.as_tensor(pmid, device=dev), torch.as_tensor(disp, device=dev))
    kw = dict(kw)
    cell = kw.get('cell_size')
    ms = tuple(conf.mesh_shape)
    if cell is not None:
        ms = tuple(int(round(s * conf.cell_size / cell)) for s in conf.mesh_shape)
        kw['mesh_shape'] = ms
    wrap, drop, grad = kw.get('wrap', True), kw.get('drop', True), kw.get('grad', False)
    out = pa.ptcl_enmesh(ptcl, conf, **kw)
    want = opm.enmesh(pmid, disp, conf.cell_size, conf.mesh_shape if wrap else None,
                      kw.get('offset', 0), cell, ms if drop else None, grad=grad)
    assert out[0].dtype == torch.int16
    np.testing.assert_array_equal(out[0].cpu().numpy(), want[0])
    np.testing.assert_array_equal(out[1].cpu().numpy(), want[1])
    if grad:
        np.testing.assert_array_equal(out[2].cpu().numpy(), want[2])


@pytest.mark.parametrize('fdt', [torch.float64, torch.float32])
def test_from_pos_pos_raveled_id(dev, fdt):
    """Particles.from_pos / pos / raveled_id (particles.py:81-107, 156-209)
    against the oracle, bit for bit: ties of rint at half cells, negative and
    beyond-the-box positions, with and without wrap."""
    import pmwd_amd as pa
    conf = pa.Configuration(1.5, (6, 5, 4), mesh_shape=2)   # mesh (12, 10, 8), box (9, 7.5, 6)
    rng = np.random.default_rng(9)
    box = np.asarray(conf.box_size)
    p = rng.uniform(-1.5, 2.5, size=(4000, 3)) * box
    p[:300] = np.round(p[:300] / conf.cell_size * 2) / 2 * conf.cell_size   # half-cell ties
    p = p.astype(np.float64 if fdt == torch.float64 else np.float32)
    for wrap in (True, False):
        ptcl = pa.Particles.from_pos(conf, torch.as_tensor(p, device=dev), wrap=wrap)
        wpm, wd = opm.from_pos(p, conf.cell_size, conf.mesh_shape, wrap=wrap)
        np.testing.assert_array_equal(ptcl.pmid.cpu().numpy(), wpm)
        np.testing.assert_array_equal(ptcl.disp.cpu().numpy(), wd)
        for rw in (False, True):
            rid = ptcl.raveled_id(wrap=rw).cpu().numpy()
            np.testing.assert_array_equal(rid.view(np.uint64), opm.raveled_id(wpm, conf.mesh_shape, rw))
        for dt, ndt in ((torch.float64, np.float64), (torch.float32, np.float32)):
            for pw in (True, False):
                got = ptcl.pos(dtype=dt, wrap=pw).cpu().numpy()
                np.testing.assert_array_equal(got, opm.pos(wpm, wd, conf.cell_size, conf.box_size,
                                                           ndt, pw))
    # positions of a grid survive the round trip pos -> from_pos -> pos
    g = pa.Particles.gen_grid(conf, device=dev)
    back = pa.Particles.from_pos(conf, g.pos())
    np.testing.assert_allclose(back.pos().cpu().numpy(), g.pos().cpu().numpy(), atol=1e-6)


def test_ptcl_pos_is_removed():
    import pmwd_amd as pa
    with pytest.raises(RuntimeError, match='replaced by ptcl.pos'):
        pa.ptcl_pos(None, None)


@pytest.mark.parametrize('spacing', [1.0, 1.953125])
def test_gather_cic3_equals_three_gathers(dev, spacing):
    import pmwd_amd as pa
    from pmwd_amd import _lib
    shape = (16, 16, 16)
    conf = pa.Configuration(spacing, shape, mesh_shape=2)
    pmid, disp = _ptcl(shape, conf.mesh_shape, conf.cell_size, 5.0, 4)
    mesh3 = np.random.default_rng(5).standard_normal((3,) + conf.mesh_shape).astype(np.float32)
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    pm, x, m3 = P(pmid), P(disp), P(mesh3)
    c = _lib.context(conf.mesh_shape, conf.cell_size, dev)
    acc = torch.empty_like(x)
    _lib.check(_lib.lib().pmx_gather_cic3(c.bind(), _lib.ptr(pm), _lib.ptr(x), len(pmid),
                                          _lib.ptr(m3), _lib.ptr(acc)))
    ptcl = pa.Particles(conf, pm, x)
    want = torch.stack([pa.gather(ptcl, conf, m3[d]) for d in range(3)], dim=1)
    np.testing.assert_array_equal(acc.cpu().numpy(), want.cpu().numpy())
    w = np.stack([opm.gather(pmid, disp, conf.cell_size, conf.mesh_shape, mesh3[d])
                  for d in range(3)], axis=1)
    np.testing.assert_array_equal(acc.cpu().numpy(), w)


@pytest.mark.parametrize('ms', [(16, 16, 16), (8, 12, 10), (7, 9, 6)])
def test_poisson_grad_adj(dev, ms):
    """pmx_poisson_grad_adj is the transpose of pmx_poisson_grad: against the
    oracle's adjoint chain and by <P d, g> = <d, P^T g> in float64."""
    from pmwd_amd import _lib
    cell, coef = 1.3, float(np.float32(0.45))
    rng = np.random.default_rng(6)
    dens = (1 + 0.3 * rng.standard_normal(ms)).astype(np.float32)
    gc = rng.standard_normal((3,) + ms).astype(np.float32)
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    c = _lib.context(ms, cell, dev)
    L = _lib.lib()
    grad = torch.empty((3,) + ms, dtype=torch.float32, device=dev)
    _lib.check(L.pmx_poisson_grad(c.bind(), _lib.ptr(P(dens)), coef, _lib.ptr(grad)))
    dc = torch.empty(ms, dtype=torch.float32, device=dev)
    _lib.check(L.pmx_poisson_grad_adj(c.bind(), _lib.ptr(P(gc)), coef, _lib.ptr(dc)))
    dc = dc.cpu().numpy()
    kvec = opm.fftfreq(ms, cell, dtype=np.float32)
    dsum = sum(opm.neg_grad(kvec[d], opm.laplace(kvec, sfft.rfftn(gc[d])), cell) for d in range(3))
    want = (-sfft.irfftn(dsum, s=ms)).astype(np.float32) * np.float32(coef)
    assert np.linalg.norm(dc - want) / np.linalg.norm(want) < 2e-6
    lhs = np.sum(grad.cpu().numpy().astype(np.float64) * gc)
    # the mean of dens does not enter (the k=0 mode is dropped): compare with dens - 1
    rhs = np.sum((dens.astype(np.float64) - 1) * dc)
    assert abs(lhs - rhs) < 1e-5 * np.sqrt(np.sum(grad.cpu().numpy().astype(np.float64)**2)
                                           * np.sum(gc.astype(np.float64)**2))


@pytest.mark.parametrize('shape,mesh', [((16, 16, 16), 2), ((16, 12, 10), 1.5),
                                        ((9, 7, 5), 3), ((8, 10, 12), (20, 25, 30)),
                                        ((6, 6, 6), 4.5), ((32, 32, 32), 1)])
@pytest.mark.parametrize('spacing', [1.0, 1.953125, 0.7])
def test_gen_grid_matches_oracle(dev, shape, mesh, spacing):
    """Particles.gen_grid (particles.py:109-144) for integer, non-integer and
    explicit mesh shapes: pmid and disp bit for bit against the oracle, and
    zero velocities/accelerations on request."""
    import pmwd_amd as pa
    conf = pa.Configuration(spacing, shape, mesh_shape=mesh)
    p = pa.Particles.gen_grid(conf, vel=True, acc=True, device=dev)
    pmid, disp = opm.gen_grid(shape, conf.mesh_shape, conf.cell_size)
    np.testing.assert_array_equal(p.pmid.cpu().numpy(), pmid)
    np.testing.assert_array_equal(p.disp.cpu().numpy(), disp)
    assert not p.vel.any() and not p.acc.any()
    # the Lagrangian positions are the particle grid
    pos = p.pos(dtype=torch.float64, wrap=False).cpu().numpy()
    grid = np.stack(np.meshgrid(*[np.arange(s) * spacing for s in shape], indexing='ij'),
                    -1).reshape(-1, 3)
    np.testing.assert_allclose(pos, grid, rtol=0, atol=1e-6 * max(conf.box_size))
