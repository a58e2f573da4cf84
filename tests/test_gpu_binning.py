"""Capacity binning on the GPU.

A binning pass after the first one sizes each tile from the counts of the
pass before (count + count/8 + 32) and writes the records straight from the
binning kernel; a tile that overflows its room makes the device re-bin
exactly.  Either way the deposit is an integer sum and the gather works per
particle, so every result must match, bit for bit, a fresh context's first
(exact) pass on the same input.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import pm as opm

pytestmark = pytest.mark.gpu


def grid(shape, ratio, cell):
    pmid, disp = opm.gen_grid(shape, tuple(ratio * s for s in shape), cell)
    return pmid, disp


def uniform(shape, ratio, cell, seed, std=0.7):
    pmid, disp = grid(shape, ratio, cell)
    rng = np.random.default_rng(seed)
    return pmid, (disp + std * cell * rng.standard_normal(disp.shape)).astype(np.float32)


def clustered(shape, ratio, cell, seed):
    """Every particle within about a cell of the box centre (one or two tiles)."""
    pmid, disp = grid(shape, ratio, cell)
    q = pmid.astype(np.float64) * cell + disp
    centre = 0.5 * np.array(shape) * ratio * cell
    rng = np.random.default_rng(seed)
    d = centre - q + 0.4 * cell * rng.standard_normal(q.shape)
    return pmid, d.astype(np.float32)


def new_ctx(conf, dev):
    from pmwd_amd import _lib
    return _lib.Context(conf.mesh_shape, conf.cell_size, dev)


def run_step(ctx, pmid, disp, vel, dev, acc=None):
    from pmwd_amd import _lib
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
    pm, d, v = P(pmid), P(disp), P(vel)
    a = P(acc) if acc is not None else torch.zeros_like(d)
    _lib.check(_lib.lib().pmx_nbody_step(ctx.bind(), _lib.ptr(pm), _lib.ptr(d), _lib.ptr(v),
                                         _lib.ptr(a), len(pmid), 1.0, 0.45, 0.3, 0.7, 0.25))
    torch.cuda.synchronize()
    return d.cpu().numpy(), v.cpu().numpy(), a.cpu().numpy()


def run_scatter(ctx, pmid, disp, dev, conf):
    from pmwd_amd import _lib
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    pm, d = P(pmid), P(disp)
    mesh = torch.empty(conf.mesh_shape, dtype=torch.float32, device=dev)
    _lib.check(_lib.lib().pmx_scatter_mass(ctx.bind(), _lib.ptr(pm), _lib.ptr(d), len(pmid), 1.0,
                                           _lib.ptr(mesh), 0))
    torch.cuda.synchronize()
    return mesh.cpu().numpy()


def assert_same(got, want):
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)


@pytest.mark.parametrize('shape,ratio', [((32, 32, 32), 1), ((20, 24, 12), 2)])
def test_step_capacity_matches_exact(dev, shape, ratio):
    import pmwd_amd as pa
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, disp = uniform(shape, ratio, conf.cell_size, 1)
    vel = (0.3 * np.random.default_rng(2).standard_normal(disp.shape)).astype(np.float32)
    ctx = new_ctx(conf, dev)
    want = run_step(ctx, pmid, disp, vel, dev)           # first pass: exact
    for _ in range(2):                                     # capacity passes
        assert_same(run_step(ctx, pmid, disp, vel, dev), want)
    # a second, different state through the same capacities
    pmid2, disp2 = uniform(shape, ratio, conf.cell_size, 3, std=1.5)
    got = run_step(ctx, pmid2, disp2, vel, dev)
    assert_same(got, run_step(new_ctx(conf, dev), pmid2, disp2, vel, dev))


def test_step_overflow_rebins_exactly(dev):
    """Uniform, then clustered (tiles overflow), then uniform again (the
    capacities of the clustered pass are too small nearly everywhere)."""
    import pmwd_amd as pa
    shape, ratio = (32, 32, 32), 1
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, du = uniform(shape, ratio, conf.cell_size, 4)
    _, dc = clustered(shape, ratio, conf.cell_size, 5)
    vel = (0.1 * np.random.default_rng(6).standard_normal(du.shape)).astype(np.float32)
    ctx = new_ctx(conf, dev)
    run_step(ctx, pmid, du, vel, dev)
    for d in (dc, du, dc):
        assert_same(run_step(ctx, pmid, d, vel, dev), run_step(new_ctx(conf, dev), pmid, d, vel, dev))


def test_step_fewer_and_more_particles(dev):
    """Capacities from a pass of a different particle count."""
    import pmwd_amd as pa
    conf = pa.Configuration(1.0, (16, 16, 16), mesh_shape=2)
    pmid_b, disp_b = uniform((16, 16, 16), 2, conf.cell_size, 7)
    pmid_s, disp_s = uniform((8, 8, 8), 4, conf.cell_size, 8)
    vb = np.zeros_like(disp_b)
    vs = np.zeros_like(disp_s)
    ctx = new_ctx(conf, dev)
    for pm, d, v in ((pmid_b, disp_b, vb), (pmid_s, disp_s, vs), (pmid_b, disp_b, vb)):
        assert_same(run_step(ctx, pm, d, v, dev), run_step(new_ctx(conf, dev), pm, d, v, dev))


def test_scatter_capacity_matches_exact(dev):
    import pmwd_amd as pa
    shape, ratio = (24, 24, 24), 2
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, du = uniform(shape, ratio, conf.cell_size, 9)
    _, dc = clustered(shape, ratio, conf.cell_size, 10)
    ctx = new_ctx(conf, dev)
    for d in (du, du, dc, du):
        got = run_scatter(ctx, pmid, d, dev, conf)
        np.testing.assert_array_equal(got, run_scatter(new_ctx(conf, dev), pmid, d, dev, conf))
    want = opm.scatter(pmid, du, conf.cell_size, conf.mesh_shape, val=1.0)
    np.testing.assert_allclose(run_scatter(ctx, pmid, du, dev, conf), want, rtol=0, atol=1e-5)


def test_step_adj_capacity_matches_exact(dev):
    """The fused reverse step bins through the same capacities."""
    import pmwd_amd as pa
    from pmwd_amd import _lib
    shape, ratio = (16, 16, 16), 2
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, disp = uniform(shape, ratio, conf.cell_size, 11)
    rng = np.random.default_rng(12)
    arrs = [disp] + [rng.standard_normal(disp.shape).astype(np.float32) for _ in range(5)]

    def adj(ctx):
        P = lambda a: torch.as_tensor(a, device=dev).clone()
        pm = P(pmid)
        d, v, a, dc, vc, ac = (P(x) for x in arrs)
        dots = (ctypes.c_double * 2)()
        _lib.check(_lib.lib().pmx_nbody_step_adj(
            ctx.bind(), _lib.ptr(pm), _lib.ptr(d), _lib.ptr(v), _lib.ptr(a), _lib.ptr(dc),
            _lib.ptr(vc), _lib.ptr(ac), len(pmid), 1.0, 0.45, 0.2, 0.3, 0.5, None, dots))
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in (d, v, a, dc, vc, ac)], list(dots)

    ctx = new_ctx(conf, dev)
    want, wdots = adj(ctx)
    for _ in range(2):
        got, dots = adj(ctx)
        assert_same(got, want)
        # per-tile float64 partial sums follow the in-tile record order
        np.testing.assert_allclose(dots, wdots, rtol=1e-12)
