"""pmx_nbody_steps: several KDK steps with the particles kept in tile order
on the device must give, bit for bit, what as many pmx_nbody_step calls
give -- including steps where tiles overflow their capacity and the device
redoes the step exactly."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import pm as opm

pytestmark = pytest.mark.gpu


def particles(shape, ratio, cell, seed, std=0.7):
    pmid, disp = opm.gen_grid(shape, tuple(ratio * s for s in shape), cell)
    rng = np.random.default_rng(seed)
    return pmid, (disp + std * cell * rng.standard_normal(disp.shape)).astype(np.float32)


def factors(k, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    f = (scale * (0.2 + 0.3 * rng.random((k, 3)))).astype(np.float32)
    return [np.ascontiguousarray(f[:, m]) for m in range(3)]


def run(conf, dev, pmid, disp, vel, k1, d, k2, multi):
    from pmwd_amd import _lib
    L = _lib.lib()
    ctx = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
    h = ctx.bind()
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
    pm, x, v = P(pmid), P(disp), P(vel)
    acc = torch.zeros_like(x)
    val, coef = 1.0, 0.45
    _lib.check(L.pmx_gravity(h, _lib.ptr(pm), _lib.ptr(x), len(pmid), val, coef, _lib.ptr(acc)))
    if multi:
        H = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _lib.check(L.pmx_nbody_steps(h, _lib.ptr(pm), _lib.ptr(x), _lib.ptr(v), _lib.ptr(acc),
                                     len(pmid), val, coef, len(k1), H(k1), H(d), H(k2)))
    else:
        for s in range(len(k1)):
            _lib.check(L.pmx_nbody_step(h, _lib.ptr(pm), _lib.ptr(x), _lib.ptr(v), _lib.ptr(acc),
                                        len(pmid), val, coef, float(k1[s]), float(d[s]),
                                        float(k2[s])))
    torch.cuda.synchronize()
    return x.cpu().numpy(), v.cpu().numpy(), acc.cpu().numpy()


def assert_same(a, b):
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize('shape,ratio,k,spacing', [((32, 32, 32), 1, 5, 1.0),
                                                   ((20, 24, 12), 2, 4, 1.0),
                                                   ((7, 8, 6), 1, 3, 1.0),
                                                   ((16, 16, 16), 2, 1, 1.0),
                                                   ((12, 10, 8), 1, 2, 1.0),
                                                   ((16, 16, 16), 1, 4, 1.3),
                                                   ((12, 8, 20), 2, 3, 0.7),
                                                   ((1056, 2, 4), 1, 2, 1.0)])
def test_steps_match_single_steps(dev, shape, ratio, k, spacing):
    """(1056, 2, 4): an axis beyond the 1024 cells the tile-ordered state
    packs, which the library runs one step at a time."""
    import pmwd_amd as pa
    conf = pa.Configuration(spacing, shape, mesh_shape=ratio)
    pmid, disp = particles(shape, ratio, conf.cell_size, 1)
    vel = (0.5 * np.random.default_rng(2).standard_normal(disp.shape)).astype(np.float32)
    k1, d, k2 = factors(k, 3)
    got = run(conf, dev, pmid, disp, vel, k1, d, k2, True)
    want = run(conf, dev, pmid, disp, vel, k1, d, k2, False)
    assert_same(got, want)


def test_steps_collapse_overflows_capacity(dev):
    """Particles streaming to the box centre: tiles near it outgrow the
    capacities sized from the step before, which forces the exact redo."""
    import pmwd_amd as pa
    shape, ratio = (32, 32, 32), 1
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, disp = particles(shape, ratio, conf.cell_size, 4, std=0.2)
    q = pmid.astype(np.float64) * conf.cell_size + disp
    centre = 0.5 * np.array(shape) * conf.cell_size
    vel = (0.3 * (centre - q)).astype(np.float32)
    k1 = np.zeros(6, np.float32)
    k2 = np.zeros(6, np.float32)
    d = np.full(6, 0.5, np.float32)
    got = run(conf, dev, pmid, disp, vel, k1, d, k2, True)
    want = run(conf, dev, pmid, disp, vel, k1, d, k2, False)
    assert_same(got, want)
    # the run did collapse: most particles end within a few cells of the centre
    x = pmid.astype(np.float64) * conf.cell_size + got[0]
    r = np.abs(((x - centre) + 16) % 32 - 16).max(axis=1)
    assert np.mean(r < 6) > 0.5


def test_steps_large_moves(dev):
    """Several cells per step in random directions: most particles change
    tile every step."""
    import pmwd_amd as pa
    shape, ratio = (24, 24, 24), 2
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, disp = particles(shape, ratio, conf.cell_size, 5)
    vel = (8.0 * np.random.default_rng(6).standard_normal(disp.shape)).astype(np.float32)
    k1, d, k2 = factors(5, 7)
    assert_same(run(conf, dev, pmid, disp, vel, k1, d, k2, True),
                run(conf, dev, pmid, disp, vel, k1, d, k2, False))


def test_nbody_driver_pipeline_matches_per_step(dev):
    """pmwd_amd.nbody's forward loop (with checkpoints between segments)."""
    import importlib
    import pmwd_amd as pa
    nbm = importlib.import_module('pmwd_amd.nbody')
    conf = pa.Configuration(4.0, (16, 16, 16), mesh_shape=2, a_start=1 / 8,
                            a_nbody_maxstep=1 / 16)
    cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
    pmid, disp = particles((16, 16, 16), 2, conf.cell_size, 8, std=1.0)
    vel = (0.1 * np.random.default_rng(9).standard_normal(disp.shape)).astype(np.float32)
    P = lambda a: torch.as_tensor(a, device=dev).clone()
    outs = []
    for pipeline in (True, False):
        for every in (None, 3):
            d, v = P(disp), P(vel)
            acc, ckpt = nbm._run_forward(P(pmid), d, v, cosmo, conf, checkpoint_every=every,
                                         pipeline=pipeline)
            outs.append((d.cpu().numpy(), v.cpu().numpy(), acc.cpu().numpy(),
                         {i: (a.cpu().numpy(), b.cpu().numpy()) for i, (a, b) in ckpt.items()}))
    ref = outs[3]
    for o in outs[:3]:
        assert_same(o[:3], ref[:3])
    for o in (outs[1],):
        assert o[3].keys() == ref[3].keys()
        for i in ref[3]:
            assert_same(o[3][i], ref[3][i])


def test_trim_between_calls(dev):
    """Freeing the pipeline copy, or all particle scratch, between calls
    changes nothing: the next call reallocates what it needs."""
    import pmwd_amd as pa
    from pmwd_amd import _lib
    shape, ratio = (16, 16, 16), 2
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, disp = particles(shape, ratio, conf.cell_size, 10)
    vel = (0.5 * np.random.default_rng(11).standard_normal(disp.shape)).astype(np.float32)
    k1, d, k2 = factors(3, 12)
    L = _lib.lib()
    ctx = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
    H = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
    outs = []
    full = ctx.nbytes
    for trim in (None, 'pipeline', 'all'):
        if trim == 'pipeline':
            ctx.trim(pipeline_only=True)
            assert ctx.nbytes < full
        elif trim == 'all':
            ctx.trim()
            assert ctx.nbytes < full
        pm, x, v = P(pmid), P(disp), P(vel)
        acc = torch.zeros_like(x)
        h = ctx.bind()
        _lib.check(L.pmx_nbody_steps(h, _lib.ptr(pm), _lib.ptr(x), _lib.ptr(v), _lib.ptr(acc),
                                     len(pmid), 1.0, 0.45, 3, H(k1), H(d), H(k2)))
        torch.cuda.synchronize()
        outs.append((x.cpu().numpy(), v.cpu().numpy(), acc.cpu().numpy()))
        full = ctx.nbytes
    assert_same(outs[1], outs[0])
    assert_same(outs[2], outs[0])


def run_adj(conf, dev, pmid, arrs, kp, ka, d, kb, multi, snap=None, coef=0.45):
    """Reverse steps from state arrs = (disp, vel, disp_cot, vel_cot): the
    force VJP at the start, then pmx_nbody_steps_adj or the chained
    pmx_nbody_step_adj calls; returns the six arrays and the dots."""
    from pmwd_amd import _lib
    L = _lib.lib()
    ctx = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
    h = ctx.bind()
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
    pm = P(pmid)
    x, v, dc, vc = (P(a) for a in arrs)
    acc, ac = torch.empty_like(x), torch.empty_like(x)
    n, val = len(pmid), 1.0
    dot = ctypes.c_double()
    _lib.check(L.pmx_gravity_vjp(h, _lib.ptr(pm), _lib.ptr(x), n, val, coef, _lib.ptr(vc),
                                 _lib.ptr(acc), _lib.ptr(ac), ctypes.byref(dot)))
    sn = P(snap) if snap is not None else None
    k = len(ka)
    dots = np.zeros(2 * k)
    if multi:
        H = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        dd = (ctypes.c_double * (2 * k))()
        _lib.check(L.pmx_nbody_steps_adj(h, _lib.ptr(pm), _lib.ptr(x), _lib.ptr(v), _lib.ptr(acc),
                                         _lib.ptr(dc), _lib.ptr(vc), _lib.ptr(ac), n, val, coef,
                                         kp, k, H(ka), H(d), H(kb), _lib.ptr(sn), dd))
        dots[:] = list(dd)
    else:
        dd = (ctypes.c_double * 2)()
        kpi = kp
        for s in range(k):
            _lib.check(L.pmx_nbody_step_adj(h, _lib.ptr(pm), _lib.ptr(x), _lib.ptr(v),
                                            _lib.ptr(acc), _lib.ptr(dc), _lib.ptr(vc),
                                            _lib.ptr(ac), n, val, coef, kpi, float(ka[s]),
                                            float(d[s]), _lib.ptr(sn) if s == k - 1 else None,
                                            dd))
            dots[2 * s:2 * s + 2] = list(dd)
            kpi = float(kb[s])
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (x, v, dc, vc, acc, ac)], dots


@pytest.mark.parametrize('shape,ratio,k,spacing,kp,snap', [
    ((16, 16, 16), 2, 4, 1.0, 0.3, False), ((32, 32, 32), 1, 3, 1.0, 0.0, True),
    ((12, 10, 8), 1, 1, 1.0, 0.2, True), ((7, 8, 6), 1, 2, 1.3, 0.0, False),
    ((12, 8, 20), 2, 3, 0.7, 0.25, True), ((1056, 2, 4), 1, 2, 1.0, 0.2, True)])
def test_steps_adj_match_step_adj(dev, shape, ratio, k, spacing, kp, snap):
    import pmwd_amd as pa
    conf = pa.Configuration(spacing, shape, mesh_shape=ratio)
    pmid, disp = particles(shape, ratio, conf.cell_size, 20)
    rng = np.random.default_rng(21)
    arrs = [disp] + [(0.5 * rng.standard_normal(disp.shape)).astype(np.float32)
                     for _ in range(3)]
    ka, d, kb = factors(k, 22)
    d = -d  # reverse time
    sn = (disp + 0.3 * rng.standard_normal(disp.shape)).astype(np.float32) if snap else None
    got, gd = run_adj(conf, dev, pmid, arrs, kp, ka, d, kb, True, sn)
    want, wd = run_adj(conf, dev, pmid, arrs, kp, ka, d, kb, False, sn)
    assert_same(got, want)
    np.testing.assert_allclose(gd, wd, rtol=1e-12, atol=1e-300)


def test_steps_adj_collapse_overflows_capacity(dev):
    import pmwd_amd as pa
    shape, ratio = (32, 32, 32), 1
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, disp = particles(shape, ratio, conf.cell_size, 23, std=0.2)
    q = pmid.astype(np.float64) * conf.cell_size + disp
    centre = 0.5 * np.array(shape) * conf.cell_size
    vel = (0.3 * (centre - q)).astype(np.float32)
    rng = np.random.default_rng(24)
    arrs = [disp, vel] + [rng.standard_normal(disp.shape).astype(np.float32) for _ in range(2)]
    ka = np.zeros(5, np.float32)
    kb = np.zeros(5, np.float32)
    d = np.full(5, 0.5, np.float32)
    got, gd = run_adj(conf, dev, pmid, arrs, 0.0, ka, d, kb, True)
    want, wd = run_adj(conf, dev, pmid, arrs, 0.0, ka, d, kb, False)
    assert_same(got, want)
    np.testing.assert_allclose(gd, wd, rtol=1e-12, atol=1e-300)


def test_adjoint_driver_pipeline_matches_per_step(dev):
    """pmwd_amd.nbody's reverse loop, with and without checkpoint snaps."""
    import importlib
    import pmwd_amd as pa
    nbm = importlib.import_module('pmwd_amd.nbody')
    conf = pa.Configuration(4.0, (16, 16, 16), mesh_shape=2, a_start=1 / 8,
                            a_nbody_maxstep=1 / 16)
    cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
    pmid, disp = particles((16, 16, 16), 2, conf.cell_size, 25, std=1.0)
    rng = np.random.default_rng(26)
    vel = (0.1 * rng.standard_normal(disp.shape)).astype(np.float32)
    dc0 = rng.standard_normal(disp.shape).astype(np.float32)
    vc0 = rng.standard_normal(disp.shape).astype(np.float32)
    P = lambda a: torch.as_tensor(a, device=dev).clone()
    for every in (None, 3):
        d, v = P(disp), P(vel)
        _, ckpt = nbm._run_forward(P(pmid), d, v, cosmo, conf, checkpoint_every=every)
        outs = []
        for pipeline in (True, False):
            x, vv, dc, vc = d.clone(), v.clone(), P(dc0), P(vc0)
            acc, ac, om = nbm._run_adjoint(P(pmid), x, vv, dc, vc, cosmo, conf, ckpt=ckpt,
                                           pipeline=pipeline)
            outs.append(([t.cpu().numpy() for t in (x, vv, dc, vc, acc, ac)], om))
        assert_same(outs[0][0], outs[1][0])
        assert abs(outs[0][1] - outs[1][1]) <= 1e-10 * abs(outs[1][1])


# row blocks of 64, 32, 16 and 8 rows (rows of 64, 128, 256 and 512 complex
# values): a block spans four, two or one tile rows in y, or half of one
@pytest.mark.parametrize('shape,ratio', [((64, 64, 64), 2), ((32, 48, 64), 4),
                                         ((16, 16, 128), 4), ((8, 16, 256), 4)])
def test_fold_on_load_matches_fold_kernel(dev, monkeypatch, shape, ratio):
    """On meshes whose Poisson row pass adds the deposit's halo shells as it
    loads the density (16^3 tiles, half-length rows of 64..1024 complex
    values), forward and reverse steps -- chained single steps and the
    multi-step entry points, the reverse ones with their three cotangent
    channels -- equal bit for bit the runs that fold with the separate
    kernel (PMX_FOLD_ON_LOAD=0 when the context is made)."""
    import pmwd_amd as pa
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, disp = particles(shape, ratio, conf.cell_size, 31, std=0.8)
    rng = np.random.default_rng(32)
    vel = (0.5 * rng.standard_normal(disp.shape)).astype(np.float32)
    k1, d, k2 = factors(2, 33)
    arrs = [disp] + [(0.5 * rng.standard_normal(disp.shape)).astype(np.float32)
                     for _ in range(3)]
    ka, da, kb = factors(2, 34)
    got = [run(conf, dev, pmid, disp, vel, k1, d, k2, m) for m in (True, False)]
    got_adj = [run_adj(conf, dev, pmid, arrs, 0.2, ka, -da, kb, m) for m in (True, False)]
    monkeypatch.setenv('PMX_FOLD_ON_LOAD', '0')
    want = run(conf, dev, pmid, disp, vel, k1, d, k2, False)
    want_adj, want_dots = run_adj(conf, dev, pmid, arrs, 0.2, ka, -da, kb, False)
    for g in got:
        assert_same(g, want)
    for g, gd in got_adj:
        assert_same(g, want_adj)
        np.testing.assert_allclose(gd, want_dots, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize('shape,ratio', [((64, 64, 64), 2), ((32, 48, 64), 4)])
def test_deposit_split_matches_channels_in_turn(dev, monkeypatch, shape, ratio):
    """The reverse steps' three-channel deposits run one workgroup per tile
    and channel; chained single steps and the multi-step entry point equal
    bit for bit the runs whose deposits take the channels in turn in one
    workgroup (PMX_DEPOSIT_SPLIT=0 when the context is made)."""
    import pmwd_amd as pa
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, disp = particles(shape, ratio, conf.cell_size, 35, std=0.8)
    rng = np.random.default_rng(36)
    arrs = [disp] + [(0.5 * rng.standard_normal(disp.shape)).astype(np.float32)
                     for _ in range(3)]
    ka, da, kb = factors(2, 37)
    got = [run_adj(conf, dev, pmid, arrs, 0.2, ka, -da, kb, m) for m in (True, False)]
    monkeypatch.setenv('PMX_DEPOSIT_SPLIT', '0')
    want = [run_adj(conf, dev, pmid, arrs, 0.2, ka, -da, kb, m) for m in (True, False)]
    for (g, gd), (w, wd) in zip(got, want):
        assert_same(g, w)
        np.testing.assert_allclose(gd, wd, rtol=1e-12, atol=1e-300)
