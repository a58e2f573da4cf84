"""Serving: each request's result is the same alone and in a concurrent mix."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _requests():
    from pmwd_amd.serving import SimRequest
    kw = dict(a_start=1 / 8, a_nbody_maxstep=1 / 8)
    return [SimRequest(seed=11, ptcl_grid=16, mesh_ratio=2, box=32.0, Omega_m=0.28, sigma8=0.75,
                       grad=True, **kw),
            SimRequest(seed=12, ptcl_grid=8, mesh_ratio=2, box=16.0, Omega_m=0.33, sigma8=0.85, **kw),
            SimRequest(seed=13, ptcl_grid=16, mesh_ratio=1, box=32.0, Omega_m=0.30, sigma8=0.80, **kw),
            SimRequest(seed=14, ptcl_grid=8, mesh_ratio=2, box=16.0, Omega_m=0.26, sigma8=0.72,
                       grad=True, **kw),
            SimRequest(seed=15, ptcl_grid=32, mesh_ratio=2, box=64.0, Omega_m=0.31, sigma8=0.88, **kw),
            SimRequest(seed=16, ptcl_grid=16, mesh_ratio=2, box=32.0, Omega_m=0.34, sigma8=0.78,
                       grad=True, **kw)]


def test_results_same_alone_and_in_the_mix(dev):
    from pmwd_amd.serving import Server
    reqs = _requests()
    solo = Server(dev, max_slots=1)
    alone = [solo.submit(r).result() for r in reqs]
    solo.close()
    assert solo.peak_active == 1
    mix_server = Server(dev, max_slots=len(reqs))
    futs = [mix_server.submit(r) for r in reqs]
    mix = [f.result() for f in futs]
    mix_server.close()
    assert mix_server.peak_active > 1
    for a, m in zip(alone, mix):
        assert a.request == m.request
        np.testing.assert_array_equal(a.dens, m.dens)
        assert np.isfinite(a.dens).all() and abs(a.dens.mean() - 1) < 1e-5
        if a.request.grad:
            assert a.loss == m.loss
            np.testing.assert_array_equal(a.grad_modes, m.grad_modes)
            # float64 tile partial sums may be added in another order
            assert abs(a.grad_Omega_m - m.grad_Omega_m) <= 1e-9 * abs(a.grad_Omega_m)
            assert a.grad_sigma8 == m.grad_sigma8 or \
                abs(a.grad_sigma8 - m.grad_sigma8) <= 1e-9 * abs(a.grad_sigma8)


def test_server_admits_under_a_tight_budget(dev):
    """With room for one request at a time the server still completes them
    all (serially), evicting idle contexts as needed."""
    from pmwd_amd.serving import Server, request_bytes
    reqs = _requests()[:3]
    budget = max(request_bytes(r) for r in reqs) + 1
    s = Server(dev, hbm_budget=budget, max_slots=3)
    res = [f.result() for f in [s.submit(r) for r in reqs]]
    s.close()
    assert len(res) == 3 and s.peak_active == 1


def test_gradient_request_uses_only_pooled_contexts(dev, monkeypatch):
    """A gradient request's backward (run by autograd, possibly on another
    thread) uses the request's pooled contexts, not the process-wide cache,
    so the server's byte accounting and eviction see every context."""
    from pmwd_amd import _lib
    from pmwd_amd.serving import Server, SimRequest
    _lib.clear_contexts()
    used = set()
    real_bind = _lib.Context.bind

    def bind(self):
        used.add(id(self))
        return real_bind(self)
    monkeypatch.setattr(_lib.Context, 'bind', bind)
    req = SimRequest(seed=21, ptcl_grid=16, mesh_ratio=2, box=32.0, Omega_m=0.3, sigma8=0.8,
                     grad=True, a_start=1 / 8, a_nbody_maxstep=1 / 8)
    s = Server(dev, max_slots=1)
    res = s.submit(req).result()
    s.close()
    assert res.grad_modes is not None and np.isfinite(res.grad_modes).all()
    assert _lib._ctx_cache == {}, list(_lib._ctx_cache)
    pooled = {id(c) for cs in s._pool.values() for c in cs}
    assert used and used <= pooled, (len(used), len(pooled))
    for cs in s._pool.values():
        for c in cs:
            c.close()


@pytest.mark.parametrize('mode', ['pack', 'replay'])
def test_packed_graph_runs_equal_the_eager_path(dev, mode):
    """Forward-only requests of one configuration submitted together replay
    their slots' captured graphs once each slot has captured its graph: in
    packs (several slots' graphs in one launch) or each on its slot's
    stream; every density equals, bit for bit, the uncaptured public-API
    path (nbody then scatter) of run_request."""
    from pmwd_amd import _lib
    from pmwd_amd.serving import Server, SimRequest, run_request
    kw = dict(a_start=1 / 8, a_nbody_maxstep=1 / 8)
    reqs = [SimRequest(seed=40 + i, ptcl_grid=16, mesh_ratio=2, box=32.0,
                       Omega_m=0.26 + 0.02 * i, sigma8=0.72 + 0.03 * i, **kw) for i in range(8)]
    with torch.cuda.stream(torch.cuda.Stream(dev)):
        want = [run_request(r, dev)['dens'] for r in reqs]
    s = Server(dev, max_slots=8, pack_window=0.05, pack_mode=mode)
    # a slot's first run is eager and captures its graph; later runs replay
    # it, several slots per launch when they load within the window
    got = []
    for _ in range(3):
        got.append([f.result() for f in [s.submit(r) for r in reqs]])
    stats = s.pack_stats()
    s.close()
    st = next(iter(stats.values()))
    assert 1 <= st['eager_runs'] <= st['slots'], stats
    assert st['packed_runs'] + st['eager_runs'] == 3 * len(reqs), stats
    if mode == 'pack':
        assert st['mean_pack_size'] > 1, stats
    else:
        assert st['mean_pack_size'] == 1, stats
    for rnd in got:
        for r, w in zip(rnd, want):
            np.testing.assert_array_equal(r.dens, w)


def test_private_streams_are_not_pooled(dev):
    """Slots capture on streams of their own: distinct from each other and
    from every stream of torch's pool, handed back for reuse on close; the
    C pair creates and destroys a stream torch never saw."""
    import ctypes
    from pmwd_amd import _lib
    pooled = {torch.cuda.Stream(dev).cuda_stream for _ in range(40)}
    ps = [_lib.PrivateStream(dev) for _ in range(3)]
    hs = [p.stream.cuda_stream for p in ps]
    assert len(set(hs)) == 3 and not set(hs) & pooled
    ps[0].close()
    again = _lib.PrivateStream(dev)
    assert again.stream.cuda_stream == hs[0]
    for p in ps[1:] + [again]:
        p.close()
    h = ctypes.c_void_p()
    _lib.check(_lib.lib().pmx_stream_create(torch.device(dev).index or 0, ctypes.byref(h)))
    assert h.value and h.value not in pooled
    _lib.check(_lib.lib().pmx_stream_destroy(h))


@pytest.mark.parametrize('mode', ['pack', 'replay'])
def test_failed_run_returns_its_slot(dev, mode):
    """A run that fails (initial conditions of the wrong shape) raises, and
    its slot goes back: afterwards every slot serves requests, with results
    bitwise those of the uncaptured path."""
    from pmwd_amd.serving import PackEngine, SimRequest, run_request, nbody_factor_table
    import pmwd_amd as pa
    kw = dict(a_start=1 / 8, a_nbody_maxstep=1 / 8)
    reqs = [SimRequest(seed=60 + i, ptcl_grid=16, mesh_ratio=2, box=32.0,
                       Omega_m=0.27 + 0.01 * i, sigma8=0.8, **kw) for i in range(3)]
    conf = reqs[0].conf()
    eng = PackEngine(conf, dev, nslots=2, window=0.0, mode=mode)
    try:
        cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
        bad = torch.zeros((5, 3), device=dev)
        for _ in range(3):
            with pytest.raises(RuntimeError):
                eng.run(bad, bad, nbody_factor_table(cosmo, conf))
        assert sorted(eng._free) == [0, 1] and eng._loading == 0
        for _ in range(2):   # eager runs that capture, then replays
            for r in reqs:
                got = run_request(r, dev, eng)['dens']
                np.testing.assert_array_equal(got, run_request(r, dev)['dens'])
        assert sorted(eng._free) == [0, 1] and eng._loading == 0
    finally:
        eng.close()
