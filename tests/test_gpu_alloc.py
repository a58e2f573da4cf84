"""GPU: contexts on torch's caching allocator (the host's default,
pmx_set_allocator) against contexts on hipMalloc."""
import os

import numpy as np
import pytest
import torch

from oracle import pm as opm

pytestmark = pytest.mark.gpu


def _hooked_by_default():
    return os.environ.get('PMX_TORCH_ALLOCATOR', '1') != '0'


def _make(conf, hooked, dev):
    """A fresh Context built with (hooked) or without torch's allocator, and
    the bytes of torch's pool it took."""
    from pmwd_amd import _lib
    L = _lib.lib()
    if hooked:
        _lib.check(L.pmx_set_allocator(_lib._torch_alloc, _lib._torch_free, None))
    else:
        _lib.check(L.pmx_set_allocator(_lib._ALLOC_T(), _lib._FREE_T(), None))
    try:
        torch.cuda.synchronize(dev)
        before = torch.cuda.memory_allocated(dev)
        ctx = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
        took = torch.cuda.memory_allocated(dev) - before
    finally:
        if _hooked_by_default():
            _lib.check(L.pmx_set_allocator(_lib._torch_alloc, _lib._torch_free, None))
        else:
            _lib.check(L.pmx_set_allocator(_lib._ALLOC_T(), _lib._FREE_T(), None))
    return ctx, took


def _run(conf, ctx, pmid, disp, vel, dev):
    import pmwd_amd as pa
    from pmwd_amd import _lib
    with _lib.context_scope(lambda key, make: ctx):
        P = lambda a: torch.as_tensor(a, device=dev)
        ptcl = pa.Particles(conf, P(pmid), P(disp), P(vel))
        dens = pa.scatter(ptcl, conf)
        cosmo = pa.boltzmann(pa.Cosmology.from_sigma8(conf, 0.8, n_s=0.96, Omega_m=0.3,
                                                      Omega_b=0.05, h=0.7), conf)
        out, _ = pa.nbody(ptcl, None, cosmo, conf)
        torch.cuda.synchronize(dev)
        return dens.cpu().numpy(), out.disp.cpu().numpy(), out.vel.cpu().numpy()


def test_torch_allocator_context_matches_hipmalloc(dev):
    """Same bits either way; the hooked context's memory is torch's, and
    returns to torch's pool when the scratch is trimmed and the context
    closed."""
    import pmwd_amd as pa
    shape = (32, 32, 32)
    conf = pa.Configuration(2.0, shape, mesh_shape=1, a_start=1 / 8, a_nbody_maxstep=1 / 8)
    pmid, disp = opm.gen_grid(shape, conf.mesh_shape, conf.cell_size)
    rng = np.random.default_rng(4)
    disp = (disp + 1.5 * rng.standard_normal(disp.shape)).astype(np.float32)
    vel = (0.1 * rng.standard_normal(disp.shape)).astype(np.float32)

    ctx_h, took_h = _make(conf, True, dev)
    ctx_p, took_p = _make(conf, False, dev)
    try:
        assert took_p == 0
        assert took_h >= ctx_h.nbytes > 0
        got_h = _run(conf, ctx_h, pmid, disp, vel, dev)
        got_p = _run(conf, ctx_p, pmid, disp, vel, dev)
        for a, b in zip(got_h, got_p):
            np.testing.assert_array_equal(a, b)
        # particle scratch went through the hook too, and goes back
        torch.cuda.synchronize(dev)
        held = torch.cuda.memory_allocated(dev)
        grown = ctx_h.nbytes
        ctx_h.trim()
        assert ctx_h.nbytes < grown
        assert torch.cuda.memory_allocated(dev) <= held - (grown - ctx_h.nbytes)
    finally:
        ctx_h.close()
        ctx_p.close()


def test_adjoint_falls_back_when_pipeline_allocation_fails(dev, monkeypatch):
    """The reverse pipeline's allocation fails (an allocator that refuses
    one request, armed just before pmx_nbody_steps_adj): the library gives
    back what it took and returns PMX_ENOMEM before touching any array, and
    the adjoint driver runs the per-step path instead, with the same
    arrays as the pipeline."""
    import importlib
    import pmwd_amd as pa
    from pmwd_amd import _lib
    nbm = importlib.import_module('pmwd_amd.nbody')
    L = _lib.lib()
    shape = (16, 16, 16)
    conf = pa.Configuration(4.0, shape, mesh_shape=2, a_start=1 / 8, a_nbody_maxstep=1 / 16)
    cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
    pmid, disp = opm.gen_grid(shape, conf.mesh_shape, conf.cell_size)
    rng = np.random.default_rng(41)
    disp = (disp + rng.standard_normal(disp.shape)).astype(np.float32)
    vel = (0.1 * rng.standard_normal(disp.shape)).astype(np.float32)
    dc0 = rng.standard_normal(disp.shape).astype(np.float32)
    vc0 = rng.standard_normal(disp.shape).astype(np.float32)
    P = lambda a: torch.as_tensor(a, device=dev).clone()

    state = {'armed': False, 'refused': 0}

    @_lib._ALLOC_T
    def refusing_alloc(nbytes, device, stream, user):
        if state['armed']:
            state['armed'] = False
            state['refused'] += 1
            return None
        try:
            return torch.cuda.caching_allocator_alloc(int(nbytes), int(device), int(stream or 0))
        except Exception:
            return None

    def run(ctx, pipeline, arm):
        real = L.pmx_nbody_steps_adj
        calls = []

        def armed_steps_adj(*a):
            state['armed'] = arm
            rc = real(*a)
            calls.append(rc)
            state['armed'] = False
            return rc

        monkeypatch.setattr(L, 'pmx_nbody_steps_adj', armed_steps_adj)
        try:
            with _lib.context_scope(lambda key, make: ctx):
                d, v = P(disp), P(vel)
                nbm._run_forward(P(pmid), d, v, cosmo, conf)
                x, vv, dc, vc = d, v, P(dc0), P(vc0)
                acc, ac, om = nbm._run_adjoint(P(pmid), x, vv, dc, vc, cosmo, conf,
                                               pipeline=pipeline)
                torch.cuda.synchronize(dev)
                return [t.cpu().numpy() for t in (x, vv, dc, vc, acc, ac)], om, calls
        finally:
            monkeypatch.setattr(L, 'pmx_nbody_steps_adj', real)

    _lib.check(L.pmx_set_allocator(refusing_alloc, _lib._torch_free, None))
    try:
        ctx_f = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
        ctx_p = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
    finally:
        if _hooked_by_default():
            _lib.check(L.pmx_set_allocator(_lib._torch_alloc, _lib._torch_free, None))
        else:
            _lib.check(L.pmx_set_allocator(_lib._ALLOC_T(), _lib._FREE_T(), None))
    try:
        got, om_got, calls = run(ctx_f, None, True)
        assert state['refused'] == 1 and calls == [_lib.ENOMEM]
        want, om_want, calls_p = run(ctx_p, True, False)
        assert calls_p and all(rc == 0 for rc in calls_p)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)
        assert abs(om_got - om_want) <= 1e-10 * abs(om_want)
    finally:
        ctx_f.close()
        ctx_p.close()


def test_process_exits_cleanly_with_live_contexts(dev):
    """A process that ends with contexts alive on torch's pool -- one in the
    library's cache, one held by the caller -- exits with status 0: the
    cached one is destroyed at exit while the allocator callbacks live, the
    other is left to the process end."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = '\n'.join([
        'import numpy as np, torch',
        'import pmwd_amd as pa',
        'from pmwd_amd import _lib',
        'conf = pa.Configuration(1.0, (16, 16, 16), mesh_shape=2)',
        'pmid = torch.zeros((16 ** 3, 3), dtype=torch.int16, device="cuda:0")',
        'pmid[:, 0] = torch.arange(16 ** 3, device="cuda:0") % 16 * 2',
        'disp = torch.rand((16 ** 3, 3), device="cuda:0")',
        'dens = pa.scatter(pa.Particles(conf, pmid, disp), conf)',
        'held = _lib.Context(conf.mesh_shape, conf.cell_size, torch.device("cuda:0"))',
        'torch.cuda.synchronize()',
        'print("ok", float(dens.sum()), held.nbytes > 0, flush=True)',
    ])
    r = subprocess.run([sys.executable, '-c', code], cwd=root, capture_output=True, text=True,
                       timeout=150)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.startswith('ok')


@pytest.mark.parametrize('entry', ['pmx_nbody_steps', 'pmx_nbody_step', 'pmx_nbody_steps_adj',
                                   'pmx_nbody_step_adj'])
def test_enomem_leaves_arrays_unchanged(dev, entry):
    """On a context whose particle scratch was given back (PMX_TRIM_ALL),
    refuse the k-th device allocation of one call, for every k in turn: each
    refused call returns PMX_ENOMEM with every caller array as it was, and
    the first call that needs no refusal gives the same arrays as a fresh
    context."""
    import ctypes
    import pmwd_amd as pa
    from pmwd_amd import _lib
    L = _lib.lib()
    shape = (16, 16, 16)
    conf = pa.Configuration(4.0, shape, mesh_shape=2)
    pmid, disp = opm.gen_grid(shape, conf.mesh_shape, conf.cell_size)
    rng = np.random.default_rng(43)
    arrays0 = [(disp + rng.standard_normal(disp.shape)).astype(np.float32)]
    arrays0 += [rng.standard_normal(disp.shape).astype(np.float32) for _ in range(5)]
    val = float(np.float32(conf.mesh_size / conf.ptcl_num))
    coef, n = 0.45, len(pmid)
    ka = np.array([-0.01, -0.02], np.float32)
    dd = np.array([-0.5, -0.4], np.float32)
    kb = np.array([-0.015, -0.01], np.float32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    pm = torch.as_tensor(pmid, device=dev)

    state = {'countdown': -1, 'refused': 0}

    @_lib._ALLOC_T
    def counting_alloc(nbytes, device, stream, user):
        if state['countdown'] == 0:
            state['countdown'] = -1
            state['refused'] += 1
            return None
        if state['countdown'] > 0:
            state['countdown'] -= 1
        try:
            return torch.cuda.caching_allocator_alloc(int(nbytes), int(device), int(stream or 0))
        except Exception:
            return None

    def call(ctx, arrs):
        x, v, a, dc, vc, ac = (_lib.ptr(t) for t in arrs)
        h = ctx.bind()
        if entry == 'pmx_nbody_steps':
            return L.pmx_nbody_steps(h, _lib.ptr(pm), x, v, a, n, val, coef, 2, P(ka), P(dd), P(kb))
        if entry == 'pmx_nbody_step':
            return L.pmx_nbody_step(h, _lib.ptr(pm), x, v, a, n, val, coef, 0.01, 0.5, 0.02)
        if entry == 'pmx_nbody_steps_adj':
            dots = (ctypes.c_double * 4)()
            return L.pmx_nbody_steps_adj(h, _lib.ptr(pm), x, v, a, dc, vc, ac, n, val, coef,
                                         -0.01, 2, P(ka), P(dd), P(kb), None, dots)
        dots = (ctypes.c_double * 2)()
        return L.pmx_nbody_step_adj(h, _lib.ptr(pm), x, v, a, dc, vc, ac, n, val, coef, -0.01,
                                    -0.02, -0.5, None, dots)

    _lib.check(L.pmx_set_allocator(counting_alloc, _lib._torch_free, None))
    try:
        ctx = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
    finally:
        if _hooked_by_default():
            _lib.check(L.pmx_set_allocator(_lib._torch_alloc, _lib._torch_free, None))
        else:
            _lib.check(L.pmx_set_allocator(_lib._ALLOC_T(), _lib._FREE_T(), None))
    fresh = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
    try:
        # a first full call grows the context (the adjoint's halo shells too)
        warm = [torch.as_tensor(a, device=dev).clone() for a in arrays0]
        _lib.check(call(ctx, warm))
        want = [torch.as_tensor(a, device=dev).clone() for a in arrays0]
        _lib.check(call(fresh, want))
        want = [t.cpu().numpy() for t in want]
        for k in range(64):
            ctx.trim()
            torch.cuda.synchronize(dev)
            arrs = [torch.as_tensor(a, device=dev).clone() for a in arrays0]
            state['countdown'], before = k, state['refused']
            rc = call(ctx, arrs)
            torch.cuda.synchronize(dev)
            got = [t.cpu().numpy() for t in arrs]
            if state['refused'] == before:
                state['countdown'] = -1
                assert rc == 0, (k, _lib.lib().pmx_last_error())
                for g, w in zip(got, want):
                    np.testing.assert_array_equal(g, w)
                break
            assert rc == _lib.ENOMEM, (k, rc)
            for g, a0 in zip(got, arrays0):
                np.testing.assert_array_equal(g, a0, err_msg=f'{entry}: allocation {k} refused')
        else:
            pytest.fail('more than 64 allocations in one call')
        assert k > 0
    finally:
        state['countdown'] = -1
        ctx.close()
        fresh.close()
