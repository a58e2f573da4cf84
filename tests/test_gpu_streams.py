"""Stream rebinding of a library context: a new stream starts after all the
work the context queued, the previous stream's handle is never used again,
and graph capture refuses allocations and host waits with an error."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import pm as opm

pytestmark = pytest.mark.gpu


def _setup(dev):
    import pmwd_amd as pa
    from pmwd_amd import _lib
    shape = (16, 16, 16)
    conf = pa.Configuration(1.0, shape, mesh_shape=2)
    pmid, disp = opm.gen_grid(shape, conf.mesh_shape, conf.cell_size)
    disp = (disp + np.random.default_rng(3).standard_normal(disp.shape)).astype(np.float32)
    pm = torch.as_tensor(pmid, device=dev)
    x = torch.as_tensor(disp, device=dev)
    val = float(np.float32(conf.mesh_size / conf.ptcl_num))
    return _lib, conf, pm, x, val


def _scatter(_lib, ctx, pm, x, val, mesh):
    _lib.check(_lib.lib().pmx_scatter_mass(ctx.h, _lib.ptr(pm), _lib.ptr(x), len(pm), val,
                                           _lib.ptr(mesh), 0))


def test_new_stream_waits_for_work_queued_on_the_old(dev):
    _lib, conf, pm, x, val = _setup(dev)
    ctx = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
    want = torch.zeros(conf.mesh_shape, device=dev)
    with torch.cuda.stream(torch.cuda.Stream(dev)):
        ctx.bind()
        _scatter(_lib, ctx, pm, x, val, want)
        torch.cuda.current_stream(dev).synchronize()
    A, B = torch.cuda.Stream(dev), torch.cuda.Stream(dev)
    got = torch.zeros(conf.mesh_shape, device=dev)
    with torch.cuda.stream(A):
        ctx.bind()
        torch.cuda._sleep(1 << 28)           # about 0.1 s of spinning on A
        _scatter(_lib, ctx, pm, x, val, got)
    with torch.cuda.stream(B):
        ctx.bind()
        ev = torch.cuda.Event()
        ev.record(B)
    assert not ev.query(), 'the rebound stream did not wait for the old one'
    ev.synchronize()
    assert A.query(), 'B ran past the work queued on A'
    torch.cuda.synchronize(dev)
    ctx.trim()
    with torch.cuda.stream(B):
        again = torch.zeros(conf.mesh_shape, device=dev)
        _scatter(_lib, ctx, pm, x, val, again)
    torch.cuda.synchronize(dev)
    assert torch.equal(got, want) and torch.equal(again, want)
    ctx.close()


def test_rebinding_after_the_old_stream_is_destroyed(dev):
    """A caller that destroys its stream before rebinding (an external
    stream of its own) is served: the old handle is not touched."""
    _lib, conf, pm, x, val = _setup(dev)
    # the HIP runtime already in the process (torch's), never a second copy
    with open('/proc/self/maps') as f:
        paths = {ln.split()[-1] for ln in f if 'libamdhip64.so' in ln}
    assert len(paths) == 1, paths
    hip = ctypes.CDLL(paths.pop())
    hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    raw = ctypes.c_void_p()
    assert hip.hipStreamCreateWithFlags(ctypes.byref(raw), 0) == 0
    ext = torch.cuda.ExternalStream(raw.value, device=dev)
    ctx = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
    m1 = torch.zeros(conf.mesh_shape, device=dev)
    with torch.cuda.stream(ext):
        ctx.bind()
        _scatter(_lib, ctx, pm, x, val, m1)
    assert hip.hipStreamSynchronize(raw) == 0
    del ext
    assert hip.hipStreamDestroy(raw) == 0
    m2 = torch.zeros(conf.mesh_shape, device=dev)
    with torch.cuda.stream(torch.cuda.Stream(dev)):
        ctx.bind()
        _scatter(_lib, ctx, pm, x, val, m2)
    torch.cuda.synchronize(dev)
    assert torch.equal(m1, m2)
    ctx.close()


def test_capture_refuses_allocation_and_host_waits(dev):
    """Under graph capture a call that would allocate, or wait on the host,
    returns PMX_EINVAL with a message instead of synchronising the
    capturing stream."""
    _lib, conf, pm, x, val = _setup(dev)
    L = _lib.lib()
    ctx = _lib.Context(conf.mesh_shape, conf.cell_size, dev)
    mesh = torch.zeros(conf.mesh_shape, device=dev)
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream(dev)
    rcs = {}
    with torch.cuda.stream(s):
        g.capture_begin()
        try:
            h = ctx.bind()
            # a fresh context has no particle scratch yet: binning would allocate
            rcs['scatter'] = L.pmx_scatter_mass(h, _lib.ptr(pm), _lib.ptr(x), len(pm), val,
                                                _lib.ptr(mesh), 0)
            rcs['msg'] = L.pmx_last_error().decode()
            d = ctypes.c_double()
            rcs['dot'] = L.pmx_dot(h, _lib.ptr(x), _lib.ptr(x), x.numel(), ctypes.byref(d))
        finally:
            g.capture_end()
    assert rcs['scatter'] == 1 and 'graph capture' in rcs['msg']
    assert rcs['dot'] == 1
    ctx.close()
