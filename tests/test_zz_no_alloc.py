"""The entry point that allocates nothing, pmx_scatter_mass_noalloc, against
pmx_scatter_mass on a context: bitwise the same density on two calls in a
row (the second smaller, on another mesh), overwrite and accumulate, and a
clear refusal of sizes beyond its static scratch."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import pm as opm

pytestmark = pytest.mark.gpu

c_p, c_ll, c_int, c_d, c_f = (ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_double,
                              ctypes.c_float)


def _noalloc():
    from pmwd_amd import _lib
    return _lib.lib().pmx_scatter_mass_noalloc


def _case(shape, mesh, cell, std, seed):
    pmid, disp = opm.gen_grid(shape, mesh, cell)
    rng = np.random.default_rng(seed)
    disp = (disp + std * rng.standard_normal(disp.shape)).astype(np.float32)
    # a clump: a tenth of the particles within a cell of the box centre
    k = len(disp) // 10
    centre = np.asarray(mesh) * cell / 2
    pos = centre + 0.3 * cell * rng.standard_normal((k, 3))
    pmid[:k] = np.rint(pos / cell).astype(np.int16) % np.asarray(mesh, np.int16)
    disp[:k] = (pos - np.rint(pos / cell) * cell).astype(np.float32)
    return pmid, disp


def test_noalloc_equals_scatter_mass_on_two_calls(dev):
    from pmwd_amd import _lib
    f = _noalloc()
    L = _lib.lib()
    cases = [((32, 32, 32), (64, 64, 64), 1.953125, 3.0, 1),
             ((12, 10, 8), (23, 20, 17), 1.0, 2.0, 2)]         # smaller, odd mesh
    for shape, mesh, cell, std, seed in cases:
        pmid, disp = _case(shape, mesh, cell, std, seed)
        n = len(pmid)
        pm = torch.as_tensor(pmid, device=dev)
        x = torch.as_tensor(disp, device=dev)
        val = float(np.float32(np.prod(mesh) / n))
        ctx = _lib.Context(mesh, cell, dev)
        base = torch.as_tensor(np.random.default_rng(seed).standard_normal(mesh).astype(np.float32),
                               device=dev)
        for acc in (0, 1):
            want = base.clone()
            _lib.check(L.pmx_scatter_mass(ctx.bind(), _lib.ptr(pm), _lib.ptr(x), n, val,
                                          _lib.ptr(want), acc))
            got = base.clone()
            stream = torch.cuda.current_stream(dev).cuda_stream
            rc = f(_lib.ptr(pm), _lib.ptr(x), n, *mesh, cell, val, _lib.ptr(got), acc, stream)
            assert rc == 0, _lib.lib().pmx_last_error()
            torch.cuda.synchronize(dev)
            np.testing.assert_array_equal(got.cpu().numpy(), want.cpu().numpy())
        ref = opm.scatter(pmid, disp, cell, mesh)
        assert np.linalg.norm(got.cpu().numpy() - base.cpu().numpy() - ref) < 1e-5 * np.linalg.norm(ref)
        ctx.close()


def test_noalloc_refuses_sizes_beyond_its_scratch(dev):
    from pmwd_amd import _lib
    f = _noalloc()
    n = 4 << 20
    mesh = torch.zeros((256, 256, 256), device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    # no particle array is read: the size check comes first
    dummy = torch.zeros((1, 3), device=dev)
    rc = f(_lib.ptr(dummy), _lib.ptr(dummy), n, 256, 256, 256, 1.0, 1.0, _lib.ptr(mesh), 0, stream)
    assert rc == 1
    assert 'more than' in _lib.lib().pmx_last_error().decode()
    assert float(mesh.abs().sum()) == 0.0
