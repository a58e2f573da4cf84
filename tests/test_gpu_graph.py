"""Device factor tables, captured graphs and packs.

pmx_nbody_steps_dev reads every factor from a device table and must give the
bits of pmx_nbody_steps with the same factors on the host.  A captured graph
of the forward model's N-body part (initial force, K KDK steps, final CIC
density) must reproduce the uncaptured calls bit for bit on every replay --
with other device work and host synchronisations between replays, and with
the factor table rewritten (another cosmology) between them -- and a pack of
several contexts' graphs must give each simulation's own result."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import pm as opm

pytestmark = pytest.mark.gpu


def particles(shape, ratio, cell, seed, std=0.7):
    pmid, disp = opm.gen_grid(shape, tuple(ratio * s for s in shape), cell)
    rng = np.random.default_rng(seed)
    disp = (disp + std * cell * rng.standard_normal(disp.shape)).astype(np.float32)
    vel = (0.05 * rng.standard_normal(disp.shape)).astype(np.float32)
    return pmid, disp, vel


def factors(k, seed):
    rng = np.random.default_rng(seed)
    return (0.2 + 0.3 * rng.random((k, 3))).astype(np.float32)


def table(coef, f, dev):
    return torch.as_tensor(np.concatenate([[coef], f.reshape(-1)]).astype(np.float32),
                           device=dev)


class Sim:
    """A context on its own stream with fixed particle and density buffers."""

    def __init__(self, dev, mesh, cell, pmid, disp, vel):
        from pmwd_amd import _lib
        self.L, self._lib = _lib.lib(), _lib
        self.stream = torch.cuda.Stream(dev)
        with torch.cuda.stream(self.stream):
            self.ctx = _lib.Context(mesh, cell, dev)
            self.h = self.ctx.bind()
        P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
        self.pm, self.x0, self.v0 = P(pmid), P(disp), P(vel)
        self.x, self.v = self.x0.clone(), self.v0.clone()
        self.acc = torch.zeros_like(self.x)
        self.dens = torch.zeros(mesh, dtype=torch.float32, device=dev)
        self.n = len(pmid)
        torch.cuda.synchronize()

    def reset(self):
        with torch.cuda.stream(self.stream):
            self.x.copy_(self.x0)
            self.v.copy_(self.v0)

    def eager_dev(self, fac, val):
        """The captured sequence, uncaptured."""
        lib, p = self._lib, self._lib.ptr
        # the initial force takes 1.5 Omega_m from the table like the steps
        coef = float(fac[0].item())
        lib.check(self.L.pmx_gravity(self.h, p(self.pm), p(self.x), self.n, val, coef, p(self.acc)))
        lib.check(self.L.pmx_nbody_steps_dev(self.h, p(self.pm), p(self.x), p(self.v), p(self.acc),
                                             self.n, val, (fac.numel() - 1) // 3, p(fac)))
        lib.check(self.L.pmx_scatter_mass(self.h, p(self.pm), p(self.x), self.n, val, p(self.dens),
                                          0))

    def capture(self, fac, val):
        lib, p = self._lib, self._lib.ptr
        lib.check(self.L.pmx_graph_capture(self.h, p(self.pm), p(self.x), p(self.v), p(self.acc),
                                           self.n, val, (fac.numel() - 1) // 3, p(fac),
                                           p(self.dens)))

    def result(self):
        self.stream.synchronize()
        return [t.cpu().numpy() for t in (self.x, self.v, self.acc, self.dens)]


def eager_host(dev, mesh, cell, pmid, disp, vel, coef, f, val):
    """Reference: pmx_gravity, pmx_nbody_steps with host factors, pmx_scatter_mass."""
    from pmwd_amd import _lib
    L, p = _lib.lib(), _lib.ptr
    ctx = _lib.Context(mesh, cell, dev)
    h = ctx.bind()
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev).clone()
    pm, x, v = P(pmid), P(disp), P(vel)
    acc = torch.zeros_like(x)
    dens = torch.zeros(mesh, dtype=torch.float32, device=dev)
    k1, d, k2 = (np.ascontiguousarray(f[:, m]) for m in range(3))
    H = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    _lib.check(L.pmx_gravity(h, p(pm), p(x), len(pmid), val, coef, p(acc)))
    _lib.check(L.pmx_nbody_steps(h, p(pm), p(x), p(v), p(acc), len(pmid), val, coef, len(f),
                                 H(k1), H(d), H(k2)))
    _lib.check(L.pmx_scatter_mass(h, p(pm), p(x), len(pmid), val, p(dens), 0))
    torch.cuda.synchronize()
    out = [t.cpu().numpy() for t in (x, v, acc, dens)]
    ctx.close()
    return out


def assert_same(a, b, what):
    for name, x, y in zip(('disp', 'vel', 'acc', 'dens'), a, b):
        np.testing.assert_array_equal(x, y, err_msg=f'{what}: {name}')


CASES = [((16, 16, 16), 2, 6), ((20, 24, 12), 1, 4)]


@pytest.mark.parametrize('shape,ratio,k', CASES)
def test_device_factor_steps_equal_host_factor_steps(dev, shape, ratio, k):
    cell = 1.0
    mesh = tuple(ratio * s for s in shape)
    pmid, disp, vel = particles(shape, ratio, cell, 1)
    f, coef, val = factors(k, 2), 0.45, 1.0 / ratio**3
    print(f'case shape={shape} ratio={ratio} k={k}', flush=True)
    want = eager_host(dev, mesh, cell, pmid, disp, vel, coef, f, val)
    sim = Sim(dev, mesh, cell, pmid, disp, vel)
    fac = table(coef, f, dev)
    with torch.cuda.stream(sim.stream):
        sim.eager_dev(fac, val)
    assert_same(sim.result(), want, 'device factors')


@pytest.mark.parametrize('shape,ratio,k', CASES)
def test_graph_replays_equal_eager(dev, shape, ratio, k):
    cell = 1.0
    mesh = tuple(ratio * s for s in shape)
    pmid, disp, vel = particles(shape, ratio, cell, 3)
    val = 1.0 / ratio**3
    fa, fb = factors(k, 4), factors(k, 5)
    want_a = eager_host(dev, mesh, cell, pmid, disp, vel, 0.45, fa, val)
    want_b = eager_host(dev, mesh, cell, pmid, disp, vel, 0.39, fb, val)
    sim = Sim(dev, mesh, cell, pmid, disp, vel)
    fac = table(0.45, fa, dev)
    print(f'case shape={shape} ratio={ratio} k={k}: warm-up, capture', flush=True)
    with torch.cuda.stream(sim.stream):
        sim.eager_dev(fac, val)      # sizes the context's scratch
        sim.reset()
        sim.capture(fac, val)
    ta, tb = fac.clone(), table(0.39, fb, dev)
    for rep, (src, want) in enumerate([(ta, want_a), (tb, want_b), (ta, want_a),
                                       (tb, want_b)]):
        # eager device work and host synchronisations between the replays
        junk = torch.randn(1 << 16, device=dev).sum()
        torch.cuda.synchronize()
        del junk
        with torch.cuda.stream(sim.stream):
            if rep:
                fac.copy_(src)
            sim.reset()
            sim._lib.check(sim.L.pmx_graph_launch(sim.h))
        assert_same(sim.result(), want, f'replay {rep}')


def test_pack_runs_each_simulation(dev):
    from pmwd_amd import _lib
    shape, ratio, k = (16, 16, 16), 2, 5
    cell = 1.0
    mesh = tuple(ratio * s for s in shape)
    val = 1.0 / ratio**3
    sims, wants, facs = [], [], []
    for i in range(3):
        pmid, disp, vel = particles(shape, ratio, cell, 10 + i)
        f, coef = factors(k, 20 + i), 0.4 + 0.02 * i
        wants.append(eager_host(dev, mesh, cell, pmid, disp, vel, coef, f, val))
        s = Sim(dev, mesh, cell, pmid, disp, vel)
        fac = table(coef, f, dev)
        with torch.cuda.stream(s.stream):
            s.eager_dev(fac, val)
            s.reset()
            s.capture(fac, val)
            s.reset()
        s.stream.synchronize()
        sims.append(s)
        facs.append(fac)
    L = _lib.lib()
    for k_pack in (3, 2):
        arr = (ctypes.c_void_p * k_pack)(*[s.h.value for s in sims[:k_pack]])
        pack = ctypes.c_void_p()
        _lib.check(L.pmx_pack_create(arr, k_pack, ctypes.byref(pack)))
        try:
            for s in sims:
                s.reset()
            torch.cuda.synchronize()
            stream = torch.cuda.Stream(dev)
            _lib.check(L.pmx_pack_launch(pack, ctypes.c_void_p(stream.cuda_stream)))
            stream.synchronize()
            for i in range(k_pack):
                assert_same(sims[i].result(), wants[i], f'pack of {k_pack}: simulation {i}')
        finally:
            L.pmx_pack_destroy(pack)


def test_trim_drops_the_graph(dev):
    from pmwd_amd import _lib
    shape, ratio, k = (16, 16, 16), 1, 2
    pmid, disp, vel = particles(shape, ratio, 1.0, 7)
    sim = Sim(dev, shape, 1.0, pmid, disp, vel)
    fac = table(0.45, factors(k, 8), dev)
    with torch.cuda.stream(sim.stream):
        sim.eager_dev(fac, 1.0)
        sim.reset()
        sim.capture(fac, 1.0)
    sim.stream.synchronize()
    sim.ctx.trim()
    with pytest.raises(_lib.PmxError, match='no captured graph'):
        _lib.check(sim.L.pmx_graph_launch(sim.h))
