"""Continuous-batching simulation server on one GPU.

A request is one pmwd forward model run from its own seed and cosmology:
white noise -> ``linear_modes`` -> ``lpt`` -> ``nbody`` -> ``scatter``
(modes.py, lpt.py, nbody.py, scatter.py), optionally followed by the
gradient of the density loss mean((rho - 1)^2) with respect to the white
modes, Omega_m and sigma8 (A_s following sigma8 as in
Cosmology.from_sigma8), taken through the reverse-time adjoint.

Scheduling:

* requests queue on arrival; the admission loop starts one whenever a slot
  is free and its estimated device bytes fit the HBM budget next to the
  requests already running (first come first served, a large request at the
  head waits rather than being overtaken);
* every admitted request runs on a worker thread with its own HIP stream
  (default priority) and its own library contexts, so the kernels of
  concurrent requests overlap on the device while the host work of one
  (tables, random numbers, launches) proceeds beside the device work of the
  others; requests arriving while others run join as soon as memory frees;
* library contexts (FFT twiddles, k tables, mesh buffers, binning scratch)
  stay resident in a pool keyed by mesh shape and are handed to the next
  request of that shape; idle ones are released only when the budget needs
  their memory;
* forward-only requests hand their initial conditions to the PackEngine of
  their configuration: slots with resident buffers and a captured HIP graph
  of nbody + scatter whose step factors live in a device table, so one graph
  serves every cosmology; each request replays its slot's graph on the
  slot's own stream, so the slots' runs proceed side by side (or, with
  pack_mode='pack', the slots loaded at the same time run as one pack: one
  graph launch holding each slot's graph as an independent branch).

Every kernel on the path is deterministic and each request owns its
buffers, so a request's density and gradients do not depend on what runs
beside it; only the float64 sums behind the Omega_m gradient may change in
their last bits with the order in which tile partial sums are added.
"""
import ctypes
import math
import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib


@dataclass(frozen=True)
class SimRequest:
    """One simulation: particles per side, mesh/particle ratio, box size in
    Mpc/h, cosmology and whether the gradient is wanted."""
    seed: int
    ptcl_grid: int
    mesh_ratio: int
    box: float
    Omega_m: float
    sigma8: float
    grad: bool = False
    a_start: float = 1 / 32
    a_nbody_maxstep: float = 1 / 32
    n_s: float = 0.96
    Omega_b: float = 0.05
    h: float = 0.7

    def conf(self):
        from .configuration import Configuration
        g = int(self.ptcl_grid)
        return Configuration(self.box / g, (g, g, g), mesh_shape=self.mesh_ratio,
                             a_start=self.a_start, a_nbody_maxstep=self.a_nbody_maxstep)


@dataclass
class SimResult:
    request: SimRequest
    dens: np.ndarray                      # final CIC density (mesh shape), float32
    loss: Optional[float] = None          # mean((dens - 1)^2) of a gradient request
    grad_modes: Optional[np.ndarray] = None
    grad_Omega_m: Optional[float] = None
    grad_sigma8: Optional[float] = None
    t_arrive: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    n_steps: int = 0

    @property
    def latency(self):
        return self.t_done - self.t_arrive


def context_bytes(mesh_shape, n_ptcl, adjoint):
    """Device bytes a library context of this mesh holds once used for n_ptcl
    particles: three complex mesh buffers, twiddles, k tables, tile arrays
    and halo shells, and per record slot (about 1.2 n_ptcl slots) the record,
    index, tile and rank words, four value channels and the multi-step
    pipeline's own particle copy.  The reverse-time pipeline adds three
    cotangent meshes, a 56-byte particle copy per slot and two more halo
    channels."""
    n0, n1, n2 = mesh_shape
    n2p = n2 // 2 + 1
    n2p = (n2p + 15) // 16 * 16 if n2p > 16 else n2p
    mesh = (6 if adjoint else 3) * n0 * n1 * n2p * 8
    tiles = math.ceil(n0 / 16) * math.ceil(n1 / 16) * math.ceil(n2 / 16)
    tile_bytes = tiles * (40 + 4 * 817 * (3 if adjoint else 1))
    cap = n_ptcl + n_ptcl // 16 + 1024
    slots = cap + cap // 8 + 33 * (tiles + 1)
    per_slot = 16 + 4 + 8 + 16 + 32 + (56 if adjoint else 0)
    return int(mesh + tile_bytes + slots * per_slot + 64 * 2**20)


def request_bytes(req):
    """Estimated device bytes of one request apart from its library contexts:
    particle state and the tensors the initial conditions, the density and
    (for gradients) the autograd graph of linear_modes and lpt hold."""
    n = req.ptcl_grid**3
    m = n * req.mesh_ratio**3
    state = n * (6 + 4 * 12)                  # pmid, disp, vel, acc, one copy
    ic = n * 8 * 16                           # complex grids of modes and lpt
    dens = m * 4 * 3
    if req.grad:
        state += n * 12 * 6                   # cotangents, saved start state
        ic *= 2                               # tensors kept for the backward
    return int(state + ic + dens)


def _contexts_needed(req):
    conf = req.conf()
    n = conf.ptcl_num
    # the particle-grid FFTs of the initial conditions and the mesh of the
    # force/scatter (one context when they coincide in shape and cell)
    out = {(tuple(conf.ptcl_grid_shape), 1.0): context_bytes(conf.ptcl_grid_shape, 0, False),
           (tuple(conf.mesh_shape), float(conf.cell_size)):
               context_bytes(conf.mesh_shape, n, req.grad)}
    return out


def run_request(req, device, engine=None):
    """Runs one request with the public API on the current stream; returns
    the SimResult fields (device work finished, results on the host).  With
    ``engine`` (a PackEngine of the request's configuration) a forward-only
    request's nbody and scatter run as a replay of the engine's captured
    graphs."""
    from .boltzmann import boltzmann
    from .cosmology import Cosmology
    from .lpt import lpt
    from .modes import linear_modes, white_noise
    from .nbody import nbody
    from .scatter import scatter
    conf = req.conf()
    # the variance table is not on the model's path (from_sigma8 normalises
    # with its own two nodes)
    cosmo = boltzmann(Cosmology.from_sigma8(conf, req.sigma8, n_s=req.n_s, Omega_m=req.Omega_m,
                                            Omega_b=req.Omega_b, h=req.h), conf, varlin=False)
    white = white_noise(req.seed, conf, real=True, device=device)
    out = {'n_steps': len(conf.a_nbody) - 1}
    if not req.grad:
        with torch.no_grad():
            modes = linear_modes(white, cosmo, conf)
            ptcl, _ = lpt(modes, cosmo, conf)
            if engine is not None:
                out['dens'] = engine.run(ptcl.disp, ptcl.vel, nbody_factor_table(cosmo, conf))
                return out
            ptcl, _ = nbody(ptcl, None, cosmo, conf)
            dens = scatter(ptcl, conf)
        out['dens'] = dens.cpu().numpy()
        return out
    white.requires_grad_(True)
    theta = {'Omega_m': torch.tensor(cosmo.Omega_m, dtype=torch.float64, requires_grad=True),
             'sigma8': torch.tensor(cosmo.sigma8, dtype=torch.float64, requires_grad=True)}
    modes = linear_modes(white, cosmo, conf, theta=theta)
    ptcl, _ = lpt(modes, cosmo, conf, theta=theta)
    ptcl, _ = nbody(ptcl, None, cosmo, conf, omega_m=theta['Omega_m'])
    dens = scatter(ptcl, conf)
    loss = ((dens.double() - 1)**2).mean()
    g_w, g_om, g_s8 = torch.autograd.grad(loss, (white, theta['Omega_m'], theta['sigma8']))
    out.update(dens=dens.detach().cpu().numpy(), loss=float(loss.detach()), grad_modes=g_w.cpu().numpy(),
               grad_Omega_m=float(g_om), grad_sigma8=float(g_s8))
    return out


def nbody_factor_table(cosmo, conf):
    """float32 {1.5 Omega_m, k1_0, d_0, k2_0, k1_1, ...} of nbody's KDK loop
    (nbody.py:205-223), the device table of pmx_nbody_steps_dev."""
    from .nbody import kdk_factor_table
    from .gravity import omega_coef
    a = conf.a_nbody
    f = kdk_factor_table(a[:-1], a[1:], cosmo, conf)
    return np.concatenate([[omega_coef(cosmo, conf)], f.reshape(-1)]).astype(np.float32)


class _Slot:
    """One simulation slot of a pack engine: a library context on a stream
    of its own (not one of torch's pooled streams, which other threads may
    be queueing work on while this one is captured), fixed particle, force
    and density buffers, a device factor table, and (after its first run)
    the captured graph of nbody plus the final scatter on those buffers."""

    def __init__(self, conf, device):
        from .particles import Particles
        self._ps = _lib.PrivateStream(device)
        self.stream = self._ps.stream
        # pinned host memory and context creation may wait on the device
        with _lib.capture_lock, torch.cuda.device(device), torch.cuda.stream(self.stream):
            self.ctx = _lib.Context(conf.mesh_shape, conf.cell_size, device)
            self.h = self.ctx.bind()
            self.pmid = Particles.gen_grid(conf, device=device).pmid
            n = conf.ptcl_num
            self.disp = torch.empty((n, 3), dtype=torch.float32, device=device)
            self.vel = torch.empty_like(self.disp)
            self.acc = torch.empty_like(self.disp)
            self.dens = torch.empty(tuple(conf.mesh_shape), dtype=torch.float32, device=device)
            self.nsteps = len(conf.a_nbody) - 1
            self.fac = torch.empty(1 + 3 * self.nsteps, dtype=torch.float32, device=device)
            self.fac_host = torch.empty(1 + 3 * self.nsteps, dtype=torch.float32, pin_memory=True)
        self.n = conf.ptcl_num
        self.val = float(np.float32(conf.mesh_size / conf.ptcl_num))
        self.captured = False
        self.loaded = None       # event: state and factors in place
        self.done = threading.Event()
        self.finished = None     # event: the launch that ran this slot has ended
        self.error = None        # the launch's failure, if it failed

    def nbytes(self):
        b = sum(t.numel() * t.element_size() for t in (self.pmid, self.disp, self.vel, self.acc,
                                                       self.dens, self.fac))
        return b + self.ctx.nbytes

    def run_eager(self):
        """The captured sequence, uncaptured (sizes the context's scratch)."""
        L, p = _lib.lib(), _lib.ptr
        coef = float(self.fac_host[0])
        _lib.check(L.pmx_gravity(self.h, p(self.pmid), p(self.disp), self.n, self.val, coef,
                                 p(self.acc)))
        _lib.check(L.pmx_nbody_steps_dev(self.h, p(self.pmid), p(self.disp), p(self.vel),
                                         p(self.acc), self.n, self.val, self.nsteps, p(self.fac)))
        _lib.check(L.pmx_scatter_mass(self.h, p(self.pmid), p(self.disp), self.n, self.val,
                                      p(self.dens), 0))

    def capture(self):
        L, p = _lib.lib(), _lib.ptr
        with _lib.capture_lock:
            _lib.check(L.pmx_graph_capture(self.h, p(self.pmid), p(self.disp), p(self.vel),
                                           p(self.acc), self.n, self.val, self.nsteps,
                                           p(self.fac), p(self.dens)))
        self.captured = True

    def close(self):
        self.ctx.close()
        self._ps.close()


class PackEngine:
    """Forward-only N-body runs of one size class (one Configuration shape)
    as replays of captured graphs.

    Each of ``nslots`` slots keeps its context, buffers and captured graph
    resident.  A request's initial conditions are copied into a free slot
    with its factor table.  A slot's first run is eager and captures its
    graph for the next ones.  Results are bitwise those of ``nbody``
    followed by ``scatter``.  Then, by ``mode``:

    * ``'replay'``: the request replays its slot's graph on the slot's own
      stream at once (``pmx_graph_launch``); the slots' streams run side by
      side.
    * ``'pack'``: the launcher thread runs every slot loaded by then as one
      pack (``pmx_pack_launch``: the slots' graphs as independent branches of
      one graph, one launch for all of them), waiting ``window`` seconds for
      slots still loading.
    """

    def __init__(self, conf, device, nslots=4, window=0.002, mode='replay'):
        if mode not in ('replay', 'pack'):
            raise ValueError(f"mode must be 'replay' or 'pack', not {mode!r}")
        self.conf = conf
        self.device = torch.device(device)
        self.mode = mode
        self.slots = [_Slot(conf, self.device) for _ in range(nslots)]
        self.window = float(window)
        self._cv = threading.Condition()
        self._free = list(range(nslots))
        self._ready = []
        self._loading = 0
        self._packs = {}
        self._closing = False
        self._ps = _lib.PrivateStream(self.device)
        self.stream = self._ps.stream
        self.launches = 0
        self.packed = 0
        self.eager_runs = 0
        self._thread = threading.Thread(target=self._loop, daemon=True)
        self._thread.start()

    def nbytes(self):
        return sum(s.nbytes() for s in self.slots)

    def run(self, disp, vel, fac, src_stream=None):
        """Runs nbody + scatter from (disp, vel) with the factor table fac
        (nbody_factor_table); returns the density on the host (numpy)."""
        src_stream = src_stream or torch.cuda.current_stream(self.device)
        with self._cv:
            while not self._free:
                self._cv.wait()
            i = self._free.pop()
            self._loading += 1
        s = self.slots[i]
        loading = True
        try:
            s.stream.wait_stream(src_stream)
            with torch.cuda.device(self.device), torch.cuda.stream(s.stream):
                s.disp.copy_(disp)
                s.vel.copy_(vel)
                s.fac_host.copy_(torch.from_numpy(np.ascontiguousarray(fac, dtype=np.float32)))
                s.fac.copy_(s.fac_host, non_blocking=True)
                eager = not s.captured
                if eager:
                    s.run_eager()
                    s.stream.synchronize()
                    dens = s.dens.cpu().numpy()
                    s.capture()
                elif self.mode == 'replay':
                    _lib.check(_lib.lib().pmx_graph_launch(s.h))
                    dens = s.dens.cpu().numpy()
                else:
                    s.loaded = torch.cuda.Event()
                    s.loaded.record(s.stream)
            if eager or self.mode == 'replay':
                with self._cv:
                    if eager:
                        self.eager_runs += 1
                    else:
                        self.launches += 1
                        self.packed += 1
                return dens
            s.done.clear()
            with self._cv:
                self._loading -= 1
                loading = False
                self._ready.append(i)
                self._cv.notify_all()
            s.done.wait()
            if s.error is not None:
                raise s.error
            s.finished.synchronize()
            with torch.cuda.device(self.device), torch.cuda.stream(s.stream):
                return s.dens.cpu().numpy()
        finally:
            # the slot goes back in every case, a failed run's included (work
            # queued on its stream orders the next user's copies after it)
            with self._cv:
                if loading:
                    self._loading -= 1
                if i in self._ready:
                    self._ready.remove(i)
                self._free.append(i)
                self._cv.notify_all()

    def _pack(self, idx):
        key = tuple(idx)
        p = self._packs.get(key)
        if p is None:
            arr = (ctypes.c_void_p * len(idx))(*[self.slots[i].h.value for i in idx])
            p = ctypes.c_void_p()
            _lib.check(_lib.lib().pmx_pack_create(arr, len(idx), ctypes.byref(p)))
            self._packs[key] = p
        return p

    def _loop(self):
        while True:
            with self._cv:
                while not self._ready and not self._closing:
                    self._cv.wait()
                if self._closing and not self._ready:
                    return
                # gather the slots still loading, for one window
                if self._loading and len(self._ready) < len(self.slots):
                    self._cv.wait(timeout=self.window)
                idx = sorted(self._ready)
                self._ready.clear()
            try:
                for i in idx:
                    self.stream.wait_event(self.slots[i].loaded)
                _lib.check(_lib.lib().pmx_pack_launch(self._pack(idx),
                                                      ctypes.c_void_p(self.stream.cuda_stream)))
                ev = torch.cuda.Event()
                ev.record(self.stream)
                err = None
            except BaseException as e:  # the waiting requests raise it
                ev, err = None, e
            self.launches += 1
            self.packed += len(idx)
            for i in idx:
                self.slots[i].finished, self.slots[i].error = ev, err
                self.slots[i].done.set()

    def close(self):
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        self._thread.join()
        torch.cuda.synchronize(self.device)
        L = _lib.lib()
        for p in self._packs.values():
            L.pmx_pack_destroy(p)
        self._packs.clear()
        for s in self.slots:
            s.close()
        self._ps.close()

    def stats(self):
        return {'slots': len(self.slots), 'pack_launches': self.launches,
                'packed_runs': self.packed, 'eager_runs': self.eager_runs,
                'mean_pack_size': self.packed / self.launches if self.launches else None}


class Server:
    """Admits SimRequests against an HBM budget and runs up to ``max_slots``
    of them concurrently on one GPU (see the module docstring)."""

    def __init__(self, device='cuda:0', hbm_budget=None, max_slots=8, packs=True,
                 pack_window=0.002, pack_slots=None, pack_mode='replay'):
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device('cuda', torch.cuda.current_device())
        if hbm_budget is None:
            free, _total = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self._cv = threading.Condition()
        self._pending = deque()
        self._active = 0
        self._active_bytes = 0
        self._pool = {}            # key -> list of idle Context
        self._pool_bytes = 0       # bytes of every pooled context, idle or in use
        self._ctx_bytes = {}       # id(Context) -> estimated bytes
        self._closing = False
        self.peak_active = 0
        # forward-only requests: one PackEngine per configuration
        self.packs = bool(packs)
        self.pack_window = float(pack_window)
        self.pack_slots = pack_slots
        self.pack_mode = pack_mode
        self._engines = {}
        self._engine_lock = threading.Lock()
        self._admitter = threading.Thread(target=self._admit_loop, daemon=True)
        self._admitter.start()

    # ------------------------------------------------------------ public --
    def submit(self, req, t_arrive=None):
        fut = Future()
        with self._cv:
            if self._closing:
                raise RuntimeError('server is closed')
            self._pending.append((req, fut, time.perf_counter() if t_arrive is None else t_arrive))
            self._cv.notify_all()
        return fut

    def close(self):
        """Waits for every submitted request, then stops the admission loop."""
        with self._cv:
            self._closing = True
            self._cv.notify_all()
            while self._pending or self._active:
                self._cv.wait()
        self._admitter.join()
        for e in self._engines.values():
            e.close()
        self._engines.clear()

    def pack_stats(self):
        return {f'{k[0]}^3/{k[0] * k[1]}^3': e.stats() for k, e in self._engines.items()}

    # ---------------------------------------------------------- internals --
    @staticmethod
    def _engine_key(req):
        return (req.ptcl_grid, req.mesh_ratio, req.box, req.a_start, req.a_nbody_maxstep)

    def _slots_for(self, req):
        return int(self.pack_slots) if self.pack_slots else self._engine_slots(req)

    @staticmethod
    def _engine_slots(req):
        # as many slots as requests in flight for small simulations (their
        # replays run side by side on the slots' streams), two from 256^3
        return 8 if req.ptcl_grid**3 <= 1 << 21 else 2

    def _engine_bytes(self, req):
        conf = req.conf()
        n, m = conf.ptcl_num, conf.mesh_size
        return self._slots_for(req) * (context_bytes(conf.mesh_shape, n, False) + 42 * n + 4 * m)

    def _engine_for(self, req):
        key = self._engine_key(req)
        with self._engine_lock:
            e = self._engines.get(key)
            if e is None:
                e = PackEngine(req.conf(), self.device, self._slots_for(req),
                               window=self.pack_window, mode=self.pack_mode)
                with self._cv:
                    self._pool_bytes += self._engine_bytes(req)
                self._engines[key] = e
        return e

    def _fits(self, req):
        need = request_bytes(req)
        for key, nb in _contexts_needed(req).items():
            if not self._pool.get(self._key(key)):
                need += nb
        if self.packs and not req.grad and self._engine_key(req) not in self._engines:
            need += self._engine_bytes(req)
        return self._active_bytes + self._pool_bytes + need <= self.hbm_budget, need

    def _key(self, k):
        shape, cell = k
        return (shape, cell, self.device.index)

    def _evict_idle(self, need):
        """Frees idle pooled contexts until `need` more bytes fit (synchronises
        the device, which only happens when the budget is tight)."""
        for key in list(self._pool):
            while self._pool[key] and self._active_bytes + self._pool_bytes + need > self.hbm_budget:
                c = self._pool[key].pop()
                self._pool_bytes -= self._ctx_bytes.pop(id(c), 0)
                c.close()

    def _admit_loop(self):
        while True:
            with self._cv:
                while True:
                    if self._closing and not self._pending:
                        return
                    if self._pending and self._active < self.max_slots:
                        req = self._pending[0][0]
                        ok, need = self._fits(req)
                        if not ok:
                            self._evict_idle(need)
                            ok, need = self._fits(req)
                        if ok or self._active == 0:
                            break
                    self._cv.wait()
                req, fut, t_arr = self._pending.popleft()
                self._active += 1
                self._active_bytes += request_bytes(req)
                self.peak_active = max(self.peak_active, self._active)
            threading.Thread(target=self._work, args=(req, fut, t_arr), daemon=True).start()

    def _provider(self, held):
        def get(key, make):
            if key in held:
                return held[key]
            with self._cv:
                idle = self._pool.setdefault(key, [])
                ctx = idle.pop() if idle else None
            if ctx is None:
                ctx = make()
                shape = key[0]
                with self._cv:
                    nb = int(ctx.nbytes)
                    self._ctx_bytes[id(ctx)] = max(nb, context_bytes(shape, 0, False))
                    self._pool_bytes += self._ctx_bytes[id(ctx)]
            held[key] = ctx
            return ctx
        return get

    def _work(self, req, fut, t_arr):
        t_admit = time.perf_counter()
        held = {}
        try:
            stream = torch.cuda.Stream(self.device)
            with torch.cuda.device(self.device), torch.cuda.stream(stream), \
                    _lib.context_scope(self._provider(held)):
                engine = self._engine_for(req) if self.packs and not req.grad else None
                out = run_request(req, self.device, engine)
                stream.synchronize()
            res = SimResult(request=req, t_arrive=t_arr, t_admit=t_admit,
                            t_done=time.perf_counter(), **out)
            fut.set_result(res)
        except BaseException as e:  # the caller sees the failure on its future
            fut.set_exception(e)
        finally:
            with self._cv:
                for key, ctx in held.items():
                    # contexts grow their binning scratch on first use
                    nb = max(int(ctx.nbytes), self._ctx_bytes.get(id(ctx), 0))
                    self._pool_bytes += nb - self._ctx_bytes.get(id(ctx), 0)
                    self._ctx_bytes[id(ctx)] = nb
                    self._pool.setdefault(key, []).append(ctx)
                self._active -= 1
                self._active_bytes -= request_bytes(req)
                self._cv.notify_all()


def c5_mix(n, rate, seed=0):
    """The serving load: n requests with Poisson arrivals at `rate` per
    second; 40% 64^3 particles on a 128^3 mesh, 40% 128^3 on 256^3, 20% 256^3
    on 512^3 (2 Mpc/h particle spacing); half forward-only, half with
    gradients; Omega_m ~ U[0.25, 0.35], sigma8 ~ U[0.7, 0.9]; 31 KDK steps
    from a = 1/32.  Returns [(arrival offset in s, SimRequest)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1 / rate)
        u = rng.random()
        g = 64 if u < 0.4 else (128 if u < 0.8 else 256)
        req = SimRequest(seed=int(rng.integers(2**31)), ptcl_grid=g, mesh_ratio=2, box=2.0 * g,
                         Omega_m=float(rng.uniform(0.25, 0.35)),
                         sigma8=float(rng.uniform(0.7, 0.9)), grad=bool(rng.random() < 0.5))
        out.append((t, req))
    return out


def run_load(server, schedule, t0=None):
    """Submits each request at its arrival offset (sleeping in between) and
    returns the results in submission order."""
    t0 = time.perf_counter() if t0 is None else t0
    futs = []
    for dt, req in schedule:
        wait = t0 + dt - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        futs.append(server.submit(req, t_arrive=t0 + dt))
    return [f.result() for f in futs]


def latency_stats(results, t_begin, t_end):
    lat = np.array([r.latency for r in results])
    upd = sum(r.request.ptcl_grid**3 * r.n_steps * (2 if r.request.grad else 1) for r in results)
    return {'requests': len(results), 'p50_s': float(np.percentile(lat, 50)),
            'p99_s': float(np.percentile(lat, 99)), 'mean_s': float(lat.mean()),
            'max_s': float(lat.max()), 'wall_s': t_end - t_begin,
            'throughput_req_per_s': len(results) / (t_end - t_begin),
            'particle_updates_per_s': upd / (t_end - t_begin)}
