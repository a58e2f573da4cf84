#!/usr/bin/env python
"""Particle-updates/s of pmwd's KDK N-body step on one MI355X.

Workload (default): 512^3 particles on a 512^3 mesh (mesh_shape=1), box
1000 Mpc/h, FP32 state, 2LPT initial conditions at a=1/32 from device white
noise (Philox, seed 0), sigma8=0.8, Omega_m=0.3, Omega_b=0.05, h=0.7,
n_s=0.96, then KDK steps with a_nbody_maxstep=1/32 (31 steps to a=1).
A "step" is one KDK step (kick, drift, CIC scatter, FFT Poisson solve,
3-component CIC gather, kick) over all particles; ``value`` = particles x
steps / wall time of the K timed steps, inputs resident in HBM.  The timed
steps are the last K of the 31-step schedule (ending at a=1, where the
particles are most clustered); the steps before them run untimed, W of
them immediately before as warm-up.  The whole 31-step ``nbody()`` call
(first force included) is timed separately.  After the forward run the
reverse-time adjoint runs K steps from the final state and
``fwd_adj_value`` = particles x steps / (forward + adjoint).

Prints one JSON line.  ``--gpus`` above 1 exits: this build runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser(description=__doc__.split('\n')[0])
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=29)
    p.add_argument('--warmup', type=int, default=2)
    p.add_argument('--grid', type=int, default=512, help='particles per dimension')
    p.add_argument('--mesh-ratio', type=int, default=1)
    p.add_argument('--box', type=float, default=1000.0)
    p.add_argument('--no-adjoint', action='store_true')
    p.add_argument('--no-cpu-baseline', action='store_true')
    p.add_argument('--no-serve', action='store_true', help='skip the serving load')
    p.add_argument('--no-nbody-call', action='store_true',
                   help='skip timing the whole nbody() call (profiles: the timed K-step call '
                        'is then the last of the forward)')
    p.add_argument('--full-adjoint', type=int, default=0, metavar='G',
                   help='instead of the default run: forward plus gradient of a density MSE '
                        'for G^3 particles on a G^3 mesh (box 2000*G/1024 Mpc/h, 31 steps, '
                        'reverse-time adjoint), with device memory per phase')
    p.add_argument('--full-adjoint-warm', action='store_true',
                   help='with --full-adjoint: run the forward and gradient a second time in '
                        'the same process and report its times too')
    p.add_argument('--full-adjoint-checkpoints', type=int, default=0, metavar='S',
                   help='with --full-adjoint: snap the reverse run to states stored every S '
                        'steps, on the device when they fit, else in pinned host memory '
                        '(0: none, pure reversal as pmwd)')
    p.add_argument('--serve-requests', type=int, default=40)
    p.add_argument('--serve-rate', type=float, default=2.0, help='Poisson arrivals per second')
    p.add_argument('--serve-slots', type=int, default=8)
    p.add_argument('--no-fft-yardstick', action='store_true',
                   help='skip timing rocFFT (torch.fft) next to the library FFTs')
    p.add_argument('--cpu-slab', type=int, default=64,
                   help='fraction 1/slab of the particles in the CPU baseline sample')
    p.add_argument('--copy-calib', action='store_true',
                   help='also run 1 GiB device copies (counter calibration for profiles)')
    return p.parse_args()


def log(msg):
    print(f'[bench] {msg}', file=sys.stderr, flush=True)


def kernel_bytes(N, mesh_shape, tile, K, fold_on_load=False):
    """HBM bytes of each timed kernel class over a K-step pmx_nbody_steps
    call: (algorithmic, by design).

    Algorithmic bytes count each logical input read once and each output
    written once: pmid 6 B (int16 x3) and disp/vel/acc 12 B per particle,
    4 B per real and 8 B per complex (r2c layout) mesh cell.  An FFT pass
    reads its input and writes its output once.  Per step the scatter reads
    the positions (pmid, disp: 18N) and writes the density (4M); the fused
    step kernel reads pmid, disp, vel and the three force meshes and writes
    disp and vel (54N + 12M) -- gather, closing kick, next opening kick and
    drift; acc is written once, by the last step (which writes acc, vel and
    disp: 66N + 12M).  The first kick/drift/bin of the call reads vel, acc,
    disp, pmid and writes vel, disp (66N).  The second figure is what the
    kernels move by design: tile-ordered particles of 32 B (s0, s1)
    in place of pmid/disp/vel, tile and rank words, the +1 halo shells.
    With fold_on_load the Poisson row pass reads those shells as it loads
    the density and no fold pass runs.
    """
    n0, n1, n2 = mesh_shape
    M = n0 * n1 * n2
    Mh = n0 * n1 * (n2 // 2 + 1)
    tx, ty, tz = tile
    ntiles = -(-n0 // tx) * -(-n1 // ty) * -(-n2 // tz)
    shell = (ty + 1) * (tz + 1) + tx * (tz + 1) + tx * ty
    faces = ty * tz + (tx - 1) * tz + (tx - 1) * (ty - 1)
    return {
        # first kick + drift + exact count: read vel, acc, disp, pmid, write
        # vel, disp; by design also the tile and rank words
        'bin': (66 * N, 74 * N),
        'scan': (0, 8 * ntiles * (2 * K - 1)),
        # first fill: read pmid, disp, vel, tile, rank; write 32 B per particle
        'fill': (0, 70 * N),
        'deposit': (K * (18 * N + 4 * M), K * (16 * N + 4 * M + 4 * shell * ntiles)),
        'fold': (0, 0 if fold_on_load else K * (4 * shell + 8 * faces) * ntiles),
        'z_r2c': (K * (4 * M + 8 * Mh),
                  K * (4 * M + 8 * Mh + (4 * shell * ntiles if fold_on_load else 0))),
        'y_fwd': (K * 16 * Mh,) * 2,
        # read one complex mesh, write two (the x component and the
        # transform shared by the y and z components)
        'poisson_x': (K * 24 * Mh,) * 2,
        # x component in place; the shared transform read once, the y and z
        # components written
        'y_inv': (K * 40 * Mh,) * 2,
        # three launches, each reads a complex and writes a real mesh
        'z_c2r': (K * 3 * (8 * Mh + 4 * M),) * 2,
        'gather': ((K - 1) * (54 * N + 12 * M) + 66 * N + 12 * M,
                   (K - 1) * (64 * N + 12 * M) + 68 * N + 12 * M),
    }


def step_bytes(N, M, Mh):
    """Algorithmic bytes of one forward KDK step by kernel group, the
    compulsory traffic of pmwd's logical operations (each input read once,
    each output written once): scatter 18N + 4M (pmid and disp in, density
    out); 3D r2c 4M + 8Mh and Poisson (Green, -ik, 3 c2r) 8Mh + 12M; the
    rest of the step -- the 3-component gather, both kicks and the drift,
    reading pmid, disp, vel, acc and the force meshes, writing disp, vel,
    acc -- 60N + 12M.  Together 78N + 8M + 16Mh + 24M (118 B per particle
    at mesh = particle grid)."""
    return {'scatter': 18 * N + 4 * M,
            'fft_poisson': 16 * M + 16 * Mh,
            'step_gather_kick_drift': 60 * N + 12 * M}


def adj_step_bytes(N, M, Mh):
    """Algorithmic bytes of one reverse KDK step (kick_adj, drift_adj,
    force_adj, kick_adj of nbody.py:49-118) by kernel group, counted the same
    way: the density and the three cotangent deposits read pmid, disp and
    vel_cot and write four meshes (48N + 16M); the forward Poisson solve
    (16M + 16Mh) and the adjoint one (three r2c of the cotangent meshes, the
    Green/-ik combination fused into the inverse: 16M + 48Mh); the reverse
    step kernel reads pmid, disp, vel, disp_cot, vel_cot, the three force
    meshes and the density cotangent and writes disp, vel, disp_cot and
    vel_cot (102N + 16M)."""
    return {'deposits': 48 * N + 16 * M,
            'fft_poisson_fwd_adj': 32 * M + 64 * Mh,
            'reverse_step': 102 * N + 16 * M}


GROUPS = {
    # re-binning, capacity scan, first fill of a call, and the scatter
    'scatter': ('bin', 'scan', 'fill', 'deposit', 'fold'),
    # 3D r2c plus Poisson (Green, -ik, 3 c2r)
    'fft_poisson': ('z_r2c', 'y_fwd', 'poisson_x', 'y_inv', 'z_c2r'),
    # gather, closing kick, next opening kick and drift, move to the new tile
    'step_gather_kick_drift': ('gather',),
}

ADJ_GROUPS = {
    'deposits': ('bin', 'scan', 'fill', 'deposit', 'fold'),
    'fft_poisson_fwd_adj': ('z_r2c', 'y_fwd', 'poisson_x', 'y_inv', 'z_c2r'),
    'reverse_step': ('grad_gather', 'kd_adj', 'vjp_gather'),
}


def cpu_info():
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for ln in f:
                if ln.startswith('model name'):
                    model = ln.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return {'cpu_count': os.cpu_count(), 'cpu_model': model}


def cpu_baseline(conf, cosmo, ptcl, dens, slab, reps=3):
    """The numpy oracle timed on one KDK step and one reverse step of the
    run's own 2LPT state, on the host cores (threads: at most 16, the
    host share of one GPU).

    Sample: the full-mesh FFT Poisson solve (scipy.fft, all threads) on the
    run's density, and the particle work -- kick, drift, CIC setup and
    scatter, 3-component gather, kick -- for 1/slab of the particles (the
    first 1/slab of them, numpy on a thread pool), scaled by slab.  The
    reverse step: the oracle's force VJP (forward recompute plus the adjoint
    chain) on the same sample, its mesh part timed with one particle.
    Median of ``reps`` runs after one warm-up.  Returns (fwd upd/s,
    fwd+adj upd/s, details)."""
    import numpy as np
    import scipy.fft as sfft
    from oracle import pm as opm
    threads = min(16, os.cpu_count() or 1)
    shape = conf.mesh_shape
    N = conf.ptcl_num
    ns = N // slab
    pmid = ptcl.pmid[:ns].cpu().numpy()
    disp0 = ptcl.disp[:ns].cpu().numpy()
    vel0 = ptcl.vel[:ns].cpu().numpy()
    acc0 = np.zeros_like(disp0)
    dens = dens.cpu().numpy()
    cell, om = conf.cell_size, cosmo.Omega_m
    f32 = np.float32

    def fwd():
        t0 = time.perf_counter()
        vel = vel0 + acc0 * f32(0.01)
        disp = disp0 + vel * f32(0.01)
        cics, _ = opm.force_particle_setup(pmid, disp, cell, shape, N)
        t1 = time.perf_counter()
        with sfft.set_workers(threads):
            g = opm.gravity_mesh(dens, om, cell, shape)
        t2 = time.perf_counter()
        acc = opm.force_particle_readout(cics, g, ns)
        vel = vel + acc * f32(0.01)
        t3 = time.perf_counter()
        return (t1 - t0) + (t3 - t2), t2 - t1

    dc0 = (1e-3 * np.random.default_rng(5).standard_normal(disp0.shape)).astype(f32)
    vc0 = (1e-3 * np.random.default_rng(6).standard_normal(disp0.shape)).astype(f32)

    def adj(n):
        """kick_adj, drift_adj, force_adj, kick_adj of n particles."""
        k, d = f32(-0.01), f32(-0.01)
        t0 = time.perf_counter()
        vel = vel0[:n] + acc0[:n] * k
        dcot = dc0[:n] - acc0[:n] * k
        disp = disp0[:n] + vel * d
        vcot = vc0[:n] - dcot * d
        dots = [float(np.sum(vcot.astype(np.float64) * acc0[:n])),
                float(np.sum(dcot.astype(np.float64) * vel))]
        with sfft.set_workers(threads):
            acc, acot, _ = opm.gravity_vjp(pmid[:n], disp, om, cell, shape, vcot, ptcl_num=N)
        vel = vel + acc * k
        dcot = dcot - acot * k
        return time.perf_counter() - t0 + 0 * sum(dots)

    fwd()
    runs = [fwd() for _ in range(reps)]
    t_ptcl = float(np.median([r[0] for r in runs]))
    t_mesh = float(np.median([r[1] for r in runs]))
    t_fwd = t_ptcl * slab + t_mesh
    adj(1)
    t_adj_mesh = adj(1)
    t_adj_all = float(np.median([adj(ns) for _ in range(reps)]))
    t_adj = max(t_adj_all - t_adj_mesh, 0.0) * slab + t_adj_mesh
    info = dict(threads=threads, t_step_s=t_fwd, t_adj_step_s=t_adj, t_mesh_s=t_mesh,
                t_ptcl_sample_s=t_ptcl, t_adj_mesh_s=t_adj_mesh, reps=reps,
                sample=f'one KDK step of the run\'s 2LPT state: full {shape[0]}^3-mesh FFT '
                       f'Poisson solve (scipy.fft, {threads} threads) on the run\'s density + '
                       f'kick/drift/CIC scatter/3-component gather of the first 1/{slab} of the '
                       f'particles ({ns}, numpy on {threads} threads), particle part scaled '
                       f'x{slab}; median of {reps} after a warm-up', **cpu_info())
    return N / t_fwd, N / (t_fwd + t_adj), info


def serving_bench(dev, n, rate, slots):
    """p50/p99 latency of the serving mix (pmwd_amd.serving.c5_mix: Poisson
    arrivals; 40% 64^3/128^3, 40% 128^3/256^3, 20% 256^3/512^3 particles/mesh;
    half with gradients; 31 steps) on one GPU after one warm-up request of
    each size and kind (contexts resident, as in steady state)."""
    import time
    import numpy as np
    from pmwd_amd import serving as sv
    server = sv.Server(dev, max_slots=slots)
    warm = [sv.SimRequest(seed=1, ptcl_grid=g, mesh_ratio=2, box=2.0 * g, Omega_m=0.3,
                          sigma8=0.8, grad=gr) for g in (64, 128, 256) for gr in (False, True)]
    for f in [server.submit(r) for r in warm]:
        f.result()
    sched = sv.c5_mix(n, rate, seed=0)
    t0 = time.perf_counter()
    res = sv.run_load(server, sched, t0)
    t_end = max(r.t_done for r in res)
    packs = server.pack_stats()
    server.close()
    out = sv.latency_stats(res, t0, t_end)
    by = {}
    for g in (64, 128, 256):
        lat = [r.latency for r in res if r.request.ptcl_grid == g]
        if lat:
            by[f'{g}^3/{2 * g}^3'] = {'n': len(lat), 'p50_s': float(np.percentile(lat, 50)),
                                      'p99_s': float(np.percentile(lat, 99))}
    out.update(rate_per_s=rate, max_slots=slots, peak_concurrent=server.peak_active,
               packs=packs,
               hbm_budget_gb=server.hbm_budget / 1e9, by_size=by,
               mix='40% 64^3/128^3, 40% 128^3/256^3, 20% 256^3/512^3; half forward+gradient; '
                   'Omega_m~U[0.25,0.35], sigma8~U[0.7,0.9]; 31 KDK steps; Poisson arrivals')
    return out


def serving_burst(dev, n=48, grid=64, slots=8):
    """Throughput of a burst of n forward-only 64^3/128^3 requests (31 steps,
    seeds and cosmologies differing) submitted at once, with the N-body runs
    replaying captured graphs (the server's default: each on its slot's
    stream) and without (each request eager on its own worker stream).
    Every slot has run and captured its graph before the timed burst."""
    import time
    import numpy as np
    from pmwd_amd import serving as sv
    rng = np.random.default_rng(5)
    out = {}
    for packs in (True, False):
        server = sv.Server(dev, max_slots=slots, packs=packs)
        warm = [sv.SimRequest(seed=100 + i, ptcl_grid=grid, mesh_ratio=2, box=2.0 * grid,
                              Omega_m=0.3, sigma8=0.8) for i in range(2 * slots)]
        for f in [server.submit(r) for r in warm]:
            f.result()
        reqs = [sv.SimRequest(seed=1000 + i, ptcl_grid=grid, mesh_ratio=2, box=2.0 * grid,
                              Omega_m=float(rng.uniform(0.25, 0.35)),
                              sigma8=float(rng.uniform(0.7, 0.9))) for i in range(n)]
        t0 = time.perf_counter()
        res = [f.result() for f in [server.submit(r) for r in reqs]]
        wall = time.perf_counter() - t0
        stats = server.pack_stats()
        server.close()
        lat = [r.latency for r in res]
        key = 'graphs' if packs else 'eager'
        steps = 31
        out[key] = {'wall_s': wall, 'req_per_s': n / wall,
                    'particle_updates_per_s': n * grid**3 * steps / wall,
                    'p50_s': float(np.percentile(lat, 50)), 'p99_s': float(np.percentile(lat, 99)),
                    'packs': stats.get(f'{grid}^3/{2 * grid}^3') if packs else None}
    out['workload'] = (f'{n} forward-only requests of {grid}^3 particles on a {2 * grid}^3 mesh '
                       f'(31 KDK steps) submitted at once, {slots} slots')
    return out


def fft_yardstick(conf, ctx, h, L, _lib, dev, reps=5):
    """rocFFT (torch.fft.rfftn/irfftn) beside the library's pmx_rfftn/irfftn on
    one mesh: mean ms per transform (HIP events on the library stream), the
    library's three passes alone (timer classes, without the layout copies
    between contiguous and padded rows), and the relative L2 difference."""
    import torch
    shape = conf.mesh_shape
    x = torch.randn(shape, device=dev)
    nh = shape[2] // 2 + 1

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / reps

    # rocFFT's r2c result is strided (not C-contiguous) and a c2r may use its
    # input as scratch: every consumer gets its own contiguous copy
    yr = torch.fft.rfftn(x).contiguous()
    xr = torch.fft.irfftn(yr.clone(), s=shape)
    yt = yr.clone()
    roc_r2c = timed(lambda: torch.fft.rfftn(x))
    roc_c2r = timed(lambda: torch.fft.irfftn(yt, s=shape))
    yo = torch.empty(shape[0], shape[1], nh, dtype=torch.complex64, device=dev)
    xo = torch.empty(shape, device=dev)
    r2c = lambda: _lib.check(L.pmx_rfftn(h, _lib.ptr(x), _lib.ptr(yo)))
    c2r = lambda: _lib.check(L.pmx_irfftn(h, _lib.ptr(yr), _lib.ptr(xo)))
    our_r2c = timed(r2c)
    our_c2r = timed(c2r)
    passes = {}
    for name, fn in (('r2c', r2c), ('c2r', c2r)):
        _lib.check(L.pmx_timing_enable(h, 1))
        _lib.timing_read(ctx)
        for _ in range(reps):
            fn()
        t = _lib.timing_read(ctx)
        _lib.check(L.pmx_timing_enable(h, 0))
        passes[name] = sum(t[k][0] for k in ('z_r2c', 'y_fwd', 'y_inv', 'z_c2r', 'other')) / reps
    r2c()
    c2r()
    torch.cuda.synchronize()
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    return {'mesh': list(shape), 'rocfft_r2c_ms': roc_r2c, 'rocfft_c2r_ms': roc_c2r,
            'pmx_r2c_ms': our_r2c, 'pmx_c2r_ms': our_c2r,
            'pmx_r2c_passes_ms': passes['r2c'], 'pmx_c2r_passes_ms': passes['c2r'],
            'r2c_rel_l2_vs_rocfft': rel(yo, yr), 'c2r_rel_l2_vs_rocfft': rel(xo, xr)}


def full_adjoint(G, dev, ckpt_every=0, warm=False):
    """Forward plus gradient at G^3 particles on a G^3 mesh (the 1024^3
    configuration at G=1024): white modes (seed 0) -> linear_modes -> 2LPT ->
    31 KDK steps -> density; loss = mean((rho - target)^2) against a seeded
    target density; gradients with respect to the modes, Omega_m and sigma8
    through the reverse-time adjoint, which reconstructs the trajectory from
    the final state, snapped to the states stored every ckpt_every steps (on
    the device when they fit, else in pinned host memory; 0: pure reversal
    as pmwd).  The modes -> 2LPT part
    is recomputed in the backward rather than kept.  Reports wall times and
    device memory per phase (torch peaks, library contexts, device-wide use).
    With warm, the same forward and gradient run a second time in the
    process (device memory already mapped, as in an optimisation loop):
    its times, whether its loss and modes gradient equal the first run's
    bit for bit, and how far its Omega_m and sigma8 gradients are from the
    first run's (float64 sums whose grouping follows the particles' order
    within a tile, which the atomics of the re-binning set)."""
    import time
    import torch
    import pmwd_amd as pa
    from pmwd_amd import _lib
    conf = pa.Configuration(2000.0 / 1024, (G, G, G), mesh_shape=1, a_start=1 / 32,
                            a_nbody_maxstep=1 / 32)
    torch.cuda.reset_peak_memory_stats(dev)
    t0 = time.perf_counter()
    cosmo = pa.boltzmann(pa.Cosmology.from_sigma8(conf, 0.8, n_s=0.96, Omega_m=0.3,
                                                  Omega_b=0.05, h=0.7), conf)
    white = pa.white_noise(0, conf, real=True, device=dev).requires_grad_(True)
    gen = torch.Generator(device=dev).manual_seed(1)
    target = 1 + 0.5 * torch.randn(conf.mesh_shape, generator=gen, device=dev)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    theta = {'Omega_m': torch.tensor(cosmo.Omega_m, dtype=torch.float64, requires_grad=True),
             'sigma8': torch.tensor(cosmo.sigma8, dtype=torch.float64, requires_grad=True)}

    def ics(w, om, s8):
        th = {'Omega_m': om, 'sigma8': s8}
        ptcl, _ = pa.lpt(pa.linear_modes(w, cosmo, conf, theta=th), cosmo, conf, theta=th)
        return ptcl.disp, ptcl.vel, ptcl.pmid

    def used():
        free, total = torch.cuda.mem_get_info(dev)
        return (total - free) / 1e9

    def once():
        # The initial conditions are recomputed in the backward rather than
        # kept (as torch.utils.checkpoint would); between the N-body adjoint
        # and that recomputation the library contexts give back their
        # particle scratch.
        torch.cuda.reset_peak_memory_stats(dev)
        t0 = time.perf_counter()
        with torch.no_grad():
            disp0, vel0, pmid = ics(white, theta['Omega_m'], theta['sigma8'])
        disp0.requires_grad_(True)
        vel0.requires_grad_(True)
        ptcl, _ = pa.nbody(pa.Particles(conf, pmid, disp0, vel0), None, cosmo, conf,
                           omega_m=theta['Omega_m'], checkpoint_every=ckpt_every or None,
                           checkpoint_device='auto')
        dens = pa.scatter(ptcl, conf)
        del ptcl
        # the difference in float32, its square and mean in float64 (fewer
        # full-mesh temporaries than casting both densities first)
        loss = ((dens - target).double()**2).mean()
        del dens
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        mem = {'library_context_gb_forward': sum(c.nbytes for c in _lib._ctx_cache.values()) / 1e9,
               'device_used_gb_after_forward': used()}
        t0 = time.perf_counter()
        g_d0, g_v0, g_om1 = torch.autograd.grad(loss, (disp0, vel0, theta['Omega_m']))
        torch.cuda.synchronize()
        t_nbody_bwd = time.perf_counter() - t0
        mem['library_context_gb_adjoint'] = sum(c.nbytes for c in _lib._ctx_cache.values()) / 1e9
        mem['device_used_gb_after_adjoint'] = used()
        mem['torch_peak_gb_through_adjoint'] = torch.cuda.max_memory_allocated(dev) / 1e9
        loss_value = float(loss.detach())
        del disp0, vel0, loss
        _lib.trim_contexts()
        torch.cuda.reset_peak_memory_stats(dev)
        d1, v1, _ = ics(white, theta['Omega_m'], theta['sigma8'])
        g_w, g_om2, g_s8 = torch.autograd.grad((d1, v1), (white, theta['Omega_m'], theta['sigma8']),
                                               grad_outputs=(g_d0, g_v0), allow_unused=True)
        g_om = g_om1 + (g_om2 if g_om2 is not None else 0)
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        mem['torch_peak_gb_ic_backward'] = torch.cuda.max_memory_allocated(dev) / 1e9
        mem['library_context_gb_ic_backward'] = sum(c.nbytes for c in _lib._ctx_cache.values()) / 1e9
        mem['device_peak_gb_reported'] = max(mem['device_used_gb_after_forward'],
                                             mem['device_used_gb_after_adjoint'])
        del d1, v1, g_d0, g_v0
        return dict(t_fwd=t_fwd, t_bwd=t_bwd, t_nbody_bwd=t_nbody_bwd, mem=mem,
                    loss=loss_value, g_w=g_w, g_om=float(g_om), g_s8=float(g_s8))

    r = once()
    g_w = r['g_w']
    # the modes gradient along three seeded directions (comparable across runs)
    gen = torch.Generator(device=dev).manual_seed(123)
    proj = []
    for _ in range(3):
        u = torch.randn(g_w.shape, generator=gen, device=dev, dtype=g_w.dtype)
        proj.append(float((g_w.double() * u.double()).sum()))
        del u
    steps = len(conf.a_nbody) - 1
    n = conf.ptcl_num
    out = {'workload': f'{G}^3 particles on a {G}^3 mesh, box {2000.0 * G / 1024:g} Mpc/h, '
                       f'{steps} KDK steps from a=1/32, 2LPT, density MSE against a seeded '
                       f'target, gradients w.r.t. modes, Omega_m, sigma8 by reverse-time '
                       f'integration from the final state'
                       + (f', snapped to states stored every {ckpt_every} steps (on the '
                          f'device when they fit, else in pinned host memory)'
                          if ckpt_every else ''),
           't_setup_s': t_setup, 't_forward_s': r['t_fwd'], 't_backward_s': r['t_bwd'],
           't_nbody_backward_s': r['t_nbody_bwd'],
           'fwd_updates_per_s': n * steps / r['t_fwd'],
           'fwd_adj_updates_per_s': n * steps / (r['t_fwd'] + r['t_bwd']),
           'memory': r['mem'], 'loss': r['loss'],
           'grad_modes_norm': float(g_w.double().norm()), 'grad_modes_proj': proj,
           'grad_Omega_m': r['g_om'],
           'grad_sigma8': r['g_s8']}
    if warm:
        # the first run's modes gradient waits in host memory (the device is
        # nearly full at the adjoint's peak)
        first = (r['loss'], r['g_om'], r['g_s8'], g_w.detach().cpu())
        del r, g_w
        w = once()
        out['warm'] = {'t_forward_s': w['t_fwd'], 't_backward_s': w['t_bwd'],
                       't_nbody_backward_s': w['t_nbody_bwd'],
                       'fwd_updates_per_s': n * steps / w['t_fwd'],
                       'fwd_adj_updates_per_s': n * steps / (w['t_fwd'] + w['t_bwd']),
                       'memory': w['mem'],
                       'same_loss': w['loss'] == first[0],
                       'same_modes_gradient': bool(torch.equal(w['g_w'].cpu(), first[3])),
                       'rel_diff_Omega_m': abs(w['g_om'] - first[1]) / abs(first[1]),
                       'rel_diff_sigma8': abs(w['g_s8'] - first[2]) / abs(first[2])}
    return out


def main():
    args = parse()
    world =int(os.environ.get('WORLD_SIZE', '1'))
    if args.gpus != 1 or world != 1:
        print(f'bench.py: this build runs on one GPU; --gpus {args.gpus} (WORLD_SIZE={world}) '
              'is not supported. The multi-GPU decomposition is a design note only.',
              file=sys.stderr)
        sys.exit(2)

    import ctypes
    import numpy as np
    import torch
    import pmwd_amd as pa
    from pmwd_amd import _lib
    nb = sys.modules['pmwd_amd.nbody']

    dev = torch.device('cuda:0')
    torch.cuda.set_device(dev)
    if args.full_adjoint:
        log(f'full forward + gradient at {args.full_adjoint}^3')
        print(json.dumps({'full_adjoint': full_adjoint(args.full_adjoint, dev,
                                                       args.full_adjoint_checkpoints,
                                                       args.full_adjoint_warm)}))
        return
    G = args.grid
    spacing = args.box / G
    conf = pa.Configuration(spacing, (G, G, G), mesh_shape=args.mesh_ratio, a_start=1 / 32,
                            a_nbody_maxstep=1 / 32)
    cosmo = pa.Cosmology.from_sigma8(conf, 0.8, n_s=0.96, Omega_m=0.3, Omega_b=0.05, h=0.7)
    cosmo = pa.boltzmann(cosmo, conf)
    N, M = conf.ptcl_num, conf.mesh_size
    Mh = conf.mesh_shape[0] * conf.mesh_shape[1] * (conf.mesh_shape[2] // 2 + 1)

    if args.copy_calib:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
        log('copy calibration: 3 x 1 GiB runtime copies, 3 x 1 GiB elementwise neg')

    log(f'initial conditions {G}^3 (2LPT)')
    torch.cuda.synchronize()
    t = time.perf_counter()
    white = pa.white_noise(0, conf, real=True, device=dev)
    modes = pa.linear_modes(white, cosmo, conf)
    ptcl_ic, _ = pa.lpt(modes, cosmo, conf)
    del modes
    torch.cuda.synchronize()
    t_ic = time.perf_counter() - t
    # warm lpt(linear_modes(...)), the quantity of pmwd's 2LPT timing
    # (docs/papers/adjoint/perf.py: min over repeats); and the host tables
    t_lpt = []
    for _ in range(5):
        torch.cuda.synchronize()
        t = time.perf_counter()
        p2, _ = pa.lpt(pa.linear_modes(white, cosmo, conf), cosmo, conf)
        torch.cuda.synchronize()
        t_lpt.append(time.perf_counter() - t)
        del p2
    t_tab = []
    for _ in range(3):
        t = time.perf_counter()
        pa.boltzmann(cosmo, conf, varlin=False)
        t_tab.append(time.perf_counter() - t)
    del white

    pmid = ptcl_ic.pmid.contiguous()
    disp = ptcl_ic.disp.clone().contiguous()
    vel = ptcl_ic.vel.clone().contiguous()
    acc = torch.empty_like(disp)
    L = _lib.lib()
    ctx = _lib.context(conf.mesh_shape, conf.cell_size, dev)
    h = ctx.bind()
    val, coef = nb._mass_val(conf), nb.omega_coef(cosmo, conf)

    a = conf.a_nbody
    nsteps_sched = len(a) - 1
    sched = [nb._kdk_factors(a[i], a[i + 1], cosmo, conf) for i in range(nsteps_sched)]

    def steps(i0, k):
        """k KDK steps of the schedule from step i0 in one pmx_nbody_steps call"""
        f = np.array([sched[i % nsteps_sched] for i in range(i0, i0 + k)], dtype=np.float32)
        k1, d, k2 = (np.ascontiguousarray(f[:, m]) for m in range(3))
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        _lib.check(L.pmx_nbody_steps(h, _lib.ptr(pmid), _lib.ptr(disp), _lib.ptr(vel),
                                     _lib.ptr(acc), N, val, coef, k, P(k1), P(d), P(k2)))

    K, W = args.steps, args.warmup
    # the timed steps end at a = 1; the steps before them run untimed (the W
    # just before as warm-up).  With fewer untimed steps than W, a W-step
    # warm-up run from the initial conditions comes first.
    i0 = max(nsteps_sched - K, 0)
    _lib.check(L.pmx_gravity(h, _lib.ptr(pmid), _lib.ptr(disp), N, val, coef, _lib.ptr(acc)))
    if W and i0 < W:
        steps(0, W)
        disp.copy_(ptcl_ic.disp)
        vel.copy_(ptcl_ic.vel)
        _lib.check(L.pmx_gravity(h, _lib.ptr(pmid), _lib.ptr(disp), N, val, coef, _lib.ptr(acc)))
    if i0 > W:
        steps(0, i0 - W)
    if i0 and min(W, i0):
        steps(i0 - min(W, i0), min(W, i0))
    torch.cuda.synchronize()

    log(f'timed forward: {K} steps from step {i0}')
    _lib.check(L.pmx_timing_enable(h, 1))
    _lib.timing_read(ctx)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    steps(i0, K)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    timing = _lib.timing_read(ctx)
    _lib.check(L.pmx_timing_enable(h, 0))

    ms_per_step = 1e3 * t_fwd / K
    value = N * K / t_fwd
    a_to = a[(i0 + K) % nsteps_sched] if i0 + K != nsteps_sched else a[-1]

    # per kernel class: time per step, launches, bytes per launch
    ts = (ctypes.c_int * 3)()
    _lib.check(L.pmx_tile_shape(h, ctypes.byref(ts)))
    # no fold launches: the row pass of the Poisson solve folds on load
    kb = kernel_bytes(N, conf.mesh_shape, tuple(ts), K,
                      fold_on_load=not timing.get('fold', (0, 0))[1])
    kernels = {}
    for k, (ms_tot, cnt) in timing.items():
        if not cnt:
            continue
        e = dict(ms_per_step=ms_tot / K, launches_per_step=cnt / K,
                 ms_per_launch=ms_tot / cnt)
        if k in kb:
            alg, impl = kb[k]
            e['alg_bytes_per_launch'] = alg / cnt
            e['impl_bytes_per_launch'] = impl / cnt
            e['alg_gbs'] = alg / (ms_tot * 1e-3) / 1e9
            e['impl_gbs'] = impl / (ms_tot * 1e-3) / 1e9
        kernels[k] = e
    sb = step_bytes(N, M, Mh)
    groups = {}
    for gname, classes in GROUPS.items():
        ms = sum(kernels[c]['ms_per_step'] for c in classes if c in kernels)
        impl = sum(kb[c][1] for c in classes) / K
        groups[gname] = dict(ms=ms, alg_bytes=sb[gname], impl_bytes=impl,
                             gbs=sb[gname] / (ms * 1e-3) / 1e9 if ms > 0 else None)
        groups[gname]['frac'] = groups[gname]['gbs'] / HBM_PEAK_GBS if ms > 0 else None
    dom = max((k for k in kernels if k in kb and kb[k][0]), key=lambda k: kernels[k]['ms_per_step'])
    B_step = sum(sb.values())

    traffic, traffic_K = None, None
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles',
                         f'traffic_{G}.json')
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj.get('kernels', {}).get(dom, {}).get('hbm_bytes_per_launch')
        traffic_K = tj.get('steps')

    # the whole 31-step nbody() call from the initial conditions: first
    # force, the host factors, one pmx_nbody_steps call (its first kick/drift
    # and binning included)
    nbody_call = None
    if not args.no_nbody_call:
        log('timed nbody(): the full schedule')
        pa.nbody(ptcl_ic, None, cosmo, conf)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fin, _ = pa.nbody(ptcl_ic, None, cosmo, conf)
        torch.cuda.synchronize()
        t_nbody = time.perf_counter() - t0
        del fin
        nbody_call = {'steps': nsteps_sched, 's': t_nbody,
                      'updates_per_s': N * nsteps_sched / t_nbody}

    out = {
        'metric': 'particle-updates/s fwd & fwd+adj at 512^3 (1 GPU); p50/p99 mixed-size latency',
        'value': value,
        'unit': 'particle-updates/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': W,
        'ms_per_step': ms_per_step,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic: 2LPT ICs from device white noise (Philox4x32-10, seed 0)',
        'config': {'workload': f'pmwd KDK N-body step, {G}^3 particles on a '
                               f'{conf.mesh_shape[0]}^3 mesh, box {args.box:g} Mpc/h, '
                               f'a_start=1/32, 31-step schedule, sigma8=0.8, Omega_m=0.3',
                   'ptcl_grid': G, 'mesh': conf.mesh_shape[0], 'box_mpc_h': args.box,
                   'parallelism': 'single GPU'},
        'timed_steps': {'first': i0, 'count': K, 'untimed_before': i0,
                        'a_from': float(a[i0 % nsteps_sched]), 'a_to': float(a_to),
                        'schedule_wraps': i0 + K > nsteps_sched},
        'roofline': {'bound': 'hbm', 'kernel': dom,
                     'achieved': kernels[dom]['alg_gbs'], 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': kernels[dom]['alg_gbs'] / HBM_PEAK_GBS, 'traffic': traffic,
                     'traffic_steps': traffic_K,
                     'algorithmic_bytes_per_launch': kernels[dom]['alg_bytes_per_launch'],
                     'ms_per_launch': kernels[dom]['ms_per_launch']},
        'whole_step_roofline': {'bytes_per_step': B_step,
                                'bytes_per_particle_update': B_step / N,
                                'frac': B_step / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        'kernel_groups': groups,
        'kernels': kernels,
        'nbody_call': nbody_call,
        't_ic_s': t_ic,
        'lpt_linear_modes_warm_ms': 1e3 * min(t_lpt),
        'growth_tables_ms': 1e3 * min(t_tab),
    }

    if not args.no_fft_yardstick:
        log('fft yardstick: rocFFT (torch.fft) vs library r2c/c2r')
        out['fft_yardstick'] = fft_yardstick(conf, ctx, h, L, _lib, dev)

    if not args.no_adjoint:
        log(f'timed adjoint: {K} steps')
        dc = torch.randn_like(disp) * 1e-3
        vc = torch.randn_like(vel) * 1e-3
        acc_cot = torch.empty_like(disp)
        dot = ctypes.c_double()
        adj_sched = []
        a_rev = list(zip(a[:0:-1], a[-2::-1]))
        for a_prev, a_next in a_rev:
            ah = a_prev * (1 - 0.5) + a_next * 0.5
            adj_sched.append((nb._f32(nb.kick_factor(a_prev, a_prev, ah, cosmo, conf)),
                              nb._f32(nb.drift_factor(ah, a_prev, a_next, cosmo, conf)),
                              nb._f32(nb.kick_factor(a_next, ah, a_next, cosmo, conf))))
        _lib.check(L.pmx_gravity_vjp(h, _lib.ptr(pmid), _lib.ptr(disp), N, val, coef, _lib.ptr(vc),
                                     _lib.ptr(acc), _lib.ptr(acc_cot), ctypes.byref(dot)))
        torch.cuda.synchronize()
        _lib.check(L.pmx_timing_enable(h, 1))
        _lib.timing_read(ctx)
        t0 = time.perf_counter()
        # K reverse KDK steps from a = 1 in one pmx_nbody_steps_adj call
        # (particles and cotangents kept in tile order), the per-step float64
        # dots the Omega_m cotangent needs read back at the end as nbody's
        # adjoint does, then the pending closing kick
        f = np.array([adj_sched[i % nsteps_sched] for i in range(K)], dtype=np.float32)
        fa, fd, fb = (np.ascontiguousarray(f[:, m]) for m in range(3))
        adots = (ctypes.c_double * (2 * K))()
        P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        _lib.check(L.pmx_nbody_steps_adj(h, _lib.ptr(pmid), _lib.ptr(disp), _lib.ptr(vel),
                                         _lib.ptr(acc), _lib.ptr(dc), _lib.ptr(vc),
                                         _lib.ptr(acc_cot), N, val, coef, 0.0, K, P(fa), P(fd),
                                         P(fb), None, adots))
        _lib.check(L.pmx_kick_adj(h, N, _lib.ptr(vel), _lib.ptr(acc), _lib.ptr(dc),
                                  _lib.ptr(acc_cot), _lib.ptr(vc), float(fb[-1]), None))
        torch.cuda.synchronize()
        t_adj = time.perf_counter() - t0
        atiming = _lib.timing_read(ctx)
        _lib.check(L.pmx_timing_enable(h, 0))
        out['fwd_adj_value'] = N * K / (t_fwd + t_adj)
        out['adj_ms_per_step'] = 1e3 * t_adj / K
        out['adj_kernels'] = {k: dict(ms_per_step=v[0] / K, launches_per_step=v[1] / K)
                              for k, v in atiming.items() if v[1]}
        ab = adj_step_bytes(N, M, Mh)
        agroups = {}
        for gname, classes in ADJ_GROUPS.items():
            ms = sum(atiming[c][0] for c in classes if c in atiming) / K
            agroups[gname] = dict(ms=ms, alg_bytes=ab[gname],
                                  gbs=ab[gname] / (ms * 1e-3) / 1e9 if ms > 0 else None)
            agroups[gname]['frac'] = agroups[gname]['gbs'] / HBM_PEAK_GBS if ms > 0 else None
        out['adj_kernel_groups'] = agroups
        B_adj = sum(ab.values())
        out['adj_step_roofline'] = {'bytes_per_step': B_adj, 'bytes_per_particle_update': B_adj / N,
                                    'frac': B_adj / (1e-3 * out['adj_ms_per_step']) / 1e9
                                    / HBM_PEAK_GBS}

    if not args.no_serve:
        log(f'serving: {args.serve_requests} requests at {args.serve_rate}/s')
        out['serving'] = serving_bench(dev, args.serve_requests, args.serve_rate, args.serve_slots)
        log('serving: a burst of small forward requests, graph replays and eager')
        out['serving_burst'] = serving_burst(dev)

    if not args.no_cpu_baseline:
        log('cpu baseline sample (numpy oracle)')
        dens_ic = pa.scatter(ptcl_ic, conf)
        upd, upd_adj, info = cpu_baseline(conf, cosmo, ptcl_ic, dens_ic, args.cpu_slab)
        out['cpu_baseline'] = {'value': upd, 'unit': 'particle-updates/s',
                               'cores': info.pop('threads'), 'kind': 'port',
                               'sample': info.pop('sample'), 'fwd_adj_value': upd_adj, **info}
        out['speedup_vs_cpu'] = value / upd
        if 'fwd_adj_value' in out:
            out['fwd_adj_speedup_vs_cpu'] = out['fwd_adj_value'] / upd_adj

    print(json.dumps(out))


if __name__ == '__main__':
    main()
