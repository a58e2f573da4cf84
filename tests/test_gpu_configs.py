"""The benchmark configurations at their stated sizes.

* 64^3 particles on a 128^3 mesh, box 256 Mpc/h, 2LPT at a = 1/64, 63 KDK
  steps: final positions, velocities and CIC density against the oracle's
  FP32 run from the same white noise, with the error against its FP64 run
  reported beside them.
* 128^3 on a 256^3 mesh, 63 steps: gradients of the density MSE against a
  second seeded run (sigma8 = 0.85) with respect to the white modes, Omega_m
  and sigma8, against the float64 oracle's central differences (committed
  fixture); the GPU's float32 differences at eps = 1e-3 are recorded
  beside them; the N-body adjoint cotangents against the oracle's
  reverse-time adjoint over three steps.
* 512^3 on a 512^3 mesh, box 1000 Mpc/h, a = 1/32, 31 steps: full-size
  properties (mass conservation, bitwise reproducibility, the multi-step
  pipeline equal to per-step calls, reversibility within pmwd's published
  FP32 figure) and the 128^3 downscaled run against the oracle in FP32 and
  FP64.
* The serving mix's size classes (64^3/128^3, 128^3/256^3, 256^3/512^3):
  each request the same alone and in the mix.

The measured errors are printed, and written as JSON to the file named by
PMX_TEST_RECORD when it is set.
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle import cosmo as oc, ic as oic, nbody as onb, pm as opm
from oracle.conf import OConf

pytestmark = pytest.mark.gpu

RECORD = {}


def _record(name, **kw):
    RECORD[name] = {k: (float(v) if isinstance(v, (np.floating, float, int)) else v)
                    for k, v in kw.items()}
    print(name, json.dumps(RECORD[name]), flush=True)
    path = os.environ.get('PMX_TEST_RECORD')
    if path:
        with open(path, 'w') as f:
            json.dump(RECORD, f, indent=1)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cosine(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


COSMO = dict(n_s=0.96, Omega_b=0.05, h=0.7)


def _confs(spacing, n, ratio, a_start, **kw):
    import pmwd_amd as pa
    kw = dict(mesh_shape=ratio, a_start=a_start, a_nbody_maxstep=a_start, **kw)
    return pa.Configuration(spacing, (n,) * 3, **kw), OConf(spacing, (n,) * 3, **kw)


def _gpu_cosmo(conf, sigma8=0.8, Omega_m=0.3):
    import pmwd_amd as pa
    return pa.boltzmann(pa.Cosmology.from_sigma8(conf, sigma8, Omega_m=Omega_m, **COSMO), conf)


def _oracle_cosmo(oconf, sigma8=0.8, Omega_m=0.3):
    return oc.boltzmann(oc.from_sigma8(oconf, sigma8, Omega_m=Omega_m, **COSMO), oconf)


def _gpu_model(white, cosmo, conf, theta=None, checkpoint_every=None):
    """white noise -> linear_modes -> lpt -> nbody -> scatter on the GPU."""
    import pmwd_amd as pa
    modes = pa.linear_modes(white, cosmo, conf, theta=theta)
    ptcl0, _ = pa.lpt(modes, cosmo, conf, theta=theta)
    om = theta['Omega_m'] if theta else None
    ptcl, _ = pa.nbody(ptcl0, None, cosmo, conf, omega_m=om, checkpoint_every=checkpoint_every)
    return ptcl0, ptcl, pa.scatter(ptcl, conf)


def _oracle_model(white, ocosmo, oconf, dtype):
    modes = oic.linear_modes(white, ocosmo, oconf, dtype=dtype)
    pmid, d0, v0 = oic.lpt(modes, ocosmo, oconf, dtype=dtype)
    d1, v1, _ = onb.nbody(pmid, d0, v0, ocosmo, oconf, dtype=dtype)
    dens = opm.scatter(pmid, d1, oconf.cell_size, oconf.mesh_shape, dtype=dtype)
    return pmid, (d0, v0), (d1, v1), dens


def _forward_vs_oracle(name, dev, spacing, n, ratio, a_start, seed=0):
    import pmwd_amd as pa
    conf, oconf = _confs(spacing, n, ratio, a_start)
    cosmo = _gpu_cosmo(conf)
    white = pa.white_noise(seed, conf, real=True, device=dev)
    ptcl0, ptcl, dens = _gpu_model(white, cosmo, conf)
    g0 = (ptcl0.disp.cpu().numpy(), ptcl0.vel.cpu().numpy())
    g1 = (ptcl.disp.cpu().numpy(), ptcl.vel.cpu().numpy())
    gdens = dens.cpu().numpy()
    w = white.cpu().numpy()
    ocosmo = _oracle_cosmo(oconf)
    pmid, o0, o1, odens = _oracle_model(w, ocosmo, oconf, np.float32)
    _, x0, x1, xdens = _oracle_model(w, ocosmo, oconf, np.float64)
    np.testing.assert_array_equal(ptcl.pmid.cpu().numpy(), pmid)
    err = dict(
        lpt_disp_fp32=rel_l2(g0[0], o0[0]), lpt_vel_fp32=rel_l2(g0[1], o0[1]),
        disp_fp32=rel_l2(g1[0], o1[0]), vel_fp32=rel_l2(g1[1], o1[1]),
        dens_fp32=rel_l2(gdens, odens),
        disp_fp64=rel_l2(g1[0], x1[0]), vel_fp64=rel_l2(g1[1], x1[1]),
        dens_fp64=rel_l2(gdens, xdens),
        oracle32_disp_fp64=rel_l2(o1[0], x1[0]), oracle32_vel_fp64=rel_l2(o1[1], x1[1]),
        oracle32_dens_fp64=rel_l2(odens, xdens),
        disp_std=float(g1[0].std()), vel_std=float(g1[1].std()),
        lpt_disp_std=float(g0[0].std()), lpt_vel_std=float(g0[1].std()),
        mass=float(gdens.astype(np.float64).mean()))
    _record(name, **err)
    return err


def test_c1_forward_vs_oracle(dev):
    """64^3 particles, 128^3 mesh, box 256 Mpc/h, seed 0, 2LPT at a = 1/64,
    63 steps.  Density within 1e-4 relative L2 of the oracle's FP32 run;
    displacements within 1e-5 and velocities within 1e-4; each result's
    error against the oracle's FP64 run is reported, next to the FP32
    oracle's own."""
    e = _forward_vs_oracle('C1', dev, 4.0, 64, 2, 1 / 64)
    assert e['lpt_disp_fp32'] < 1e-6 and e['lpt_vel_fp32'] < 1e-6, e
    assert e['dens_fp32'] < 1e-4, e
    assert e['disp_fp32'] < 1e-5 and e['vel_fp32'] < 1e-4, e
    assert abs(e['mass'] - 1) < 1e-6
    # no worse against FP64 than the FP32 oracle itself, within a factor 2
    assert e['dens_fp64'] < 2 * e['oracle32_dens_fp64'] + 1e-6, e


def test_c3_downscaled_vs_oracle(dev):
    """The 512^3 configuration downscaled to 128^3 particles on a 128^3
    mesh in the same 1000 Mpc/h box (7.8125 Mpc/h spacing), a = 1/32, 31
    steps, against the oracle's FP32 and FP64 runs."""
    e = _forward_vs_oracle('C3_128', dev, 1000 / 128, 128, 1, 1 / 32)
    assert e['lpt_disp_fp32'] < 1e-6 and e['lpt_vel_fp32'] < 1e-6, e
    assert e['dens_fp32'] < 1e-4, e
    assert e['disp_fp32'] < 1e-5 and e['vel_fp32'] < 1e-4, e
    assert abs(e['mass'] - 1) < 1e-6
    assert e['dens_fp64'] < 2 * e['oracle32_dens_fp64'] + 1e-6, e


def test_c3_full_size_properties(dev):
    """512^3 particles on a 512^3 mesh, box 1000 Mpc/h, a = 1/32, 31 steps:
    mass conservation; two runs bitwise equal (pmwd's FP32 run-to-run RMSD
    is 1.4e-7, adjoint.tex:1466); three steps of the multi-step pipeline
    bitwise equal to three per-step calls; forward to a = 1 then back to
    a = 1/32 within pmwd's FP32 reversibility, 1.4e-3 / 2.2e-3 of the LPT
    disp/vel std (adjoint.tex:1538)."""
    import importlib
    import pmwd_amd as pa
    nbm = importlib.import_module('pmwd_amd.nbody')
    conf, _ = _confs(1000 / 512, 512, 1, 1 / 32)
    cosmo = _gpu_cosmo(conf)
    white = pa.white_noise(0, conf, real=True, device=dev)
    with torch.no_grad():
        ptcl0, _ = pa.lpt(pa.linear_modes(white, cosmo, conf), cosmo, conf)
    del white
    fin, _ = pa.nbody(ptcl0, None, cosmo, conf)
    dens = pa.scatter(fin, conf)
    mass = float(dens.double().mean())
    fin2, _ = pa.nbody(ptcl0, None, cosmo, conf)
    same = bool(torch.equal(fin.disp, fin2.disp) and torch.equal(fin.vel, fin2.vel))
    del fin2, dens
    back, _ = pa.nbody(fin, None, cosmo, conf, reverse=True)
    rev_d = float((back.disp - ptcl0.disp).double().std() / ptcl0.disp.double().std())
    rev_v = float((back.vel - ptcl0.vel).double().std() / ptcl0.vel.double().std())
    del back
    # three steps: one pmx_nbody_steps call against three pmx_nbody_step calls
    three = conf.replace(a_stop=conf.a_start + 3 * conf.a_nbody_maxstep)
    cosmo3 = _gpu_cosmo(three)
    runs = []
    for pipeline in (True, False):
        d, v = ptcl0.disp.clone(), ptcl0.vel.clone()
        acc, _ = nbm._run_forward(ptcl0.pmid, d, v, cosmo3, three, pipeline=pipeline)
        runs.append((d, v, acc))
    pipe_equal = all(torch.equal(a, b) for a, b in zip(*runs))
    _record('C3_512', mass=mass, reproducible=same, rev_disp=rev_d, rev_vel=rev_v,
            pipeline_equals_steps=pipe_equal, disp_std=float(fin.disp.std()),
            vel_std=float(fin.vel.std()), steps=len(conf.a_nbody) - 1)
    assert abs(mass - 1) < 1e-6
    assert same and pipe_equal
    assert rev_d < 1.4e-3 and rev_v < 2.2e-3, (rev_d, rev_v)


def _c2_fixture():
    return np.load(os.path.join(os.path.dirname(__file__), 'golden', 'c2_grad.npz'))


def test_c2_gradients(dev):
    """128^3 particles on a 256^3 mesh, box 256 Mpc/h, 63 steps; loss
    mean((rho - rho_tgt)^2) with the target from seed 1 at sigma8 = 0.85.
    The adjoint gradients with respect to Omega_m, sigma8 and along three
    random directions of the white modes, against the float64 oracle's
    central differences (tests/golden/c2_grad.npz, made by make_c2_grad.py).

    Two backward schemes: pure reverse-time integration from the final state
    (the default, as the reference) and the same snapped to states stored
    every 8 steps.  The float32 reversal retraces the initial state to
    1.0e-3 (disp) and 2.5e-3 (vel) relative, as the oracle's float32
    reversal does (0.99e-3 and 2.4e-3); over 63 steps that error moves the
    Omega_m gradient by 1.1% (its N-body part 9.21 against 9.31 in float64)
    and the white-mode directions by up to 0.8%, which the stored states
    remove (Omega_m 1.3e-3, directions 6e-4).  sigma8 acts through the
    initial conditions only, where both schemes agree with float64 to
    5e-5."""
    import pmwd_amd as pa
    z = _c2_fixture()
    conf, _ = _confs(2.0, 128, 2, 1 / 64)
    with torch.no_grad():
        tcos = _gpu_cosmo(conf, sigma8=0.85)
        _, _, tgt = _gpu_model(pa.white_noise(1, conf, real=True, device=dev), tcos, conf)
        tgt = tgt.double()
    white0 = pa.white_noise(0, conf, real=True, device=dev)
    cosmo = _gpu_cosmo(conf)
    rng = np.random.default_rng(7)
    us = [torch.as_tensor(rng.standard_normal(conf.ptcl_grid_shape).astype(np.float32),
                          device=dev).double() for _ in range(3)]

    def grads(every):
        w = white0.clone().requires_grad_(True)
        theta = {'Omega_m': torch.tensor(cosmo.Omega_m, dtype=torch.float64, requires_grad=True),
                 'sigma8': torch.tensor(cosmo.sigma8, dtype=torch.float64, requires_grad=True)}
        _, _, dens = _gpu_model(w, cosmo, conf, theta=theta, checkpoint_every=every)
        loss = ((dens.double() - tgt)**2).mean()
        gw, gom, gs8 = torch.autograd.grad(loss, (w, theta['Omega_m'], theta['sigma8']))
        return dict(loss=float(loss.detach()), Omega_m=float(gom), sigma8=float(gs8),
                    dirs=[float((gw.double() * u).sum()) for u in us])

    rev, ck = grads(None), grads(8)
    res = dict(f64=dict(loss=float(z['loss']), Omega_m=float(z['loss_dOm']),
                        sigma8=float(z['loss_ds8']), dirs=[float(x) for x in z['loss_dirs']]),
               reversal=rev, checkpoints_8=ck)
    _record('C2_grad', **res)
    scale = np.abs(z['loss_dirs']).max()
    for g, tol_om in ((rev, 2e-2), (ck, 2e-3)):
        assert abs(g['loss'] - z['loss']) < 1e-5 * z['loss'], res
        assert abs(g['Omega_m'] - z['loss_dOm']) < tol_om * abs(z['loss_dOm']), res
        assert abs(g['sigma8'] - z['loss_ds8']) < 1e-3 * abs(z['loss_ds8']), res
        for an, want in zip(g['dirs'], z['loss_dirs']):
            assert abs(an - want) < 1e-2 * scale, res


def test_c2_drift_checkpoints(dev):
    """reversal_checkpoints on the 128^3-particle / 256^3-mesh configuration
    (63 steps): the relative drift of pure float32 reversal grows toward the
    start (measured: 1.4e-8 one step back from the end, 3.2e-4 at step 1),
    so at tolerance 3e-5 one checkpoint is placed (step 3).  With that one
    stored state the gradients match the float64 fixture nearly as well as
    with seven every 8 steps (Omega_m 1.5e-3 and 1.3e-3 relative, against
    1.1e-2 by pure reversal)."""
    import pmwd_amd as pa
    z = _c2_fixture()
    conf, _ = _confs(2.0, 128, 2, 1 / 64)
    cosmo = _gpu_cosmo(conf)
    white0 = pa.white_noise(0, conf, real=True, device=dev)
    with torch.no_grad():
        ptcl0, _, _ = _gpu_model(white0, cosmo, conf)
        tcos = _gpu_cosmo(conf, sigma8=0.85)
        _, _, tgt = _gpu_model(pa.white_noise(1, conf, real=True, device=dev), tcos, conf)
        tgt = tgt.double()
    nsteps = len(conf.a_nbody) - 1
    none, drift = pa.reversal_checkpoints(ptcl0, cosmo, conf, tol=float('inf'))
    assert none == [] and sorted(drift) == list(range(1, nsteps))
    assert all(np.isfinite(v) for v in drift.values())
    assert drift[1] > drift[nsteps // 2] > drift[nsteps - 1] and drift[1] > 1e-4
    every, _ = pa.reversal_checkpoints(ptcl0, cosmo, conf, tol=0.0, candidates=8)
    assert every == list(range(8, nsteps, 8))
    marks, drift_t = pa.reversal_checkpoints(ptcl0, cosmo, conf, tol=3e-5)
    assert 1 <= len(marks) <= 3 and max(drift_t.values()) < 10 * 3e-5

    w = white0.clone().requires_grad_(True)
    theta = {'Omega_m': torch.tensor(cosmo.Omega_m, dtype=torch.float64, requires_grad=True),
             'sigma8': torch.tensor(cosmo.sigma8, dtype=torch.float64, requires_grad=True)}
    _, _, dens = _gpu_model(w, cosmo, conf, theta=theta, checkpoint_every=marks)
    loss = ((dens.double() - tgt)**2).mean()
    _, gom, gs8 = torch.autograd.grad(loss, (w, theta['Omega_m'], theta['sigma8']))
    res = dict(marks=marks, drift_no_snaps={k: drift[k] for k in (1, 2, 3, 8, 31, 62)},
               grad_Omega_m=float(gom), f64_Omega_m=float(z['loss_dOm']),
               grad_sigma8=float(gs8), f64_sigma8=float(z['loss_ds8']))
    _record('C2_drift_checkpoints', **res)
    assert abs(float(gom) - z['loss_dOm']) < 2.5e-3 * abs(z['loss_dOm']), res
    assert abs(float(gs8) - z['loss_ds8']) < 1e-3 * abs(z['loss_ds8']), res


def _ic_objective(conf, dev, seed_ab=11):
    """J(white, theta) = <a, disp> + <b, vel> of the 2LPT initial conditions
    (linear_modes -> lpt) for fixed standard normal a, b (numpy
    default_rng(seed_ab), shape (N,3)): a polynomial of degree two in the
    white modes and in sigma8, smooth in Omega_m."""
    import pmwd_amd as pa
    rng = np.random.default_rng(seed_ab)
    n = conf.ptcl_num
    a = torch.as_tensor(rng.standard_normal((n, 3)), device=dev)
    b = torch.as_tensor(rng.standard_normal((n, 3)), device=dev)

    def J(white, s8=0.8, Om=0.3, grad=False):
        cosmo = _gpu_cosmo(conf, s8, Om)
        theta = None
        if grad:
            theta = {'Omega_m': torch.tensor(cosmo.Omega_m, dtype=torch.float64,
                                             requires_grad=True),
                     'sigma8': torch.tensor(cosmo.sigma8, dtype=torch.float64, requires_grad=True)}
        ptcl, _ = pa.lpt(pa.linear_modes(white, cosmo, conf, theta=theta), cosmo, conf,
                         theta=theta)
        return (a * ptcl.disp.double()).sum() + (b * ptcl.vel.double()).sum(), theta
    return J


def test_c2_initial_condition_gradients_exact(dev):
    """The gradient of the 2LPT initial conditions (linear_modes -> lpt, the
    part of the C2 chain before the N-body adjoint) with respect to Omega_m,
    sigma8 and along three directions of the white modes, against the
    float64 oracle's central differences (tests/golden/c2_grad.npz): within
    1e-4 for Omega_m, 1e-5 for sigma8 and 1e-5 of the largest directional
    derivative for the modes.  (The GPU's own float32 differences over
    Omega_m carry about 1e-3 of rounding noise.)"""
    import pmwd_amd as pa
    z = _c2_fixture()
    conf, _ = _confs(2.0, 128, 2, 1 / 64)
    J = _ic_objective(conf, dev)
    white0 = pa.white_noise(0, conf, real=True, device=dev)
    w = white0.clone().requires_grad_(True)
    j, theta = J(w, grad=True)
    gw, gom, gs8 = torch.autograd.grad(j, (w, theta['Omega_m'], theta['sigma8']))
    rng = np.random.default_rng(12)
    dirs = []
    for _ in range(3):
        u = torch.as_tensor(rng.standard_normal(conf.ptcl_grid_shape).astype(np.float32),
                            device=dev)
        dirs.append(float((gw.double() * u.double()).sum()))
    res = dict(J=float(j.detach()), J_f64=float(z['ic_J']), grad_sigma8=float(gs8),
               f64_sigma8=float(z['ic_ds8']), grad_Omega_m=float(gom),
               f64_Omega_m=float(z['ic_dOm']), modes_dirs=dirs,
               f64_modes_dirs=[float(x) for x in z['ic_dirs']])
    _record('C2_ic_exact', **res)
    assert abs(res['J'] - res['J_f64']) < 1e-5 * abs(res['J_f64']), res
    assert abs(float(gs8) - z['ic_ds8']) < 1e-5 * abs(z['ic_ds8']), res
    assert abs(float(gom) - z['ic_dOm']) < 1e-4 * abs(z['ic_dOm']), res
    scale = np.abs(z['ic_dirs']).max()
    for an, want in zip(dirs, z['ic_dirs']):
        assert abs(an - want) < 1e-5 * scale, res


def test_c2_adjoint_vs_oracle(dev):
    """The N-body adjoint of the 128^3 / 256^3 configuration against the
    oracle's reverse-time adjoint (nbody.py:226-260) over three steps from
    the C2 initial conditions, with the density-MSE cotangent of the final
    state: cosine similarity of the disp and vel cotangents >= 0.999, and
    the Omega_m cotangent within 1e-3."""
    import pmwd_amd as pa
    conf, oconf = _confs(2.0, 128, 2, 1 / 64, a_stop=4 / 64)
    cosmo = _gpu_cosmo(conf)
    ocosmo = _oracle_cosmo(oconf)
    white = pa.white_noise(0, conf, real=True, device=dev)
    with torch.no_grad():
        ptcl0, _ = pa.lpt(pa.linear_modes(white, cosmo, conf), cosmo, conf)
    fin, _ = pa.nbody(ptcl0, None, cosmo, conf)
    # cotangent of mean((rho - 1)^2) at the final state: the scatter adjoint
    # of 2 (rho - 1) / M, taken once (oracle) and fed to both sides
    pmid = ptcl0.pmid.cpu().numpy()
    d1, v1 = fin.disp.cpu().numpy(), fin.vel.cpu().numpy()
    rho = opm.scatter(pmid, d1, oconf.cell_size, oconf.mesh_shape)
    mc = (2 * (rho - 1) / rho.size).astype(np.float32)
    dcot, _ = opm.scatter_adj(pmid, d1, oconf.cell_size, oconf.mesh_shape, mc)
    vcot = np.zeros_like(dcot)
    P = lambda a: torch.as_tensor(a, device=dev)
    _, pc, cc = pa.nbody_adj(fin, pa.Particles(conf, fin.pmid, P(dcot), P(vcot)), cosmo, conf)
    od0, ov0 = ptcl0.disp.cpu().numpy(), ptcl0.vel.cpu().numpy()
    wd, wv, wa = onb.nbody(pmid, od0, ov0, ocosmo, oconf)
    _, (wdc, wvc, _), wom = onb.nbody_adj(pmid, wd, wv, wa, dcot, vcot, ocosmo, oconf)
    r = dict(cos_disp_cot=cosine(pc.disp.cpu().numpy(), wdc),
             cos_vel_cot=cosine(pc.vel.cpu().numpy(), wvc),
             rel_disp_cot=rel_l2(pc.disp.cpu().numpy(), wdc),
             rel_vel_cot=rel_l2(pc.vel.cpu().numpy(), wvc),
             Omega_m_cot=cc.Omega_m, oracle_Omega_m_cot=wom,
             final_disp_vs_oracle=rel_l2(d1, wd))
    _record('C2_adjoint', **r)
    assert r['cos_disp_cot'] >= 0.999 and r['cos_vel_cot'] >= 0.999, r
    assert abs(cc.Omega_m - wom) < 1e-3 * abs(wom), r


def test_c5_size_classes_alone_and_in_the_mix(dev):
    """One request of each size class of the serving mix (64^3/128^3,
    128^3/256^3, 256^3/512^3 particles/mesh, 2 Mpc/h spacing, 31 steps), with
    and without gradients: each result is bitwise the same alone and in the
    mix (the float64 Omega_m sum within 1e-9)."""
    from pmwd_amd.serving import Server, SimRequest
    reqs = [SimRequest(seed=31, ptcl_grid=64, mesh_ratio=2, box=128.0, Omega_m=0.27,
                       sigma8=0.75, grad=True),
            SimRequest(seed=32, ptcl_grid=128, mesh_ratio=2, box=256.0, Omega_m=0.33,
                       sigma8=0.86),
            SimRequest(seed=33, ptcl_grid=256, mesh_ratio=2, box=512.0, Omega_m=0.30,
                       sigma8=0.80, grad=True),
            SimRequest(seed=34, ptcl_grid=128, mesh_ratio=2, box=256.0, Omega_m=0.29,
                       sigma8=0.72, grad=True)]
    solo = Server(dev, max_slots=1)
    alone = [solo.submit(r).result() for r in reqs]
    solo.close()
    mix = Server(dev, max_slots=len(reqs))
    got = [f.result() for f in [mix.submit(r) for r in reqs]]
    mix.close()
    assert mix.peak_active > 1
    for a, m in zip(alone, got):
        np.testing.assert_array_equal(a.dens, m.dens)
        assert abs(a.dens.astype(np.float64).mean() - 1) < 1e-6
        if a.request.grad:
            assert a.loss == m.loss
            np.testing.assert_array_equal(a.grad_modes, m.grad_modes)
            assert abs(a.grad_Omega_m - m.grad_Omega_m) <= 1e-9 * abs(a.grad_Omega_m)
            assert abs(a.grad_sigma8 - m.grad_sigma8) <= 1e-9 * abs(a.grad_sigma8)
    _record('C5_classes', n=len(reqs), peak_active=mix.peak_active,
            latencies_alone=[r.latency for r in alone], latencies_mix=[r.latency for r in got])


def _full_adjoint_run(ckpt):
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cmd = [sys.executable, os.path.join(root, 'bench.py'), '--full-adjoint', '1024',
           '--full-adjoint-checkpoints', str(ckpt)]
    print('running', ' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=500, cwd=root)
    assert r.returncode == 0, r.stderr[-3000:]
    return json.loads(r.stdout.strip().splitlines()[-1])['full_adjoint']


def test_c4_checkpointed_gradient_equals_pure_reversal(dev):
    """1024^3 particles on a 1024^3 mesh, box 2000 Mpc/h, 31 steps, density
    MSE against a seeded target: the gradients with the reverse run snapped
    to states stored every 8 steps equal those of pure reversal within 1e-4
    (Omega_m, sigma8, the modes gradient's norm) and 1e-3 of the largest of
    three projections of the modes gradient (measured: 2.9e-5, 4.9e-7,
    1.6e-6 and 1.8e-4; the float32 reversal does not retrace the forward
    exactly, which the stored states correct), all finite; peak device
    memory is recorded.  Each run is a process of its own
    (bench.py --full-adjoint), as the model fills most of the device."""
    from pmwd_amd import _lib
    _lib.clear_contexts()
    torch.cuda.empty_cache()
    ck = _full_adjoint_run(8)
    pure = _full_adjoint_run(0)
    rel = lambda a, b: abs(a - b) / abs(b)
    res = dict(ckpt8={k: ck[k] for k in ('t_forward_s', 't_backward_s', 'grad_Omega_m',
                                         'grad_sigma8', 'grad_modes_norm', 'grad_modes_proj')},
               pure={k: pure[k] for k in ('t_forward_s', 't_backward_s', 'grad_Omega_m',
                                          'grad_sigma8', 'grad_modes_norm', 'grad_modes_proj')},
               memory_ckpt8=ck['memory'], memory_pure=pure['memory'])
    _record('C4_1024', **res)
    for k in ('grad_Omega_m', 'grad_sigma8', 'grad_modes_norm'):
        assert np.isfinite(ck[k]) and np.isfinite(pure[k])
        assert rel(ck[k], pure[k]) < 1e-4, (k, res)
    scale = max(abs(x) for x in pure['grad_modes_proj'])
    for a, b in zip(ck['grad_modes_proj'], pure['grad_modes_proj']):
        assert abs(a - b) < 1e-3 * scale, res
