"""GPU parity of gravity, the KDK integrator and the reverse-time adjoint."""
import numpy as np
import pytest
import torch

from oracle import cosmo as oc, nbody as onb, pm as opm
from oracle.conf import OConf

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make_ptcl(shape, mesh_shape, cell, std, seed):
    pmid, disp = opm.gen_grid(shape, mesh_shape, cell)
    rng = np.random.default_rng(seed)
    return pmid, (disp + std * rng.standard_normal(disp.shape)).astype(np.float32)


@pytest.mark.parametrize('shape,ratio,std', [((16, 16, 16), 2, 1.5), ((4, 9, 8), 1, 7.0),
                                             ((7, 8, 6), 1, 7.0), ((32, 32, 32), 1, 0.5)])
def test_gravity(dev, shape, ratio, std):
    import pmwd_amd as pa
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    cosmo = pa.SimpleLCDM(conf)
    pmid, disp = make_ptcl(shape, conf.mesh_shape, conf.cell_size, std, 1)
    ptcl = pa.Particles(conf, torch.as_tensor(pmid, device=dev), torch.as_tensor(disp, device=dev))
    acc = pa.gravity(None, ptcl, cosmo, conf).cpu().numpy()
    want = opm.gravity(pmid, disp, 0.3, conf.cell_size, conf.mesh_shape)
    assert rel_l2(acc, want) < 2e-5


@pytest.mark.parametrize('spacing', [1.953125, 1.7, 0.7, 3.0, 2.0])
def test_gravity_tile_path_matches_generic_ops(dev, spacing):
    """pmx_gravity (binning, tile deposit, Poisson, tile gather) equals, bit
    for bit, pmx_scatter_mass + pmx_poisson_grad + the generic gather on the
    same particles, for power-of-two and other cell sizes.  Displacements
    spread over binades from 2^-70 to 2^8, with zeros of both signs."""
    import pmwd_amd as pa
    from pmwd_amd import _lib
    L = _lib.lib()
    shape = (16, 16, 16)
    conf = pa.Configuration(spacing, shape, mesh_shape=2)
    pmid, disp = opm.gen_grid(shape, conf.mesh_shape, conf.cell_size)
    rng = np.random.default_rng(41)
    mag = 2.0 ** rng.uniform(-70, 8, size=disp.shape)
    d = (np.where(rng.random(disp.shape) < 0.5, -1.0, 1.0) * mag).astype(np.float32)
    d[::7] = 0.0
    d[1::11] = -0.0
    n = len(pmid)
    c = _lib.context(conf.mesh_shape, conf.cell_size, dev)
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    pm, x = P(pmid), P(d)
    val = float(np.float32(conf.mesh_size / conf.ptcl_num))
    coef = float(np.float32(0.45))
    acc = torch.empty_like(x)
    _lib.check(L.pmx_gravity(c.bind(), _lib.ptr(pm), _lib.ptr(x), n, val, coef, _lib.ptr(acc)))
    dens = torch.zeros(conf.mesh_shape, dtype=torch.float32, device=dev)
    _lib.check(L.pmx_scatter_mass(c.bind(), _lib.ptr(pm), _lib.ptr(x), n, val, _lib.ptr(dens), 0))
    F = torch.empty((3,) + tuple(conf.mesh_shape), dtype=torch.float32, device=dev)
    _lib.check(L.pmx_poisson_grad(c.bind(), _lib.ptr(dens), coef, _lib.ptr(F)))
    ptcl = pa.Particles(conf, pm, x)
    want = torch.stack([pa.gather(ptcl, conf, F[a]) for a in range(3)], dim=1)
    np.testing.assert_array_equal(acc.cpu().numpy(), want.cpu().numpy())


@pytest.mark.parametrize('shape,ratio', [((8, 8, 8), 2), ((4, 9, 8), 1), ((7, 8, 6), 1)])
def test_gravity_vjp(dev, shape, ratio):
    import pmwd_amd as pa
    from pmwd_amd import _lib
    import ctypes
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    pmid, disp = make_ptcl(shape, conf.mesh_shape, conf.cell_size, 2.0, 2)
    vc = np.random.default_rng(5).standard_normal(disp.shape).astype(np.float32)
    c = _lib.context(conf.mesh_shape, conf.cell_size, dev)
    P = lambda a: torch.as_tensor(a, device=dev)
    acc = torch.empty_like(P(disp))
    ac = torch.empty_like(P(disp))
    dot = ctypes.c_double()
    coef = float(np.float32(1.5) * np.float32(0.3))
    val = float(np.float32(conf.mesh_size / conf.ptcl_num))
    pm_t, d_t, v_t = P(pmid), P(disp), P(vc)
    _lib.check(_lib.lib().pmx_gravity_vjp(c.bind(), _lib.ptr(pm_t), _lib.ptr(d_t), len(pmid), val,
                                          coef, _lib.ptr(v_t), _lib.ptr(acc), _lib.ptr(ac),
                                          ctypes.byref(dot)))
    wacc, wdc, wom = opm.gravity_vjp(pmid, disp, 0.3, conf.cell_size, conf.mesh_shape, vc)
    assert rel_l2(acc.cpu().numpy(), wacc) < 2e-5
    assert rel_l2(ac.cpu().numpy(), wdc) < 1e-4
    assert abs(dot.value / 0.3 - wom) < 1e-4 * abs(wom)


def test_gravity_vjp_finite_difference(dev):
    """<v, J dx> from central differences of the GPU gravity matches <J^T v, dx>."""
    import pmwd_amd as pa
    conf = pa.Configuration(1.0, (8, 8, 8), mesh_shape=2)
    cosmo = pa.SimpleLCDM(conf)
    pmid, disp = make_ptcl((8, 8, 8), conf.mesh_shape, conf.cell_size, 1.0, 3)
    rng = np.random.default_rng(6)
    v = rng.standard_normal(disp.shape)
    dx = rng.standard_normal(disp.shape)
    P = lambda a: torch.as_tensor(a, device=dev)

    def g(x):
        return pa.gravity(None, pa.Particles(conf, P(pmid), P(x.astype(np.float32))), cosmo,
                          conf).cpu().numpy().astype(np.float64)
    # small enough that few particles cross a CIC kink, large enough for f32
    eps = 1e-3
    fd = (np.sum(v * g(disp + eps * dx)) - np.sum(v * g(disp - eps * dx))) / (2 * eps)
    d = P(disp).requires_grad_(True)
    acc = pa.gravity(None, pa.Particles(conf, P(pmid), d), cosmo, conf)
    (gd,) = torch.autograd.grad(acc, d, P(v.astype(np.float32)))
    an = np.sum(gd.cpu().numpy() * dx)
    assert abs(fd - an) < 2e-2 * abs(an)


def small_setup(shape=(16, 16, 16), ratio=2, a_start=1 / 16, steps_max=1 / 16):
    import pmwd_amd as pa
    conf = pa.Configuration(4.0, shape, mesh_shape=ratio, a_start=a_start,
                            a_nbody_maxstep=steps_max)
    oconf = OConf(4.0, shape, mesh_shape=ratio, a_start=a_start, a_nbody_maxstep=steps_max)
    cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
    ocosmo = oc.boltzmann(oc.SimpleLCDM(oconf), oconf)
    return pa, conf, oconf, cosmo, ocosmo


def test_nbody_forward(dev):
    pa, conf, oconf, cosmo, ocosmo = small_setup()
    pmid, disp = make_ptcl(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, 0.5, 4)
    vel = (0.05 * np.random.default_rng(7).standard_normal(disp.shape)).astype(np.float32)
    P = lambda a: torch.as_tensor(a, device=dev)
    ptcl = pa.Particles(conf, P(pmid), P(disp), P(vel))
    out, _ = pa.nbody(ptcl, None, cosmo, conf)
    wd, wv, wa = onb.nbody(pmid, disp, vel, ocosmo, oconf)
    assert rel_l2(out.disp.cpu().numpy(), wd) < 1e-5
    assert rel_l2(out.vel.cpu().numpy(), wv) < 1e-4
    assert rel_l2(out.acc.cpu().numpy(), wa) < 1e-4


def test_nbody_reversibility(dev):
    """Forward then reverse returns to the start (adjoint.tex:1529-1549 tolerances)."""
    pa, conf, oconf, cosmo, ocosmo = small_setup()
    pmid, disp = make_ptcl(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, 0.5, 5)
    vel = (0.05 * np.random.default_rng(8).standard_normal(disp.shape)).astype(np.float32)
    P = lambda a: torch.as_tensor(a, device=dev)
    ptcl = pa.Particles(conf, P(pmid), P(disp), P(vel))
    fwd, _ = pa.nbody(ptcl, None, cosmo, conf)
    back, _ = pa.nbody(fwd, None, cosmo, conf, reverse=True)
    rd = np.sqrt(np.mean((back.disp.cpu().numpy() - disp)**2)) / np.std(disp - opm.gen_grid(
        conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size)[1])
    assert rd < 1e-3


def test_nbody_adjoint_vs_oracle(dev):
    pa, conf, oconf, cosmo, ocosmo = small_setup(shape=(8, 8, 8), a_start=1 / 4, steps_max=1 / 4)
    pmid, disp = make_ptcl(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, 0.5, 6)
    vel = (0.05 * np.random.default_rng(9).standard_normal(disp.shape)).astype(np.float32)
    rng = np.random.default_rng(10)
    dcot = rng.standard_normal(disp.shape).astype(np.float32)
    vcot = rng.standard_normal(disp.shape).astype(np.float32)
    P = lambda a: torch.as_tensor(a, device=dev)
    ptcl = pa.Particles(conf, P(pmid), P(disp), P(vel))
    fin, _ = pa.nbody(ptcl, None, cosmo, conf)
    p0, pc, cc = pa.nbody_adj(fin, pa.Particles(conf, P(pmid), P(dcot), P(vcot)), cosmo, conf)
    wd, wv, wa = onb.nbody(pmid, disp, vel, ocosmo, oconf)
    (r0d, r0v, _), (wdc, wvc, _), wom = onb.nbody_adj(pmid, wd, wv, wa, dcot, vcot, ocosmo, oconf)
    assert rel_l2(pc.disp.cpu().numpy(), wdc) < 1e-3
    assert rel_l2(pc.vel.cpu().numpy(), wvc) < 1e-3
    assert abs(cc.Omega_m - wom) < 1e-3 * abs(wom)
    assert rel_l2(p0.disp.cpu().numpy(), disp) < 1e-4


@pytest.mark.parametrize('de', [dict(w_0_=-0.9), dict(Omega_k_=0.03, w_a_=0.2)])
def test_nbody_adjoint_dark_energy_cotangents(dev, de):
    """nbody_adj's cotangents of the set parameters among Omega_k_, w_0_,
    w_a_ (nbody.py:267-274: value_and_grad of the factors w.r.t. the whole
    cosmology) against central differences of the oracle's FP64 forward run,
    growth table re-integrated; unset parameters stay None and A_s, n_s,
    Omega_b, h are 0."""
    pa, conf, oconf, _, _ = small_setup(shape=(8, 8, 8), a_start=1 / 4, steps_max=1 / 8)
    cosmo = pa.boltzmann(pa.SimpleLCDM(conf, **de), conf)
    pmid, disp = make_ptcl(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, 0.5, 6)
    vel = (0.05 * np.random.default_rng(9).standard_normal(disp.shape)).astype(np.float32)
    rng = np.random.default_rng(10)
    dcot = rng.standard_normal(disp.shape).astype(np.float32)
    vcot = rng.standard_normal(disp.shape).astype(np.float32)
    P = lambda a: torch.as_tensor(a, device=dev)
    fin, _ = pa.nbody(pa.Particles(conf, P(pmid), P(disp), P(vel)), None, cosmo, conf)
    _, _, cc = pa.nbody_adj(fin, pa.Particles(conf, P(pmid), P(dcot), P(vcot)), cosmo, conf)
    tight = oconf.replace(growth_rtol=1e-12, growth_atol=1e-12)

    def J(**kw):
        c = oc.boltzmann(oc.SimpleLCDM(tight, **{**de, **kw}), tight)
        d1, v1, _ = onb.nbody(pmid, disp, vel, c, tight, dtype=np.float64)
        return np.sum(dcot.astype(np.float64) * d1) + np.sum(vcot.astype(np.float64) * v1)
    h = 1e-5
    for p in ('Omega_k_', 'w_0_', 'w_a_'):
        if p not in de:
            assert getattr(cc, p) is None
            continue
        fd = (J(**{p: de[p] + h}) - J(**{p: de[p] - h})) / (2 * h)
        assert abs(getattr(cc, p) - fd) < 2e-3 * abs(fd), (p, getattr(cc, p), fd)
    fd = (J(Omega_m=0.3 + h) - J(Omega_m=0.3 - h)) / (2 * h)
    assert abs(cc.Omega_m - fd) < 2e-3 * abs(fd), (cc.Omega_m, fd)
    assert (cc.A_s_1e9, cc.n_s, cc.Omega_b, cc.h) == (0.0, 0.0, 0.0, 0.0)


@pytest.mark.parametrize('splits', [((0.5, 1), (0.5, 0)), ((0, 0.25), (0.5, 0.5), (0.5, 0.25))])
def test_nbody_other_symplectic_splits_vs_oracle(dev, splits):
    """integrate with other symp_splits (nbody.py:121-140,
    configuration.py:144, 187-191) -- DKD and a three-stage splitting -- run op
    by op on the GPU: the forward against the oracle (disp 1e-5, vel and acc
    1e-4), the reverse-time adjoint against the oracle's (1e-3, and the
    Omega_m cotangent), and forward-then-reverse back to the start."""
    import pmwd_amd as pa
    shape = (16, 16, 16)
    kw = dict(mesh_shape=2, a_start=1 / 8, a_nbody_maxstep=1 / 8, symp_splits=splits)
    conf = pa.Configuration(4.0, shape, **kw)
    oconf = OConf(4.0, shape, **kw)
    cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
    ocosmo = oc.boltzmann(oc.SimpleLCDM(oconf), oconf)
    pmid, disp = make_ptcl(shape, conf.mesh_shape, conf.cell_size, 0.5, 21)
    vel = (0.05 * np.random.default_rng(22).standard_normal(disp.shape)).astype(np.float32)
    P = lambda a: torch.as_tensor(a, device=dev)
    out, _ = pa.nbody(pa.Particles(conf, P(pmid), P(disp), P(vel)), None, cosmo, conf)
    wd, wv, wa = onb.nbody(pmid, disp, vel, ocosmo, oconf)
    assert rel_l2(out.disp.cpu().numpy(), wd) < 1e-5
    assert rel_l2(out.vel.cpu().numpy(), wv) < 1e-4
    assert rel_l2(out.acc.cpu().numpy(), wa) < 1e-4
    rng = np.random.default_rng(23)
    dcot = rng.standard_normal(disp.shape).astype(np.float32)
    vcot = rng.standard_normal(disp.shape).astype(np.float32)
    p0, pc, cc = pa.nbody_adj(out, pa.Particles(conf, P(pmid), P(dcot), P(vcot)), cosmo, conf)
    (r0d, _, _), (wdc, wvc, _), wom = onb.nbody_adj(pmid, wd, wv, wa, dcot, vcot, ocosmo, oconf)
    assert rel_l2(pc.disp.cpu().numpy(), wdc) < 1e-3
    assert rel_l2(pc.vel.cpu().numpy(), wvc) < 1e-3
    assert abs(cc.Omega_m - wom) < 1e-3 * abs(wom)
    assert rel_l2(p0.disp.cpu().numpy(), disp) < 1e-4


def test_nbody_autograd_matches_adjoint(dev):
    pa, conf, oconf, cosmo, ocosmo = small_setup(shape=(8, 8, 8), a_start=1 / 4, steps_max=1 / 4)
    pmid, disp = make_ptcl(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, 0.5, 6)
    P = lambda a: torch.as_tensor(a, device=dev)
    d = P(disp).requires_grad_(True)
    v = torch.zeros_like(d).requires_grad_(True)
    om = torch.tensor(cosmo.Omega_m, dtype=torch.float64, requires_grad=True)
    out, _ = pa.nbody(pa.Particles(conf, P(pmid), d, v), None, cosmo, conf, omega_m=om)
    dens = pa.scatter(out, conf)
    loss = (dens**2).mean()
    gd, gv, gom = torch.autograd.grad(loss, (d, v, om))
    assert torch.isfinite(gd).all() and torch.isfinite(gv).all()
    # Omega_m derivative by central differences of the whole pipeline
    vals = []
    for s in (1, -1):
        c2 = pa.boltzmann(pa.SimpleLCDM(conf, Omega_m=0.3 + s * 1e-3), conf)
        o2, _ = pa.nbody(pa.Particles(conf, P(pmid), P(disp), torch.zeros_like(P(disp))), None,
                         c2, conf)
        vals.append(float((pa.scatter(o2, conf).double()**2).mean()))
    fd = (vals[0] - vals[1]) / 2e-3
    assert abs(float(gom) - fd) < 0.02 * abs(fd) + 1e-6


@pytest.mark.parametrize('shape,ratio,snap', [((16, 16, 16), 2, False), ((12, 10, 8), 1, True),
                                              ((32, 32, 32), 1, False)])
def test_nbody_step_adj_matches_separate_ops(dev, shape, ratio, snap):
    """Two reverse KDK steps of pmx_nbody_step_adj (the first step's closing
    kick deferred into the second call, the last applied by pmx_kick_adj)
    equal kick_adj, drift_adj, gravity_vjp, kick_adj called one by one
    (nbody.py:143-162): bitwise for every array, and the dots match in
    float64."""
    import ctypes
    import pmwd_amd as pa
    from pmwd_amd import _lib
    L = _lib.lib()
    conf = pa.Configuration(1.0, shape, mesh_shape=ratio)
    c = _lib.context(conf.mesh_shape, conf.cell_size, dev)
    h = c.bind()
    pmid, disp = make_ptcl(shape, conf.mesh_shape, conf.cell_size, 2.5, 11)
    rng = np.random.default_rng(12)
    f32 = lambda: rng.standard_normal(disp.shape).astype(np.float32)
    P = lambda a: torch.as_tensor(a, device=dev).contiguous()
    vel, dcot, vcot = f32(), f32(), f32()
    snap_t = P(disp + f32()) if snap else None
    n = len(pmid)
    val = float(np.float32(conf.mesh_size / conf.ptcl_num))
    coef = float(np.float32(0.45))
    steps = [(0.37, -0.81, 0.29), (0.21, -0.64, 0.33)]
    dot = ctypes.c_double()

    def start():
        st = dict(pmid=P(pmid), disp=P(disp), vel=P(vel), dc=P(dcot), vc=P(vcot))
        st['acc'] = torch.empty_like(st['disp'])
        st['ac'] = torch.empty_like(st['disp'])
        _lib.check(L.pmx_gravity_vjp(h, _lib.ptr(st['pmid']), _lib.ptr(st['disp']), n, val, coef,
                                     _lib.ptr(st['vc']), _lib.ptr(st['acc']), _lib.ptr(st['ac']),
                                     ctypes.byref(dot)))
        return st, dot.value

    a, dot0 = start()
    p = lambda k: _lib.ptr(a[k])
    want_dots, kick_dots = [], []
    for i, (ka, d, kb) in enumerate(steps):
        _lib.check(L.pmx_kick_adj(h, n, p('vel'), p('acc'), p('dc'), p('ac'), p('vc'), ka,
                                  ctypes.byref(dot)))
        kick_dots.append(dot.value)
        _lib.check(L.pmx_drift_adj(h, n, p('disp'), p('vel'), p('dc'), p('vc'), d,
                                   ctypes.byref(dot)))
        ddot = dot.value
        if snap and i == 0:
            a['disp'].copy_(snap_t)
        _lib.check(L.pmx_gravity_vjp(h, p('pmid'), p('disp'), n, val, coef, p('vc'), p('acc'),
                                     p('ac'), ctypes.byref(dot)))
        want_dots.append((ddot, dot.value))
        _lib.check(L.pmx_kick_adj(h, n, p('vel'), p('acc'), p('dc'), p('ac'), p('vc'), kb,
                                  ctypes.byref(dot)))
        kick_dots.append(dot.value)

    b, _ = start()
    q = lambda k: _lib.ptr(b[k])
    dots = (ctypes.c_double * 2)()
    kp = 0.0
    for i, (ka, d, kb) in enumerate(steps):
        _lib.check(L.pmx_nbody_step_adj(h, q('pmid'), q('disp'), q('vel'), q('acc'), q('dc'),
                                        q('vc'), q('ac'), n, val, coef, kp, ka, d,
                                        _lib.ptr(snap_t) if i == 0 else None, dots))
        for got, want in zip(dots, want_dots[i]):
            assert abs(got - want) <= 1e-12 * abs(want)
        kp = kb
    _lib.check(L.pmx_kick_adj(h, n, q('vel'), q('acc'), q('dc'), q('ac'), q('vc'), kp, None))
    for k in ('disp', 'vel', 'acc', 'dc', 'vc', 'ac'):
        np.testing.assert_array_equal(b[k].cpu().numpy(), a[k].cpu().numpy(), err_msg=k)
    if snap:
        # the snapshot copied into disp first and disp passed as its own
        # snapshot (no second array): the same bits
        e, _ = start()
        r = lambda k: _lib.ptr(e[k])
        kp = 0.0
        for i, (ka, d, kb) in enumerate(steps):
            if i == 0:
                e['disp'].copy_(snap_t)
            _lib.check(L.pmx_nbody_step_adj(h, r('pmid'), r('disp'), r('vel'), r('acc'), r('dc'),
                                            r('vc'), r('ac'), n, val, coef, kp, ka, d,
                                            r('disp') if i == 0 else None, None))
            kp = kb
        _lib.check(L.pmx_kick_adj(h, n, r('vel'), r('acc'), r('dc'), r('ac'), r('vc'), kp, None))
        for k in ('disp', 'vel', 'acc', 'dc', 'vc', 'ac'):
            np.testing.assert_array_equal(e[k].cpu().numpy(), a[k].cpu().numpy(), err_msg=k)
    # every kick adjoint's dot is the dot of the force it uses
    force_dots = [dot0, want_dots[0][1], want_dots[0][1], want_dots[1][1]]
    for kd, fd in zip(kick_dots, force_dots):
        assert abs(kd - fd) <= 1e-12 * abs(fd)


@pytest.mark.parametrize('every', [None, 1, 2])
def test_adjoint_driver_fused_matches_op_by_op(dev, every):
    """nbody's reverse-time adjoint with the fused KDK call (pending closing
    kicks, checkpoint snaps) equals the op-by-op loop bit for bit."""
    import importlib
    nbm = importlib.import_module('pmwd_amd.nbody')
    pa_, conf, oconf, cosmo, ocosmo = small_setup(shape=(8, 8, 8), a_start=1 / 8, steps_max=1 / 8)
    pmid, disp = make_ptcl(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, 0.5, 21)
    P = lambda a: torch.as_tensor(a, device=dev).contiguous()
    rng = np.random.default_rng(22)
    d0, v0 = P(disp), P(0.1 * rng.standard_normal(disp.shape).astype(np.float32))
    dc0 = P(rng.standard_normal(disp.shape).astype(np.float32))
    vc0 = P(rng.standard_normal(disp.shape).astype(np.float32))
    pm = P(pmid)
    d, v = d0.clone(), v0.clone()
    _, ckpt = nbm._run_forward(pm, d, v, cosmo, conf, checkpoint_every=every)
    outs = []
    for fused in (True, False):
        st = [d.clone(), v.clone(), dc0.clone(), vc0.clone()]
        acc, acc_cot, om = nbm._run_adjoint(pm, *st, cosmo, conf, ckpt=ckpt, fused=fused)
        outs.append((st + [acc, acc_cot], om))
    for x, y in zip(outs[0][0], outs[1][0]):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
    assert abs(outs[0][1] - outs[1][1]) <= 1e-10 * abs(outs[1][1]) + 1e-10


def test_host_checkpoints_match_device_checkpoints(dev):
    """Reversal checkpoints kept in pinned host memory give the same adjoint,
    bit for bit, as checkpoints kept on the device."""
    import importlib
    nbm = importlib.import_module('pmwd_amd.nbody')
    pa_, conf, oconf, cosmo, ocosmo = small_setup(shape=(8, 8, 8), a_start=1 / 8, steps_max=1 / 8)
    pmid, disp = make_ptcl(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, 0.5, 31)
    P = lambda a: torch.as_tensor(a, device=dev).contiguous()
    rng = np.random.default_rng(32)
    v0 = P(0.1 * rng.standard_normal(disp.shape).astype(np.float32))
    dc0 = P(rng.standard_normal(disp.shape).astype(np.float32))
    vc0 = P(rng.standard_normal(disp.shape).astype(np.float32))
    outs = []
    for where in (None, 'cpu'):
        d, v = P(disp), v0.clone()
        _, ckpt = nbm._run_forward(P(pmid), d, v, cosmo, conf, checkpoint_every=2,
                                   checkpoint_device=where)
        assert ckpt and all(t.device.type == ('cpu' if where else 'cuda') for t in ckpt[2])
        st = [d, v, dc0.clone(), vc0.clone()]
        acc, acc_cot, om = nbm._run_adjoint(P(pmid), *st, cosmo, conf, ckpt=ckpt)
        outs.append([t.cpu().numpy() for t in st + [acc, acc_cot]] + [om])
    for x, y in zip(outs[0][:-1], outs[1][:-1]):
        np.testing.assert_array_equal(x, y)
    assert abs(outs[0][-1] - outs[1][-1]) <= 1e-10 * abs(outs[1][-1]) + 1e-10
