"""Pins of the CPU oracle: known answers, closed forms, finite differences,
brute force and independent library routines (no GPU)."""
import numpy as np
import pytest
import scipy.fft as sfft
from scipy.integrate import quad, solve_ivp

from oracle import cosmo as oc, ic, nbody as onb, pm as opm
from oracle.conf import OConf


def ptcl(shape, mesh_shape, cell, std, seed, dtype=np.float32):
    pmid, disp = opm.gen_grid(shape, mesh_shape, cell, dtype=dtype)
    rng = np.random.default_rng(seed)
    return pmid, (disp + std * rng.standard_normal(disp.shape)).astype(dtype)


# ------------------------------------------------------------- enmesh ---

def brute_enmesh(pmid, disp, cell, shape):
    """Independent per-particle CIC: Python floats, explicit floor and mod."""
    out_i, out_w = [], []
    for p, d in zip(pmid, disp):
        ii, ww = [], []
        for j in range(2**len(shape)):
            idx, w = [], 1.0
            for a, n in enumerate(shape):
                x = np.float32(np.float32(d[a]) / np.float32(cell))
                lo = int(np.floor(x))
                nb = (j >> a) & 1
                frac = np.float32(1) - abs(np.float32(x - np.float32(lo + nb)))
                w = np.float32(w * frac) if a else frac
                idx.append((lo + nb + int(p[a])) % n)
            ii.append(idx)
            ww.append(w)
        out_i.append(ii)
        out_w.append(ww)
    return np.array(out_i), np.array(out_w, dtype=np.float32)


@pytest.mark.parametrize('shape', [(4, 9, 7), (8, 8, 8), (3, 5, 2)])
def test_enmesh_brute_force(shape):
    pmid, disp = ptcl(shape, shape, 1.5, 6.0, 0)
    i, w = opm.enmesh(pmid, disp, 1.5, shape, 0, None, shape)
    bi, bw = brute_enmesh(pmid, disp, 1.5, shape)
    np.testing.assert_array_equal(i, bi)
    np.testing.assert_array_equal(w, bw)


def test_enmesh_special_cases():
    shape = (4, 4, 4)
    pmid = np.array([[0, 0, 0], [3, 3, 3], [1, 2, 3]], np.int16)
    disp = np.zeros((3, 3), np.float32)
    i, w = opm.enmesh(pmid, disp, 2.0, shape, 0, None, shape)
    np.testing.assert_array_equal(w[:, 0], 1)      # zero disp: own cell only
    np.testing.assert_array_equal(w[:, 1:], 0)
    np.testing.assert_array_equal(i[1, 7], [0, 0, 0])  # wrap at the box edge
    disp = np.array([[-0.5, 0, 0]], np.float32)     # negative disp: floor goes down
    i, w = opm.enmesh(pmid[:1], disp, 1.0, shape, 0, None, shape)
    np.testing.assert_array_equal(i[0, 0], [3, 0, 0])
    np.testing.assert_allclose(w[0, [0, 1]], [0.5, 0.5])


@pytest.mark.parametrize('pos,n', [((-1.,), 3), ((1., -3.), 5), ((-3., 5., 7.), 7),
                                   ((3., -5., 7.), 7), ((3., 5., -7.), 7)])
def test_scatter_centered_ptcl(pos, n):
    """pm_test.py:43-71: odd offsets on a 2^d mesh of cell 2 give n 2^-d exactly."""
    d = len(pos)
    pmid = np.zeros((n, d), np.int16)
    disp = np.tile(np.array(pos, np.float32), (n, 1))
    m = opm.scatter(pmid, disp, 2.0, (2,) * d, val=1.0)
    np.testing.assert_array_equal(m, np.full((2,) * d, n * 2.0**-d, np.float32))


@pytest.mark.parametrize('n,d', [(3, 1), (5, 2), (7, 3)])
def test_scatter_sum_and_gather_uniform(n, d):
    """pm_test.py:86-118: mass conservation and a uniform mesh gathering to 1."""
    shape = (3,) * d
    pmid, disp = ptcl((n,) + (1,) * (d - 1), shape, 2.0, 7.0, 1)
    m = opm.scatter(pmid, disp, 2.0, shape, val=1.0)
    assert abs(m.astype(np.float64).sum() - n) < 1e-5
    g = opm.gather(pmid, disp, 2.0, shape, np.ones(shape, np.float32))
    np.testing.assert_allclose(g, 1.0, atol=3e-7)


def test_scatter_gather_transpose():
    """<scatter(v), m> = <v, gather(m)> in float64."""
    shape = (6, 5, 7)
    pmid, disp = ptcl(shape, shape, 1.0, 3.0, 2, np.float64)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(len(pmid))
    m = rng.standard_normal(shape)
    s = opm.scatter(pmid, disp, 1.0, shape, val=v[:, None], mesh=np.zeros(shape + (1,)),
                    dtype=np.float64)[..., 0]
    g = opm.gather(pmid, disp, 1.0, shape, m, dtype=np.float64)
    assert abs(np.sum(s * m) - np.sum(v * g)) < 1e-10 * np.sum(np.abs(v * g))


def _fd(f, x, dx, eps=1e-6):
    return (f(x + eps * dx) - f(x - eps * dx)) / (2 * eps)


def test_scatter_adj_finite_difference():
    shape = (5, 6, 4)
    pmid, disp = ptcl(shape, shape, 1.0, 2.0, 4, np.float64)
    rng = np.random.default_rng(5)
    mc = rng.standard_normal(shape)
    dx = rng.standard_normal(disp.shape)
    f = lambda x: np.sum(mc * opm.scatter(pmid, x, 1.0, shape, dtype=np.float64))
    dc, vc = opm.scatter_adj(pmid, disp, 1.0, shape, mc, dtype=np.float64)
    assert abs(_fd(f, disp, dx) - np.sum(dc * dx)) < 1e-6 * np.abs(dc * dx).sum()


def test_gather_adj_finite_difference():
    shape = (5, 6, 4)
    pmid, disp = ptcl(shape, shape, 1.0, 2.0, 6, np.float64)
    rng = np.random.default_rng(7)
    m = rng.standard_normal(shape)
    vc = rng.standard_normal(len(pmid))
    dx = rng.standard_normal(disp.shape)
    dm = rng.standard_normal(shape)
    dc, mc, _ = opm.gather_adj(pmid, disp, 1.0, shape, m, vc, dtype=np.float64)
    fx = lambda x: np.sum(vc * opm.gather(pmid, x, 1.0, shape, m, dtype=np.float64))
    fm = lambda mm: np.sum(vc * opm.gather(pmid, disp, 1.0, shape, mm, dtype=np.float64))
    assert abs(_fd(fx, disp, dx) - np.sum(dc * dx)) < 1e-6 * np.abs(dc * dx).sum()
    assert abs(_fd(fm, m, dm) - np.sum(mc * dm)) < 1e-8 * np.abs(mc * dm).sum()


# --------------------------------------------------------- FFT/Poisson ---

def test_fftfreq_matches_numpy():
    k = opm.fftfreq((4, 9, 8), 2.0, np.float64)
    np.testing.assert_allclose(k[1].ravel(), 2 * np.pi * np.fft.fftfreq(9) / 2.0)
    np.testing.assert_allclose(k[2].ravel(), 2 * np.pi * np.fft.rfftfreq(8) / 2.0)


@pytest.mark.parametrize('shape', [(4, 9, 8), (7, 8, 5)])
def test_neg_grad_nyquist(shape):
    """gravity.py:37-44: even axes lose their Nyquist plane, odd axes have none."""
    kvec = opm.fftfreq(shape, 1.0, np.float32)
    pot = np.ones(np.broadcast_shapes(*[k.shape for k in kvec]), np.complex64)
    for a, k in enumerate(kvec):
        g = opm.neg_grad(k, pot, 1.0)
        nyq = opm.nyquist_mask(k, 1.0).ravel()
        assert nyq.sum() == (1 if shape[a] % 2 == 0 else 0)
        gz = np.moveaxis(g, a, 0)
        for j, isn in enumerate(nyq):
            assert (np.all(gz[j] == 0)) == (isn or k.ravel()[j] == 0)


def test_gravity_plane_wave():
    """delta = A cos(k.x): g = -1.5 Om A sin(k.x) k/|k|^2 (below Nyquist)."""
    shape, cell, Om, A = (16, 12, 10), 2.0, 0.3, 0.01
    m = (3, 2, 1)
    X = np.meshgrid(*[np.arange(n) * cell for n in shape], indexing='ij')
    kv = [2 * np.pi * mi / (n * cell) for mi, n in zip(m, shape)]
    ph = sum(k * x for k, x in zip(kv, X))
    g = opm.gravity_mesh(1 + A * np.cos(ph), Om, cell, shape, np.float64)
    k2 = sum(k * k for k in kv)
    for d in range(3):
        np.testing.assert_allclose(g[d], -1.5 * Om * A * np.sin(ph) * kv[d] / k2, atol=1e-14)


def test_gravity_vjp_finite_difference():
    shape, ms, cell = (8, 8, 8), (16, 16, 16), 0.5
    pmid, disp = ptcl(shape, ms, cell, 1.0, 3, np.float64)
    rng = np.random.default_rng(6)
    v = rng.standard_normal(disp.shape)
    dx = rng.standard_normal(disp.shape)
    f = lambda x: np.sum(v * opm.gravity(pmid, x, 0.3, cell, ms, dtype=np.float64))
    _, dc, om = opm.gravity_vjp(pmid, disp, 0.3, cell, ms, v, dtype=np.float64)
    assert abs(_fd(f, disp, dx, 1e-7) - np.sum(dc * dx)) < 1e-7 * np.abs(dc * dx).sum()
    # Omega_m enters only through the 1.5*Omega_m scale
    g = lambda o: np.sum(v * opm.gravity(pmid, disp, o, cell, ms, dtype=np.float64))
    assert abs((g(0.3 + 1e-6) - g(0.3 - 1e-6)) / 2e-6 - om) < 1e-6 * abs(om)


# ---------------------------------------------------------- cosmology ---

def test_growth_eds_closed_form():
    conf = OConf(1.0, (8, 8, 8), a_start=1 / 64)
    c = oc.growth_integ(oc.OCosmo(conf, 2.0, 0.96, 1.0, 0.05, 0.7), conf)
    a = conf.a_nbody
    np.testing.assert_allclose(oc.growth(a, c, conf), a, rtol=1e-7)
    np.testing.assert_allclose(oc.growth(a, c, conf, order=2), 3 / 7 * a**2, rtol=1e-7)
    np.testing.assert_allclose(oc.growth(a, c, conf, deriv=1), a, rtol=1e-7)


def test_growth_lcdm_vs_solve_ivp():
    conf = OConf(1.0, (8, 8, 8), a_start=1 / 32, a_nbody_maxstep=1 / 32)
    c = oc.growth_integ(oc.SimpleLCDM(conf), conf)
    a_ic = 0.5 * np.cbrt(np.finfo(np.float64).eps)
    a = conf.growth_a.copy()
    a[0] = a_ic
    sol = solve_ivp(lambda t, y: oc.growth_ode(y, t, c), (np.log(a_ic), 0.0),
                    [1, 0, 3 / 7, 0], t_eval=np.log(a), rtol=1e-12, atol=1e-14, method='DOP853')
    np.testing.assert_allclose(c.growth[0][0], sol.y[0], rtol=1e-6)
    np.testing.assert_allclose(c.growth[1][0], sol.y[2], rtol=1e-6)


def test_sigma8_vs_quad():
    conf = OConf(4.0, (8, 8, 8))
    c = oc.boltzmann(oc.SimpleLCDM(conf), conf)

    def integrand(lnk):
        # analytic Eisenstein-Hu transfer, not the piecewise-linear table
        k = np.exp(lnk)
        x = 8 * k
        W = 3 * (np.sin(x) - x * np.cos(x)) / x**3
        T = oc.transfer_fit(k, c, conf)
        P = (0.32 * c.A_s * c.k_pivot * (k / c.k_pivot)**c.n_s
             * (np.pi * (conf.c / conf.H_0)**2 / c.Omega_m * T)**2)
        return k**3 * P / (2 * np.pi**2) * W**2
    s2 = quad(integrand, np.log(1e-4), np.log(1e3), limit=5000, epsabs=0, epsrel=1e-12,
              points=np.log([0.01, 0.1, 1, 10]))[0]
    D1 = oc.growth(1.0, c, conf)
    # the Simpson quadrature at R = 8 itself
    assert abs(oc.tophat_var(8.0, c, conf)[0] / s2 - 1) < 1e-6
    # sigma8 interpolates the table linearly in R (1.8% spacing), as pmwd does
    assert abs(c.sigma8 - np.sqrt(s2) * D1) < 3e-4 * c.sigma8


def test_transfer_low_k_limit():
    conf = OConf(4.0, (8, 8, 8))
    c = oc.SimpleLCDM(conf)
    T = oc.transfer_fit(conf.transfer_k, c, conf)
    assert T[0] == pytest.approx(1.0, abs=1e-12)
    assert np.all(np.diff(T[1:200]) <= 1e-12)


# ------------------------------------------------------------- N-body ---

def small_conf(**kw):
    kw.setdefault('a_start', 1 / 4)
    kw.setdefault('a_nbody_maxstep', 1 / 4)
    return OConf(4.0, (4, 4, 4), mesh_shape=2, **kw)


def test_nbody_fp64_reversibility():
    """FP64 reversibility ~1e-12 (adjoint.tex:1546-1549)."""
    conf = small_conf()
    c = oc.boltzmann(oc.SimpleLCDM(conf), conf)
    pmid, disp = ptcl(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, 0.4, 1, np.float64)
    vel = 0.05 * np.random.default_rng(2).standard_normal(disp.shape)
    d1, v1, _ = onb.nbody(pmid, disp, vel, c, conf, dtype=np.float64)
    d0, v0, _ = onb.nbody(pmid, d1, v1, c, conf, reverse=True, dtype=np.float64)
    assert np.abs(d0 - disp).max() < 1e-10
    assert np.abs(v0 - vel).max() < 1e-10


def test_nbody_adjoint_finite_difference():
    """Reverse-time adjoint against central differences in FP64."""
    conf = small_conf()
    c = oc.boltzmann(oc.SimpleLCDM(conf), conf)
    pmid, disp = ptcl(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, 0.4, 3, np.float64)
    vel = 0.05 * np.random.default_rng(4).standard_normal(disp.shape)
    rng = np.random.default_rng(5)
    wd, wv = rng.standard_normal(disp.shape), rng.standard_normal(disp.shape)
    dx, dv = rng.standard_normal(disp.shape), rng.standard_normal(disp.shape)

    def loss(d, v, cc=c):
        d1, v1, _ = onb.nbody(pmid, d, v, cc, conf, dtype=np.float64)
        return np.sum(wd * d1) + np.sum(wv * v1)
    d1, v1, a1 = onb.nbody(pmid, disp, vel, c, conf, dtype=np.float64)
    _, (dc, vc, _), om = onb.nbody_adj(pmid, d1, v1, a1, wd, wv, c, conf, dtype=np.float64)
    eps = 1e-6
    fd = (loss(disp + eps * dx, vel + eps * dv) - loss(disp - eps * dx, vel - eps * dv)) / (2 * eps)
    an = np.sum(dc * dx) + np.sum(vc * dv)
    assert abs(fd - an) < 1e-6 * (np.abs(dc * dx).sum() + np.abs(vc * dv).sum())
    h = 1e-5
    tight = conf.replace(growth_rtol=1e-12, growth_atol=1e-12)
    lp = [loss(disp, vel, oc.boltzmann(oc.SimpleLCDM(tight, Omega_m=0.3 + s * h), tight))
          for s in (1, -1)]
    assert abs((lp[0] - lp[1]) / (2 * h) - om) < 1e-4 * abs(om)


def test_lpt_first_order_is_zeldovich():
    """With lpt_order=1 the displacement is D1(a) times -grad(phi), laplace(phi) = delta."""
    conf = OConf(2.0, (8, 8, 8), lpt_order=1)
    c = oc.boltzmann(oc.SimpleLCDM(conf), conf)
    modes = ic.linear_modes(ic.white_noise(0, conf, real=True), c, conf)
    pmid, disp, vel = ic.lpt(modes, c, conf)
    # divergence of the displacement field in Fourier space is -D1 delta
    d = disp.reshape(conf.ptcl_grid_shape + (3,)).astype(np.float64)
    kv = opm.fftfreq(conf.ptcl_grid_shape, conf.ptcl_spacing, np.float64)
    div = sum(1j * kv[a] * sfft.rfftn(d[..., a]) for a in range(3))
    D1 = oc.growth(conf.a_start, c, conf)
    delta = modes / conf.ptcl_cell_vol
    nyq = np.zeros(div.shape, bool)
    for a, k in enumerate(kv):
        nyq |= opm.nyquist_mask(k, conf.ptcl_spacing)
    ok = ~nyq
    ok[0, 0, 0] = False
    np.testing.assert_allclose(div[ok], -D1 * delta[ok], rtol=1e-4, atol=1e-6 * np.abs(delta).max())


# ------------------------------- pins of the cosmology functions ---
#
# growth_ode (boltzmann.py:197-205), the 2LPT source _L (lpt.py:40-76) and
# transfer_fit (boltzmann.py:32-123) are pinned by what mathematics and
# published fits fix, independently of their own code:
#   * D1(a) of a flat LCDM (w = -1) cosmology is Heath's integral
#     D1 = 5/2 Om E(a) int_0^a da' / (a' E(a'))^3 (Heath 1977);
#   * D2/D1^2 follows 3/7 Om(a)^(-1/143) (Bouchet et al. 1995);
#   * _L of a sum of axis-aligned cosines is the product of their second
#     derivatives, and vanishes for a single plane wave;
#   * transfer_fit: the two EH98 fits (full, eq. 16-24; no-wiggle, eq. 29-31)
#     agree within a few per cent, and within 1.5 % without baryons; the
#     sound horizon of eq. 6 is within 2 % of the eq. 26 fit; at high k the
#     CDM piece falls as ln(q)/q^2.
# A second group of tests plants plausible mistakes in copies of these
# functions and checks that each one fails a pin.

def _heath_D1(a, c):
    E = lambda x: np.sqrt(oc.E2(x, c))
    a = np.atleast_1d(a)
    I = np.array([quad(lambda x: (x * E(x))**-3, 0, ai, epsabs=0, epsrel=1e-13, limit=500)[0]
                  for ai in a])
    dE = 0.5 * oc._dE2_da(a, c) / E(a)
    D = 2.5 * c.Omega_m * E(a) * I
    dlnD = 2.5 * c.Omega_m * (a * dE * I + a * E(a) / (a * E(a))**3)   # a dD/da
    return D, dlnD


def _growth_pins(growth_ode_fn):
    """Max relative errors of D1, dD1/dlna against Heath's integral and of
    D2/D1^2 against Bouchet's fit, on the growth-table nodes (no
    interpolation), for three LCDM cosmologies."""
    conf = OConf(1.0, (8, 8, 8), a_start=1 / 64, a_nbody_maxstep=1 / 16)
    errs = [0.0, 0.0, 0.0]
    orig = oc.growth_ode
    oc.growth_ode = growth_ode_fn
    try:
        for Om in (0.25, 0.3, 0.4):
            c = oc.growth_integ(oc.SimpleLCDM(conf, Omega_m=Om), conf)
            a = conf.growth_a[1:]
            D, dD = _heath_D1(a, c)
            errs[0] = max(errs[0], np.abs(oc.growth(a, c, conf) / D - 1).max())
            errs[1] = max(errs[1], np.abs(oc.growth(a, c, conf, deriv=1) / dD - 1).max())
            r = oc.growth(a, c, conf, order=2) / oc.growth(a, c, conf)**2
            errs[2] = max(errs[2], np.abs(r / (3 / 7 * oc.Omega_m_a(a, c)**(-1 / 143)) - 1).max())
    finally:
        oc.growth_ode = orig
    return errs


GROWTH_TOL = (1e-6, 1e-6, 1e-3)


def test_growth_lcdm_vs_heath_and_bouchet():
    e = _growth_pins(oc.growth_ode)
    assert e[0] < GROWTH_TOL[0] and e[1] < GROWTH_TOL[1], e
    assert e[2] < GROWTH_TOL[2], e


def _mutant(fn, old, new):
    """A copy of oracle function ``fn`` with ``old`` replaced by ``new`` in
    its source, compiled in fn's module namespace."""
    import inspect
    src = inspect.getsource(fn)
    assert old in src, old
    ns = dict(vars(inspect.getmodule(fn)))
    exec(compile(src.replace(old, new, 1), f'<mutant {fn.__name__}>', 'exec'), ns)
    return ns[fn.__name__]


@pytest.mark.parametrize('old,new', [
    ('Omf = 1.5 * Omega_m_a(a, cosmo)', 'Omf = 1.5 * cosmo.Omega_m'),          # Om for Om(a)
    ('G2pp = Omf * G1**2 -', 'G2pp = -'),                                      # dropped source
    ('-(3 + dlnH - Omf) * G1', '-(3 + dlnH + Omf) * G1'),                      # sign
    ('(4 + dlnH) * G1p', '(4 - dlnH) * G1p'),                                  # sign
    ('(8 + 2 * dlnH - Omf) * G2', '(8 + dlnH - Omf) * G2'),                    # dropped factor
    ('(6 + dlnH) * G2p', '(6 + dlnH) * G1p'),                                  # wrong index
])
def test_growth_pins_catch_planted_mistakes(old, new):
    e = _growth_pins(_mutant(oc.growth_ode, old, new))
    assert any(x >= t for x, t in zip(e, GROWTH_TOL)), (old, new, e)


def _L_closed_form_err(L_fn, dtype):
    conf = OConf(2.0, (16, 16, 16))
    n, h = 16, 2.0
    x = np.arange(n) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    k1, k2, k3 = (2 * np.pi / (n * h) * m for m in (2, 3, 1))
    A, B, C = 0.7, -1.3, 0.4
    kv = opm.fftfreq(conf.ptcl_grid_shape, conf.ptcl_spacing, dtype=dtype)
    cx, cy, cz = np.cos(k1 * X), np.cos(k2 * Y), np.cos(k3 * Z)
    phi = (A * cx + B * cy + C * cz).astype(dtype)
    L = L_fn(kv, sfft.rfftn(phi), conf, dtype)
    want = A * B * k1**2 * k2**2 * cx * cy + A * C * k1**2 * k3**2 * cx * cz \
        + B * C * k2**2 * k3**2 * cy * cz
    e1 = np.abs(L - want).max() / np.abs(want).max()
    # a single oblique plane wave: phi_ij = -A k_i k_j cos, every 2x2 minor vanishes
    wave = (A * np.cos(k1 * X + k2 * Y + k3 * Z)).astype(dtype)
    L = L_fn(kv, sfft.rfftn(wave), conf, dtype)
    e2 = np.abs(L).max() / (A**2 * (k1 * k2)**2)
    return e1, e2


@pytest.mark.parametrize('dtype,tol', [(np.float64, 1e-12), (np.float32, 1e-5)])
def test_lpt_source_closed_form(dtype, tol):
    e1, e2 = _L_closed_form_err(ic._L, dtype)
    assert e1 < tol and e2 < tol, (e1, e2)


@pytest.mark.parametrize('old,new', [
    ('L = L - s * s', 'L = L + s * s'),                                         # sign
    ('L = L - s * s', 'L = L - s'),                                             # dropped square
    ('for j in range(dim - 1, i, -1):', 'for j in range(dim - 1, i - 1, -1):'),  # diagonal squared
    ('for j in range(i + 1, dim):', 'for j in range(i + 2, dim):'),             # dropped pair
])
def test_lpt_source_pins_catch_planted_mistakes(old, new):
    e1, e2 = _L_closed_form_err(_mutant(ic._L, old, new), np.float64)
    assert e1 > 1e-3 or e2 > 1e-3, (old, new, e1, e2)


def _transfer_pins(fit, scales):
    """Names of the transfer pins that fail for three cosmologies."""
    conf = OConf(4.0, (8, 8, 8))
    nw = conf.replace(transfer_fit_nowiggle=True)
    k = np.logspace(-3, 1, 800)                      # h/Mpc
    bad = set()
    with np.errstate(all='ignore'):
        for Om, Ob, h in [(0.3, 0.05, 0.7), (0.25, 0.04, 0.65), (0.35, 0.06, 0.75)]:
            c = oc.SimpleLCDM(conf, Omega_m=Om, Omega_b=Ob, h=h)
            if not abs(fit(conf.transfer_k[:1], c, conf)[0] - 1) < 1e-9:
                bad.add('T(0)=1')
            r = np.abs(fit(k, c, conf) / fit(k, c, nw) - 1)
            for lo, hi, tol in [(1e-3, 0.02, 0.02), (0.02, 0.5, 0.05), (0.5, 10, 0.025)]:
                m = (k >= lo) & (k < hi)
                if not r[m].max() < tol:
                    bad.add(f'full~nowiggle [{lo},{hi})')
            c0 = c.replace(Omega_b=1e-8)
            if not np.abs(fit(k, c0, conf) / fit(k, c0, nw) - 1).max() < 0.015:
                bad.add('cdm-only full~nowiggle')
            sc = scales(c, conf)
            s26 = 44.5 * np.log(9.83 / sc['w_m']) / np.sqrt(1 + 10 * sc['w_b']**0.75)
            if not abs(sc['s'] / s26 - 1) < 0.02:
                bad.add('sound horizon')
            kk = np.logspace(4, 5, 5)
            slope = np.gradient(np.log(fit(kk, c, conf)), np.log(kk))
            q = kk * c.h / (13.41 * sc['k_eq'])
            if not np.abs(slope - (-2 + 1 / np.log(q))).max() < 0.01:
                bad.add('high-k ln q/q^2')
    return bad


def test_transfer_fit_pins():
    assert _transfer_pins(oc.transfer_fit, oc.eh98_scales) == set()


@pytest.mark.parametrize('fn,old,new', [
    ('fit', 'return f_c * T_c + f_b * T_b', 'return f_c * T_c'),                 # dropped term
    ('fit', 'T0_tilde(k, alpha_c, beta_c)', 'T0_tilde(k, beta_c, alpha_c)'),     # transposed
    ('fit', 'q = k / (13.41 * k_eq)', 'q = k / 13.41'),                          # dropped factor
    ('fit', 'C = 14.2 / alpha_c', 'C = 14.2 * alpha_c'),                         # inverted
    ('fit', 'L = np.log(np.e + 1.8 * beta_c * q)', 'L = np.log(np.e + 1.8 * q)'),
    ('fit', 'L0 / (L0 + C0 * q_eff**2)', 'L0 / (L0 + C0 * q_eff)'),              # power
    ('fit', 'alpha_c = a1**-f_b * a2**-f_b**3', 'alpha_c = a1**f_b * a2**f_b**3'),  # sign
    ('fit', 'f = 1 / (1 + (k * s / 5.4)**4)', 'f = 1 / (1 + (k * s / 5.4)**2)'),
    ('fit', '(1 - f) * T0_tilde', '(1 + f) * T0_tilde'),
    ('scales', 'k_eq = 7.46e-2 * w_m', 'k_eq = 7.46e-2 * w_b'),                   # wrong operand
    ('scales', '* np.sqrt(6 / R_eq)', '* np.sqrt(6 * R_eq)'),
    ('scales', '(1 + np.sqrt(R_eq))', '(1 + np.sqrt(R_d))'),                      # wrong index
    ('scales', 'R_d = 31.5 * w_b / T2_cmb_norm**2 * (1e3 / z_d)',
     'R_d = 31.5 * w_b / T2_cmb_norm**2 * (1e3 / z_eq)'),
])
def test_transfer_pins_catch_planted_mistakes(fn, old, new):
    fit, scales = oc.transfer_fit, oc.eh98_scales
    if fn == 'fit':
        fit = _mutant(oc.transfer_fit, old, new)
    else:
        scales = _mutant(oc.eh98_scales, old, new)
        # the fit reads the scales through the module namespace
        fit = _mutant(oc.transfer_fit, 'sc = eh98_scales(cosmo, conf)',
                      'sc = _scales(cosmo, conf)')
        fit.__globals__['_scales'] = scales
    assert _transfer_pins(fit, scales), (fn, old, new)


@pytest.mark.parametrize('shape,ratio,cell,std', [((16, 16, 16), 2, 1.7, 3.0),
                                                  ((7, 9, 5), 1, 1.0, 9.0),
                                                  ((8, 8, 8), 2, 0.5, 40.0)])
@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_shared_cic_equals_enmesh_and_composition(shape, ratio, cell, std, dtype):
    """gravity/gravity_vjp compute the CIC geometry once (cic_default); they
    equal enmesh, and scatter + three gathers (and the adjoint chain of
    gather_adj/scatter_adj), bit for bit."""
    ms = tuple(s * ratio for s in shape)
    pmid, disp = ptcl(shape, ms, cell, std, 7, dtype)
    disp[::13] = 0
    disp[1::17] = -0.0
    ind, w, fg = opm.enmesh(pmid, disp, cell, ms, 0, None, ms, grad=True)
    flat, w2, fg2 = opm.cic_default(pmid, disp, cell, ms, grad=True)
    np.testing.assert_array_equal(opm.raveled(ind, ms), flat)
    np.testing.assert_array_equal(w, w2)
    np.testing.assert_array_equal(fg, fg2)
    dens = opm.scatter(pmid, disp, cell, ms, dtype=dtype)
    g = opm.gravity_mesh(dens, 0.3, cell, ms, dtype)
    want = np.stack([opm.gather(pmid, disp, cell, ms, gi, dtype=dtype) for gi in g], -1)
    np.testing.assert_array_equal(opm.gravity(pmid, disp, 0.3, cell, ms, dtype=dtype), want)
    ac = np.random.default_rng(8).standard_normal(disp.shape).astype(dtype)
    kvec = opm.fftfreq(ms, cell, dtype=dtype)
    dcot, dsum = np.zeros_like(disp), None
    for d in range(3):
        dc, mc, _ = opm.gather_adj(pmid, disp, cell, ms, g[d], ac[:, d], dtype=dtype)
        dcot = dcot + dc
        t = opm.neg_grad(kvec[d], opm.laplace(kvec, sfft.rfftn(mc)), cell)
        dsum = t if dsum is None else dsum + t
    dens_cot = (-sfft.irfftn(dsum, s=ms)).astype(dtype) * (dtype(1.5) * dtype(0.3))
    dc2, _ = opm.scatter_adj(pmid, disp, cell, ms, dens_cot, dtype=dtype)
    _, got, _ = opm.gravity_vjp(pmid, disp, 0.3, cell, ms, ac, dtype=dtype)
    np.testing.assert_array_equal(got, (dcot + dc2).astype(dtype))


# ------------------------------------------------------- white noise ---

# Philox4x32-10 known-answer vectors published with Random123 (Salmon et al.
# 2011): counter, key, output
PHILOX_KAT = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
              ((0xffffffff,) * 4, (0xffffffff,) * 2,
               (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
              ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
               (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]


@pytest.mark.parametrize('ctr,key,want', PHILOX_KAT)
def test_philox4x32_known_answers(ctr, key, want):
    assert tuple(ic.philox(ctr, key, 32)) == want


def test_philox4x64_matches_numpy():
    """The same round function with the 4x64 constants reproduces numpy's
    independent Philox (which increments its counter before each block)."""
    key = (123456789, 987654321)
    raw = [int(x) for x in np.random.Philox(counter=0, key=list(key)).random_raw(12)]
    got = sum((ic.philox((b, 0, 0, 0), key, 64) for b in (1, 2, 3)), [])
    assert got == raw


def test_philox_vectorised_and_normals():
    """The vectorised generator equals the Python-int rounds (low and high
    counter words, seeds above 2^32); the normals follow the documented
    Box-Muller of each block's word pairs, and have the moments of N(0, 1)."""
    seed = (7 << 32) + 99
    blocks = np.array([0, 1, 2, 12345, 2**32 - 1, 2**32 + 5, 2**40 + 3])
    got = ic.philox4x32_blocks(seed, blocks)
    for b, row in zip(blocks, got):
        want = ic.philox((int(b) & 0xffffffff, int(b) >> 32, 0, 0),
                         (seed & 0xffffffff, seed >> 32), 32)
        assert [int(x) for x in row] == want
    w = ic.philox(((5), 0, 0, 0), (seed & 0xffffffff, seed >> 32), 32)
    u = [(x + 0.5) / 2**32 for x in w]
    import math
    r0, r1 = math.sqrt(-2 * math.log(u[0])), math.sqrt(-2 * math.log(u[2]))
    want = np.float32([r0 * math.cos(2 * math.pi * u[1]), r0 * math.sin(2 * math.pi * u[1]),
                       r1 * math.cos(2 * math.pi * u[3]), r1 * math.sin(2 * math.pi * u[3])])
    np.testing.assert_array_equal(ic.normals(seed, np.arange(20, 24)), want)
    z = ic.normals(3, np.arange(1 << 21)).astype(np.float64)
    n = z.size
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(z.var() - 1) < 5 * np.sqrt(2 / n)
    assert abs(np.mean(z**3)) < 5 * np.sqrt(15 / n)
    assert abs(np.mean(z**4) - 3) < 5 * np.sqrt(96 / n)
    # different seeds and neighbouring elements are uncorrelated
    z2 = ic.normals(4, np.arange(1 << 21)).astype(np.float64)
    assert abs(np.mean(z * z2)) < 5 / np.sqrt(n)
    assert abs(np.mean(z[1:] * z[:-1])) < 5 / np.sqrt(n)


# ---------------------------------------------------- power spectrum ---

def test_powspec_white_noise_and_parseval():
    """Unit white noise has P = spacing^3 in every bin (within five standard
    errors, sqrt(2/N) of the mean); with no cuts the binned power summed
    over all modes equals sum f^2 (Parseval) and the mode counts add up to
    the full grid."""
    from oracle import spec
    for shape in ((32, 32, 32), (24, 20, 17)):
        f = np.random.default_rng(1).standard_normal(shape)
        h = 1.5
        k, P, N, bins = spec.powspec(f, h, bins=1.0, cut_zero=False, cut_nyq=False)
        M = np.prod(shape)
        assert N.sum() == M
        assert np.isclose((P * N).sum() * M / h**3, M * np.sum(f**2), rtol=1e-10)
        k, P, N, bins = spec.powspec(f, h)
        assert np.all(np.abs(P / h**3 - 1) < 5 * np.sqrt(2 / N) + 1e-12)
        assert np.all(np.diff(k) > 0) and np.all((k[1:] > bins[1:len(k)]))


def test_powspec_plane_wave():
    """f = A cos(2 pi m x / n) puts all power in the bin holding |k| = 2 pi m
    / (n h): that bin's P is 2 (A n^3 / 2)^2 / N_bin h^3 / n^3, every other
    bin's is zero, and its mean k includes the wave's modes with weight 1
    each (kz = 0 plane)."""
    from oracle import spec
    n, m, A, h = 32, 3, 0.7, 2.0
    x = np.arange(n)
    f = np.broadcast_to(A * np.cos(2 * np.pi * m * x / n)[:, None, None], (n, n, n))
    k, P, N, bins = spec.powspec(f, h, bins=1.0)
    kw = 2 * np.pi * m / (n * h)
    b = int(np.argmin(np.abs(k - kw)))
    want = 2 * (A * n**3 / 2)**2 / N[b] * h**3 / n**3
    assert np.isclose(P[b], want, rtol=1e-10)
    assert np.allclose(np.delete(P, b), 0, atol=1e-20 * want)
