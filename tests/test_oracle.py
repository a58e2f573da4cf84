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
