"""powspec on the GPU (pmx_powspec) against the oracle's restatement of
pmwd/spec_util.py, and the linear power spectrum of the initial conditions
of the 512^3 configuration (downscaled to 128^3) against linear theory."""
import math

import numpy as np
import pytest
import torch

from oracle import spec as ospec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('shape', [(32, 32, 32), (24, 20, 16), (7, 9, 5)])
@pytest.mark.parametrize('kw', [dict(), dict(bins=1.0), dict(bins=2.5, cut_nyq=False),
                                dict(bins=(0, 0.05, 0.1, 0.2, 0.3, 0.5, 0.9)),
                                dict(cut_zero=False), dict(deconv=2), dict(cross=True)])
def test_powspec_matches_oracle(dev, shape, kw):
    import pmwd_amd as pa
    kw = dict(kw)
    rng = np.random.default_rng(sum(shape))
    f = rng.standard_normal(shape).astype(np.float32)
    g = rng.standard_normal(shape).astype(np.float32) if kw.pop('cross', False) else None
    h = 1.7
    out = pa.powspec(torch.as_tensor(f, device=dev), h,
                     g=torch.as_tensor(g, device=dev) if g is not None else None, **kw)
    want = ospec.powspec(f, h, g=g, **kw)
    k, P, N, bins = (t.cpu().numpy() for t in out)
    np.testing.assert_array_equal(N, want[2])
    np.testing.assert_array_equal(bins, want[3])
    np.testing.assert_allclose(k, want[0], rtol=1e-12)
    scale = np.nanmax(np.abs(want[1]))
    tol = 1e-5 if kw.get('deconv') else 2e-6
    np.testing.assert_allclose(P, want[1], rtol=tol, atol=tol * scale)


def test_powspec_channels_sum(dev):
    """Leading axes are reduced by the sum of the power (spec_util.py:114-115)."""
    import pmwd_amd as pa
    f = np.random.default_rng(3).standard_normal((2, 3, 16, 16, 16)).astype(np.float32)
    k, P, N, _ = (t.cpu().numpy() for t in pa.powspec(torch.as_tensor(f, device=dev), 1.0))
    wk, wP, wN, _ = ospec.powspec(f, 1.0)
    np.testing.assert_array_equal(N, wN)
    np.testing.assert_allclose(P, wP, rtol=2e-6)


def test_linear_power_of_initial_conditions(dev):
    """The linear density of the 512^3 configuration's initial conditions
    (downscaled to 128^3 particles in the same 1000 Mpc/h box), white noise
    -> linear_modes -> fftinv, grown to a = 1/32: its power spectrum in each
    bin equals linear theory averaged over the bin's modes, within five
    standard errors of a Gaussian field's estimate, sqrt(2/N)."""
    import pmwd_amd as pa
    from pmwd_amd.boltzmann import growth
    from pmwd_amd.pm_util import fftinv
    n, box = 128, 1000.0
    conf = pa.Configuration(box / n, (n,) * 3, mesh_shape=1, a_start=1 / 32,
                            a_nbody_maxstep=1 / 32)
    cosmo = pa.boltzmann(pa.Cosmology.from_sigma8(conf, 0.8, n_s=0.96, Omega_m=0.3,
                                                  Omega_b=0.05, h=0.7), conf)
    a = conf.a_start
    modes = pa.linear_modes(pa.white_noise(0, conf, real=True, device=dev), cosmo, conf)
    dens = fftinv(modes, shape=conf.ptcl_grid_shape, norm=conf.ptcl_spacing)
    dens = dens * float(growth(a, cosmo, conf))
    k, P, N, bins = (t.cpu().numpy() for t in pa.powspec(dens, conf.ptcl_spacing))
    # linear theory on the same modes and weights, binned the same way
    kv = np.meshgrid(*[np.fft.fftfreq(n), np.fft.fftfreq(n), np.fft.rfftfreq(n)],
                     sparse=True, indexing='ij')
    kc = np.sqrt(sum(x.astype(np.float32)**2 for x in kv)).astype(np.float32)
    w = np.full(kc.shape, 2.0)
    w[..., 0] = 1
    w[..., -1] = 1
    kang = kc.astype(np.float64) * 2 * math.pi / conf.ptcl_spacing
    Plin = pa.linear_power(kang.ravel(), a, cosmo, conf).reshape(kc.shape)
    bnum, bcut, edges, right = ospec.getbins((n,) * 3, 1j / 3, True)
    b = np.digitize(kc.ravel().astype(np.float64), edges, right=right)
    keep = b < bnum
    num = np.bincount(b[keep], weights=(Plin * w).ravel()[keep], minlength=bnum)
    den = np.bincount(b[keep], weights=w.ravel()[keep], minlength=bnum)
    want = (num / np.maximum(den, 1))[1:bcut]
    np.testing.assert_array_equal(N, den[1:bcut])
    z = (P / want - 1) / np.sqrt(2 / N)
    assert np.all(np.abs(z) < 5), z
    # the bins hold enough modes for the estimate to be meaningful overall
    assert abs(np.sum(P * N) / np.sum(want * N) - 1) < 0.02


def test_c3_large_scale_growth(dev):
    """The 512^3 configuration at full size (512^3 particles, 512^3 mesh,
    box 1000 Mpc/h, 2LPT at a = 1/32, 31 KDK steps to a = 1): on the largest
    scales the final density's power (CIC window deconvolved) equals that of
    the linear density of the same modes at a = 1.  The step factors follow
    the growth functions, so linear modes grow exactly; with the same modes
    on both sides cosmic variance cancels, and what remains is mode
    coupling, small at k < 0.02 h/Mpc (the five bins below 3.2 k_f;
    measured ratios 0.984-0.997, the limit is 3%)."""
    import pmwd_amd as pa
    from pmwd_amd.boltzmann import growth
    from pmwd_amd.pm_util import fftinv
    from test_gpu_configs import _record
    n, box = 512, 1000.0
    conf = pa.Configuration(box / n, (n,) * 3, mesh_shape=1, a_start=1 / 32,
                            a_nbody_maxstep=1 / 32)
    cosmo = pa.boltzmann(pa.Cosmology.from_sigma8(conf, 0.8, n_s=0.96, Omega_m=0.3,
                                                  Omega_b=0.05, h=0.7), conf)
    with torch.no_grad():
        modes = pa.linear_modes(pa.white_noise(0, conf, real=True, device=dev), cosmo, conf)
        lin = fftinv(modes, shape=conf.ptcl_grid_shape, norm=conf.ptcl_spacing)
        lin = lin * float(growth(1.0, cosmo, conf))
        ptcl, _ = pa.lpt(modes, cosmo, conf)
        del modes
        ptcl, _ = pa.nbody(ptcl, None, cosmo, conf)
        dens = pa.scatter(ptcl, conf)
        del ptcl
        dens = dens - 1
        k, P, N, _ = (t.cpu().numpy() for t in pa.powspec(dens, conf.cell_size, deconv=2))
        kl, Pl, Nl, _ = (t.cpu().numpy() for t in pa.powspec(lin, conf.ptcl_spacing))
    np.testing.assert_array_equal(N, Nl)
    # k in h/Mpc (angular, per unit of the spacing passed, Mpc/h)
    kf = 2 * math.pi / box
    low = k < 3.2 * kf
    ratio = P[low] / Pl[low]
    kh = k[low]
    _record('C3_large_scale_growth', k_h_per_Mpc=kh.tolist(), ratio=ratio.tolist(),
            modes=N[low].tolist())
    assert low.sum() >= 4
    assert np.all(np.abs(ratio - 1) < 0.03), dict(k=kh, ratio=ratio)
