"""Gradients through the whole forward model on the GPU: white modes ->
linear_modes -> lpt -> nbody -> scatter -> density MSE against a seeded
target (the forward-plus-adjoint configuration at reduced size), with
respect to the modes, Omega_m and sigma8 (A_s following sigma8 as in
Cosmology.from_sigma8), checked against central finite differences of the
same pipeline."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _conf():
    import pmwd_amd as pa
    return pa.Configuration(4.0, (16, 16, 16), mesh_shape=2, a_start=1 / 8, a_nbody_maxstep=1 / 8)


def _model(white, conf, Om, s8, with_theta=False):
    import pmwd_amd as pa
    cosmo = pa.boltzmann(pa.Cosmology.from_sigma8(conf, s8, n_s=0.96, Omega_m=Om, Omega_b=0.05,
                                                  h=0.7), conf)
    theta = None
    if with_theta:
        theta = {'Omega_m': torch.tensor(cosmo.Omega_m, dtype=torch.float64, requires_grad=True),
                 'sigma8': torch.tensor(cosmo.sigma8, dtype=torch.float64, requires_grad=True)}
    modes = pa.linear_modes(white, cosmo, conf, theta=theta)
    ptcl, _ = pa.lpt(modes, cosmo, conf, theta=theta)
    ptcl, _ = pa.nbody(ptcl, None, cosmo, conf, omega_m=theta['Omega_m'] if theta else None)
    return pa.scatter(ptcl, conf), theta


def _loss(dens, tgt):
    return ((dens.double() - tgt.double())**2).mean()


def test_forward_model_gradients_finite_difference(dev):
    import pmwd_amd as pa
    conf = _conf()
    tgt, _ = _model(pa.white_noise(1, conf, real=True, device=dev), conf, 0.3, 0.85)
    white = pa.white_noise(0, conf, real=True, device=dev).requires_grad_(True)
    dens, theta = _model(white, conf, 0.3, 0.8, with_theta=True)
    loss = _loss(dens, tgt)
    g_w, g_om, g_s8 = torch.autograd.grad(loss, (white, theta['Omega_m'], theta['sigma8']))
    assert torch.isfinite(g_w).all()

    def L(w=None, Om=0.3, s8=0.8):
        with torch.no_grad():
            d, _ = _model(white.detach() if w is None else w, conf, Om, s8)
            return float(_loss(d, tgt))

    eps = 1e-3
    fd_om = (L(Om=0.3 + eps) - L(Om=0.3 - eps)) / (2 * eps)
    fd_s8 = (L(s8=0.8 + eps) - L(s8=0.8 - eps)) / (2 * eps)
    v = torch.randn(white.shape, generator=torch.Generator().manual_seed(7)).to(dev)
    w0 = white.detach()
    # the density is piecewise linear in the positions (CIC): the difference
    # quotient carries the kinks its perturbation crosses and moves with the
    # step, so the modes gradient has to lie within the spread of several
    fds = [(L(w=w0 + e * v) - L(w=w0 - e * v)) / (2 * e) for e in (1e-3, 5e-4, 2.5e-4)]
    an_w = float((g_w.double() * v.double()).sum())
    print(f'Omega_m {float(g_om):.6e} fd {fd_om:.6e}; sigma8 {float(g_s8):.6e} fd {fd_s8:.6e}; '
          f'modes.v {an_w:.6e} fd {fds}')
    assert abs(float(g_om) - fd_om) < 0.03 * abs(fd_om)
    assert abs(float(g_s8) - fd_s8) < 0.03 * abs(fd_s8)
    pad = 0.03 * max(abs(f) for f in fds)
    assert min(fds) - pad <= an_w <= max(fds) + pad, (an_w, fds)


def test_library_fft_autograd_matches_rocfft_autograd(dev):
    """Backward of the library r2c/c2r equals torch's autograd through
    torch.fft (rocFFT) on the same device."""
    from pmwd_amd.pm_util import fftfwd, fftinv
    g = torch.Generator().manual_seed(4)
    for shape in ((8, 6, 10), (5, 6, 7)):
        x = torch.randn(shape, generator=g).to(dev).requires_grad_(True)
        gc = torch.randn(shape[0], shape[1], shape[2] // 2 + 1, 2, generator=g).to(dev)
        gc = torch.view_as_complex(gc)
        (got,) = torch.autograd.grad(fftfwd(x), x, gc)
        (want,) = torch.autograd.grad(torch.fft.rfftn(x), x, gc)
        torch.testing.assert_close(got, want, rtol=1e-4, atol=1e-4)
        h = gc.clone().requires_grad_(True)
        y = torch.randn(shape, generator=g).to(dev)
        (got,) = torch.autograd.grad(fftinv(h, shape=shape), h, y)
        (want,) = torch.autograd.grad(torch.fft.irfftn(h, s=shape), h, y)
        torch.testing.assert_close(got, want, rtol=1e-4, atol=1e-5)
