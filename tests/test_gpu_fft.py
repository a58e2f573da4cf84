"""GPU parity of the hand-written FFTs and the fused Poisson solve."""
import numpy as np
import pytest
import scipy.fft as sfft
import torch

from oracle import pm as opm

pytestmark = pytest.mark.gpu

FFT_SHAPES = [(8, 8, 8), (16, 32, 64), (4, 9, 7), (7, 8, 5), (12, 10, 6), (64, 64, 64),
              (2, 4, 1024), (1024, 2, 4), (3, 5, 2), (1, 1, 8), (128, 128, 128),
              (256, 64, 512), (2, 4, 2048), (3, 256, 8)]


def rel_l2(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize('shape', FFT_SHAPES)
def test_rfftn_irfftn(dev, shape):
    from pmwd_amd.pm_util import fftfwd, fftinv
    x = np.random.default_rng(0).standard_normal(shape).astype(np.float32)
    X = fftfwd(torch.as_tensor(x, device=dev)).cpu().numpy()
    Xw = sfft.rfftn(x.astype(np.float64))
    assert X.shape == Xw.shape
    assert rel_l2(X, Xw) < 2e-6
    y = fftinv(torch.as_tensor(Xw.astype(np.complex64), device=dev), shape=shape).cpu().numpy()
    yw = sfft.irfftn(Xw, s=shape)
    assert rel_l2(y, yw) < 2e-6


def test_irfftn_drops_imag_of_self_conjugate_bins(dev):
    """irfftn ignores Im X[k_z=0] and Im X[k_z=n/2] as numpy does."""
    from pmwd_amd.pm_util import fftinv
    shape = (6, 4, 8)
    rng = np.random.default_rng(1)
    X = (rng.standard_normal((6, 4, 5)) + 1j * rng.standard_normal((6, 4, 5))).astype(np.complex64)
    got = fftinv(torch.as_tensor(X, device=dev), shape=shape).cpu().numpy()
    want = sfft.irfftn(X.astype(np.complex128), s=shape)
    assert rel_l2(got, want) < 2e-6


def test_fft_norms(dev):
    from pmwd_amd.pm_util import fftfwd, fftinv
    x = np.random.default_rng(2).standard_normal((8, 12, 10)).astype(np.float32)
    xt = torch.as_tensor(x, device=dev)
    for norm in ('ortho', 'forward'):
        X = fftfwd(xt, norm=norm).cpu().numpy()
        assert rel_l2(X, sfft.rfftn(x.astype(np.float64), norm=norm)) < 2e-6
    X = fftfwd(xt, norm=2.0).cpu().numpy()
    assert rel_l2(X, 8.0 * sfft.rfftn(x.astype(np.float64))) < 2e-6
    y = fftinv(torch.as_tensor(X.astype(np.complex64), device=dev), shape=x.shape, norm=2.0)
    assert rel_l2(y.cpu().numpy(), x) < 2e-6


# (128, 64, 256) and (64, 1024, 16) run the fixed-length kernel variants
# (x 64/128, y 64/1024, z half-length 128); (2, 4096, 4) has the longest y
# axis, one pencil pair per block in the paired y pass
POISSON = [((16, 16, 16), 2.0), ((4, 9, 8), 7.0), ((7, 8, 6), 7.0), ((32, 16, 24), 1.0),
           ((64, 64, 64), 3.0), ((1, 1, 2), 1.0), ((128, 64, 256), 1.0), ((64, 1024, 16), 2.0),
           ((2, 4096, 4), 1.0)]


@pytest.mark.parametrize('shape,cell', POISSON)
def test_poisson_grad(dev, shape, cell):
    from pmwd_amd import _lib
    rng = np.random.default_rng(3)
    dens = (1 + 0.3 * rng.standard_normal(shape)).astype(np.float32)
    coef = float(np.float32(1.5) * np.float32(0.3))
    c = _lib.context(shape, cell, dev)
    out = torch.empty((3,) + shape, dtype=torch.float32, device=dev)
    _lib.check(_lib.lib().pmx_poisson_grad(c.bind(), _lib.ptr(torch.as_tensor(dens, device=dev)),
                                           coef, _lib.ptr(out)))
    want = opm.gravity_mesh(dens, 0.3, cell, shape)
    for d in range(3):
        w = want[d]
        if np.linalg.norm(w) == 0:
            assert np.abs(out[d].cpu().numpy()).max() < 1e-6
        else:
            assert rel_l2(out[d].cpu().numpy(), w) < 5e-6, d


def test_poisson_plane_wave(dev):
    """Closed form: delta = A cos(k.x) gives g = -1.5 Om A sin(k.x) k/|k|^2."""
    from pmwd_amd import _lib
    shape, cell, Om, A = (32, 16, 24), 1.5, 0.3, 0.01
    m = (3, 2, 5)
    x = [np.arange(n) * cell for n in shape]
    X = np.meshgrid(*x, indexing='ij')
    kv = [2 * np.pi * mi / (n * cell) for mi, n in zip(m, shape)]
    phase = sum(k * xx for k, xx in zip(kv, X))
    dens = (1 + A * np.cos(phase)).astype(np.float32)
    coef = float(np.float32(1.5) * np.float32(Om))
    c = _lib.context(shape, cell, dev)
    out = torch.empty((3,) + shape, dtype=torch.float32, device=dev)
    _lib.check(_lib.lib().pmx_poisson_grad(c.bind(), _lib.ptr(torch.as_tensor(dens, device=dev)),
                                           coef, _lib.ptr(out)))
    k2 = sum(k * k for k in kv)
    for d in range(3):
        want = -1.5 * Om * A * np.sin(phase) * kv[d] / k2
        np.testing.assert_allclose(out[d].cpu().numpy(), want, atol=2e-7 + 2e-5 * np.abs(want).max())


@pytest.mark.parametrize('n', [256, 512])
def test_fft_production_sizes_vs_rocfft(dev, n):
    """r2c/c2r at the benchmark's mesh sizes against torch.fft (rocFFT) on the
    same device, float32 both."""
    import torch
    from pmwd_amd.pm_util import fftfwd, fftinv
    shape = (n, n, n)
    x = torch.randn(shape, generator=torch.Generator(device=dev).manual_seed(n), device=dev)
    ref = torch.fft.rfftn(x).contiguous()
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    assert rel(fftfwd(x), ref) < 1e-6
    assert rel(fftinv(ref, shape=shape), torch.fft.irfftn(ref.clone(), s=shape)) < 1e-6


def _poisson_torch(dens, coef, cell):
    """gravity.py:51-67 in plain torch float32 (rocFFT): the three force meshes."""
    import torch
    shape = dens.shape
    ks = []
    for d, n in enumerate(shape):
        f = (torch.fft.rfftfreq(n, d=cell) if d == 2 else torch.fft.fftfreq(n, d=cell))
        k = (2 * np.pi * f.double()).float().to(dens.device)
        ks.append(k.reshape([-1 if i == d else 1 for i in range(3)]))
    src = torch.fft.rfftn((dens - 1) * coef)
    k2 = ks[0]**2 + ks[1]**2 + ks[2]**2
    pot = torch.where(k2 != 0, -src / torch.where(k2 != 0, k2, 1), 0)
    out = []
    for d, n in enumerate(shape):
        knyq = np.pi / cell
        kd = torch.where((ks[d].abs() - knyq).abs() <= knyq * np.finfo(np.float32).eps,
                         0, ks[d])
        out.append(torch.fft.irfftn(-1j * kd * pot, s=shape))
    return torch.stack(out)


@pytest.mark.parametrize('n,cell', [(256, 2.0), (512, 1.953125)])
def test_poisson_production_sizes_vs_torch(dev, n, cell):
    """The fused Poisson solve at the benchmark's sizes against a plain
    torch float32 restatement of gravity.py:51-67 on rocFFT."""
    import torch
    from pmwd_amd import _lib
    shape = (n, n, n)
    g = torch.Generator(device=dev).manual_seed(5)
    dens = 1 + 0.5 * torch.randn(shape, generator=g, device=dev)
    coef = float(np.float32(1.5) * np.float32(0.3))
    c = _lib.context(shape, cell, dev)
    out = torch.empty((3,) + shape, dtype=torch.float32, device=dev)
    _lib.check(_lib.lib().pmx_poisson_grad(c.bind(), _lib.ptr(dens.contiguous()), coef,
                                           _lib.ptr(out)))
    want = _poisson_torch(dens, coef, cell)
    rel = float(torch.linalg.vector_norm(out - want) / torch.linalg.vector_norm(want))
    assert rel < 2e-6


def test_poisson_longest_x_axis(dev):
    """The longest supported axis (4096) in the Poisson x pass, whose pencil
    block and twiddles use the whole 160 KB of a workgroup's LDS."""
    import torch
    from pmwd_amd import _lib
    shape, cell = (4096, 2, 4), 1.0
    g = torch.Generator(device=dev).manual_seed(6)
    dens = 1 + 0.5 * torch.randn(shape, generator=g, device=dev)
    coef = float(np.float32(0.45))
    c = _lib.context(shape, cell, dev)
    out = torch.empty((3,) + shape, dtype=torch.float32, device=dev)
    _lib.check(_lib.lib().pmx_poisson_grad(c.bind(), _lib.ptr(dens), coef, _lib.ptr(out)))
    want = _poisson_torch(dens, coef, cell)
    rel = float(torch.linalg.vector_norm(out - want) / torch.linalg.vector_norm(want))
    assert rel < 2e-6
