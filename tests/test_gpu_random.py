"""The device white noise (pmx_white_noise, pmx_random_bits) against the
oracle's Philox4x32-10 and Box-Muller restatement: raw words bit for bit,
normals within one float32 ulp (the float64 log/sin/cos of the device and
host libraries may round differently, which changes a float32 result only
when it lies within ~1e-16 of a rounding boundary), moments at full size."""
import numpy as np
import pytest
import torch

from oracle import ic as oic

pytestmark = pytest.mark.gpu


def _ctx(dev):
    from pmwd_amd import _lib
    return _lib, _lib.context_any_cell((8, 8, 8), dev)


@pytest.mark.parametrize('seed', [0, 1, 2**33 + 7, 2**64 - 1])
def test_random_bits_match_oracle(dev, seed):
    _lib, c = _ctx(dev)
    nb = 4099
    out = torch.empty(4 * nb, dtype=torch.int32, device=dev)
    _lib.check(_lib.lib().pmx_random_bits(c.bind(), seed, nb, _lib.ptr(out)))
    got = out.cpu().numpy().view(np.uint32).reshape(nb, 4)
    np.testing.assert_array_equal(got, oic.philox4x32_blocks(seed, np.arange(nb)))
    if seed == 0:
        assert [int(x) for x in got[0]] == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]


def _ulp_compare(got, want):
    got = np.asarray(got, np.float32)
    want = np.asarray(want, np.float32)
    diff = np.abs(got.astype(np.float64) - want)
    ulp = np.spacing(np.abs(want)).astype(np.float64)
    assert np.all(diff <= ulp), float((diff / ulp).max())
    return int(np.count_nonzero(diff))


def test_white_noise_matches_oracle(dev):
    import pmwd_amd as pa
    for shape, seed in [((64, 64, 64), 0), ((7, 9, 5), 12345), ((33, 16, 10), 2**40 + 1)]:
        conf = pa.Configuration(1.0, shape)
        w = pa.white_noise(seed, conf, real=True, device=dev).cpu().numpy()
        want = oic.white_noise(seed, _oconf(shape), real=True)
        mism = _ulp_compare(w, want)
        assert mism <= max(1, w.size // 10**6), mism


def _oconf(shape):
    from oracle.conf import OConf
    return OConf(1.0, shape)


def test_white_noise_full_size_moments_and_samples(dev):
    """512^3 normals: mean, variance, skewness and kurtosis of N(0, 1) within
    five standard errors; neighbour correlation nil; 2^20 sampled elements
    equal to the oracle's within an ulp; the Fourier path is the 'ortho'
    rfftn of the real noise."""
    import pmwd_amd as pa
    conf = pa.Configuration(1.0, (512, 512, 512))
    w = pa.white_noise(7, conf, real=True, device=dev)
    z = w.reshape(-1).double()
    n = z.numel()
    m = float(z.mean())
    v = float(z.var())
    s3 = float((z**3).mean())
    s4 = float((z**4).mean())
    c1 = float((z[1:] * z[:-1]).mean())
    assert abs(m) < 5 / np.sqrt(n) and abs(v - 1) < 5 * np.sqrt(2 / n)
    assert abs(s3) < 5 * np.sqrt(15 / n) and abs(s4 - 3) < 5 * np.sqrt(96 / n)
    assert abs(c1) < 5 / np.sqrt(n)
    idx = np.random.default_rng(1).integers(0, n, size=1 << 20)
    got = w.reshape(-1)[torch.as_tensor(idx, device=dev)].cpu().numpy()
    mism = _ulp_compare(got, oic.normals(7, idx))
    assert mism <= 4, mism
    del z
    small = pa.Configuration(1.0, (32, 32, 32))
    wr = pa.white_noise(3, small, real=True, device=dev)
    wk = pa.white_noise(3, small, device=dev)
    np.testing.assert_allclose(wk.cpu().numpy(), np.fft.rfftn(wr.cpu().numpy(), norm='ortho'),
                               rtol=1e-5, atol=1e-5)
