"""Host-side logic of pmwd_amd (no GPU): configuration, FP64 tables and
sensitivities, time-step factors, and the C ABI's exported symbols."""
import ctypes
import os
import re

import numpy as np
import pytest

import importlib

import pmwd_amd as pa
from oracle import cosmo as oc, nbody as onb
from oracle.conf import OConf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pnb = importlib.import_module('pmwd_amd.nbody')  # the package re-exports nbody()


def test_configuration_errors():
    """configuration.py:155-191 validation."""
    with pytest.raises(ValueError):
        pa.Configuration(1.0, (8, 8, 8), mesh_shape=(4, 8, 8))
    with pytest.raises(ValueError):
        pa.Configuration(1.0, (8, 8, 8), mesh_shape=(16, 8, 8))
    with pytest.raises(ValueError):
        pa.Configuration(1.0, (8, 8, 8), mesh_shape=(8, 8))
    with pytest.raises(ValueError):
        pa.Configuration(1.0, (8, 8, 8), symp_splits=((0, 0.5), (1, 0.4)))
    with pytest.raises(ValueError):
        pa.Configuration(1.0, (8, 8, 8), pmid_dtype=np.float32)


@pytest.mark.parametrize('kw', [dict(), dict(mesh_shape=2, a_start=1 / 32, a_nbody_maxstep=1 / 32),
                                dict(mesh_shape=1.5)])
def test_configuration_matches_oracle(kw):
    c = pa.Configuration(1000 / 64, (64, 64, 64), **kw)
    o = OConf(1000 / 64, (64, 64, 64), **kw)
    assert c.mesh_shape == o.mesh_shape
    assert c.cell_size == o.cell_size
    np.testing.assert_array_equal(c.a_nbody, o.a_nbody)
    np.testing.assert_array_equal(c.growth_a, o.growth_a)
    np.testing.assert_array_equal(c.transfer_k, o.transfer_k)
    assert c.a_nbody_num == o.a_nbody_num


def _cosmos(Om=0.3, **kw):
    conf = pa.Configuration(4.0, (16, 16, 16), mesh_shape=2, **kw)
    oconf = OConf(4.0, (16, 16, 16), mesh_shape=2, **kw)
    return (conf, pa.boltzmann(pa.SimpleLCDM(conf, Omega_m=Om), conf), oconf,
            oc.boltzmann(oc.SimpleLCDM(oconf, Omega_m=Om), oconf))


def test_tables_match_oracle():
    conf, c, oconf, o = _cosmos()
    np.testing.assert_allclose(c.transfer, o.transfer, rtol=1e-13)
    # rtol: the accepted ODE steps follow each implementation's rounding
    # (see test_growth_accuracy_matches_oracle)
    np.testing.assert_allclose(c.growth, o.growth, rtol=1e-7, atol=1e-14)
    np.testing.assert_allclose(c.varlin, o.varlin, rtol=1e-12)
    assert abs(c.sigma8 - o.sigma8) < 1e-12


def test_growth_accuracy_matches_oracle():
    """The library's Dormand-Prince integration and the oracle's are equally
    close to the converged solution (both integrators at rtol = atol = 1e-13):
    the error of each table entry at the default tolerance agrees within 5%,
    and the two differ from each other by far less than that error."""
    conf, c, oconf, o = _cosmos()
    tconf = conf.replace(growth_rtol=1e-13, growth_atol=1e-13)
    tight = pa.boltzmann(pa.SimpleLCDM(tconf), tconf).growth
    err_lib = np.abs(c.growth - tight).reshape(6, -1).max(axis=1)
    err_orc = np.abs(o.growth - tight).reshape(6, -1).max(axis=1)
    np.testing.assert_allclose(err_lib, err_orc, rtol=0.05)
    assert (np.abs(c.growth - o.growth).reshape(6, -1).max(axis=1) < 0.1 * err_orc).all()


def test_from_sigma8():
    conf = pa.Configuration(4.0, (16, 16, 16))
    c = pa.boltzmann(pa.Cosmology.from_sigma8(conf, 0.8, n_s=0.96, Omega_m=0.3,
                                              Omega_b=0.05, h=0.7), conf)
    assert abs(c.sigma8 - 0.8) < 1e-12


def test_growth_sensitivity_finite_difference():
    conf, c, _, _ = _cosmos()
    tight = conf.replace(growth_rtol=1e-12, growth_atol=1e-12)
    h = 1e-5
    gp = pa.boltzmann(pa.SimpleLCDM(tight, Omega_m=0.3 + h), tight).growth
    gm = pa.boltzmann(pa.SimpleLCDM(tight, Omega_m=0.3 - h), tight).growth
    fd = (gp - gm) / (2 * h)
    # the sensitivity rides on steps chosen for G alone (rtol 1.5e-8), so its
    # error is absolute, about 1e-7 of the table's O(1e-2..1) scale
    scale = np.abs(fd).max()
    np.testing.assert_allclose(c.growth_dOm[:, :, 1:], fd[:, :, 1:], rtol=2e-5, atol=1e-6 * scale)


def test_factors_match_oracle():
    # rel 1e-7: the factors read the growth tables (test_tables_match_oracle)
    conf, c, oconf, o = _cosmos()
    for a0, a1 in zip(conf.a_nbody[:-1], conf.a_nbody[1:]):
        ah = a0 * 0.5 + a1 * 0.5
        assert pnb.kick_factor(a0, a0, ah, c, conf) == pytest.approx(
            onb.kick_factor(a0, a0, ah, o, oconf), rel=1e-7)
        assert pnb.drift_factor(ah, a0, a1, c, conf) == pytest.approx(
            onb.drift_factor(ah, a0, a1, o, oconf), rel=1e-7)


def test_factor_derivatives_finite_difference():
    conf, c, _, _ = _cosmos()
    tight = conf.replace(growth_rtol=1e-12, growth_atol=1e-12)
    h = 1e-5
    cp = pa.boltzmann(pa.SimpleLCDM(tight, Omega_m=0.3 + h), tight)
    cm = pa.boltzmann(pa.SimpleLCDM(tight, Omega_m=0.3 - h), tight)
    for a0, a1 in [(0.1, 0.2), (0.5, 0.75), (0.9, 1.0)]:
        ah = 0.5 * (a0 + a1)
        for fn, dfn, args in [(pnb.kick_factor, pnb.kick_factor_dOm, (a0, a0, ah)),
                              (pnb.drift_factor, pnb.drift_factor_dOm, (ah, a0, a1))]:
            fd = (fn(*args, cp, tight) - fn(*args, cm, tight)) / (2 * h)
            assert dfn(*args, c, conf) == pytest.approx(fd, rel=1e-4)


def test_library_exports_header_symbols():
    """libpmx.so loads and exports every function include/pmx.h declares."""
    from pmwd_amd import _lib
    hdr = open(os.path.join(ROOT, 'include', 'pmx.h')).read()
    declared = set(re.findall(r'\b(pmx_[a-z0-9_]+)\s*\(', hdr))
    assert declared, 'no declarations parsed'
    L = _lib.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared == set(_lib.exported_symbols())
    assert L.pmx_version() >= 100


def test_library_rejects_bad_arguments():
    """Argument checks return PMX_EINVAL with a message, before any device work."""
    from pmwd_amd import _lib
    L = _lib.lib()
    out = ctypes.c_void_p()
    assert L.pmx_ctx_create(0, 4, 4, 1.0, None, ctypes.byref(out)) == 1
    assert b'positive' in L.pmx_last_error()
    assert L.pmx_ctx_create(4, 4, 4, -1.0, None, ctypes.byref(out)) == 1
    assert L.pmx_scatter_mass(None, None, None, 0, 1.0, None, 0) == 1


def test_product_does_not_import_oracle():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, 'pmwd_amd')
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith('.py'):
                src = open(os.path.join(dp, f)).read()
                assert 'oracle' not in re.findall(r'^\s*(?:from|import)\s+(\w+)', src, re.M), f


@pytest.mark.parametrize('shape', [(4, 6, 8), (5, 4, 7), (3, 3, 2)])
def test_fft_adjoint_formulas_match_torch_autograd(shape):
    """The r2c/c2r adjoints used for the library FFTs' backward equal torch's
    own autograd through torch.fft (CPU), for even and odd last axes."""
    import torch
    from pmwd_amd.pm_util import c2r_adjoint, r2c_adjoint
    g = torch.Generator().manual_seed(3)
    x = torch.randn(shape, dtype=torch.float64, generator=g, requires_grad=True)
    hshape = shape[:-1] + (shape[-1] // 2 + 1,)
    gc = torch.complex(torch.randn(hshape, dtype=torch.float64, generator=g),
                       torch.randn(hshape, dtype=torch.float64, generator=g))
    (want,) = torch.autograd.grad(torch.fft.rfftn(x), x, gc)
    got = r2c_adjoint(gc, shape, lambda h, s: torch.fft.irfftn(h, s=s))
    torch.testing.assert_close(got, want, rtol=1e-12, atol=1e-12)
    h = gc.clone().requires_grad_(True)
    y = torch.randn(shape, dtype=torch.float64, generator=g)
    (want,) = torch.autograd.grad(torch.fft.irfftn(h, s=shape), h, y)
    got = c2r_adjoint(y, shape, torch.fft.rfftn)
    torch.testing.assert_close(got, want, rtol=1e-12, atol=1e-12)


def test_fixed_sigma8_renormalisation():
    """at_fixed_sigma8 moves Omega_m and keeps sigma8 (from_sigma8 chain)."""
    import pmwd_amd as pa
    from pmwd_amd.cosmo_param import at_fixed_sigma8
    conf = pa.Configuration(4.0, (16, 16, 16), mesh_shape=2)
    c = pa.boltzmann(pa.Cosmology.from_sigma8(conf, 0.8, n_s=0.96, Omega_m=0.3,
                                              Omega_b=0.05, h=0.7), conf)
    c2 = at_fixed_sigma8(c, conf, 0.31)
    assert c2.Omega_m == 0.31
    assert abs(c2.sigma8 - c.sigma8) < 1e-12
    ref = pa.boltzmann(pa.Cosmology.from_sigma8(conf, c.sigma8, n_s=0.96, Omega_m=0.31,
                                                Omega_b=0.05, h=0.7), conf)
    assert abs(c2.A_s_1e9 / ref.A_s_1e9 - 1) < 1e-10


def test_serving_load_mix():
    """The serving load generator follows the stated mix and ranges."""
    from pmwd_amd.serving import c5_mix, context_bytes, request_bytes
    sched = c5_mix(4000, rate=2.0, seed=1)
    t = np.array([s[0] for s in sched])
    assert np.all(np.diff(t) > 0)
    assert abs(np.mean(np.diff(t)) - 0.5) < 0.03
    g = np.array([r.ptcl_grid for _, r in sched])
    for size, frac in ((64, 0.4), (128, 0.4), (256, 0.2)):
        assert abs(np.mean(g == size) - frac) < 0.03
    assert abs(np.mean([r.grad for _, r in sched]) - 0.5) < 0.03
    assert all(0.25 <= r.Omega_m <= 0.35 and 0.7 <= r.sigma8 <= 0.9 for _, r in sched)
    assert all(r.mesh_ratio == 2 and r.box == 2.0 * r.ptcl_grid for _, r in sched)
    small = next(r for _, r in sched if r.ptcl_grid == 64)
    big = next(r for _, r in sched if r.ptcl_grid == 256)
    assert request_bytes(big) > request_bytes(small)
    assert context_bytes((512, 512, 512), 256**3, True) > context_bytes((512, 512, 512), 256**3, False)


def test_vectorised_step_factors_equal_per_step():
    """kdk_factor_table evaluates every step's factors on arrays at once;
    the float32 results equal the per-step scalar evaluation bit for bit."""
    conf, c, _, _ = _cosmos()
    a = conf.a_nbody
    tab = pnb.kdk_factor_table(a[:-1], a[1:], c, conf)
    per = np.array([pnb._kdk_factors(a[i], a[i + 1], c, conf) for i in range(len(a) - 1)],
                   dtype=np.float32)
    np.testing.assert_array_equal(tab, per)


def test_checkpoint_steps():
    """nbody's checkpoint_every: every k-th step, or the steps given (as
    reversal_checkpoints returns them), within [1, nsteps - 1]."""
    assert sorted(pnb.checkpoint_steps(8, 31)) == [8, 16, 24]
    assert sorted(pnb.checkpoint_steps(8, 32)) == [8, 16, 24]
    assert sorted(pnb.checkpoint_steps(8, 33)) == [8, 16, 24, 32]
    assert pnb.checkpoint_steps(None, 31) == frozenset()
    assert pnb.checkpoint_steps(0, 31) == frozenset()
    assert pnb.checkpoint_steps([], 31) == frozenset()
    assert sorted(pnb.checkpoint_steps([30, 3, 3], 31)) == [3, 30]
    for bad in ([0], [31], -2):
        with pytest.raises(ValueError):
            pnb.checkpoint_steps(bad, 31)
