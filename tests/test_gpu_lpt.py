"""lpt() without gradients is one library call (pmx_lpt).  It must agree
with the oracle's float32 2LPT (pmwd/lpt.py:136-212) and with the
differentiable tensor-op path of the same function, on power-of-two and
other grids, at first and second order, for integer and non-integer mesh
ratios (non-zero gen_grid displacements)."""
import numpy as np
import pytest
import torch

from oracle import cosmo as oc, ic as oic
from oracle.conf import OConf

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    dt = np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64
    a, b = np.asarray(a, dt), np.asarray(b, dt)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


CASES = [((16, 16, 16), 2, 2), ((32, 32, 32), 1, 2), ((12, 10, 8), 1, 2), ((9, 8, 10), 2, 2),
         ((16, 16, 16), 1.5, 2), ((16, 8, 12), 1, 1)]


@pytest.mark.parametrize('shape,ratio,order', CASES)
def test_fused_lpt_matches_oracle_and_tensor_path(dev, shape, ratio, order):
    import pmwd_amd as pa
    kw = dict(mesh_shape=ratio, a_start=1 / 16, lpt_order=order)
    conf = pa.Configuration(3.0, shape, **kw)
    oconf = OConf(3.0, shape, **kw)
    cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
    ocosmo = oc.boltzmann(oc.SimpleLCDM(oconf), oconf)
    white = np.random.default_rng(sum(shape) + order).standard_normal(shape).astype(np.float32)
    print(f'case shape={shape} ratio={ratio} order={order}', flush=True)
    with torch.no_grad():
        modes = pa.linear_modes(torch.as_tensor(white, device=dev), cosmo, conf)
        fused, _ = pa.lpt(modes, cosmo, conf)
    # the tensor-op path (taken when the modes require gradients)
    m = modes.clone().requires_grad_(True)
    tens, _ = pa.lpt(m, cosmo, conf)
    omodes = oic.linear_modes(white, ocosmo, oconf)
    pmid, odisp, ovel = oic.lpt(omodes, ocosmo, oconf)
    np.testing.assert_array_equal(fused.pmid.cpu().numpy(), pmid)
    fd, fv = fused.disp.cpu().numpy(), fused.vel.cpu().numpy()
    td, tv = tens.disp.detach().cpu().numpy(), tens.vel.detach().cpu().numpy()
    res = dict(fused_vs_oracle=(rel_l2(fd, odisp), rel_l2(fv, ovel)),
               fused_vs_tensor=(rel_l2(fd, td), rel_l2(fv, tv)))
    print(res, flush=True)
    assert max(res['fused_vs_oracle']) < 1e-5, res
    assert max(res['fused_vs_tensor']) < 1e-5, res


def test_fused_lpt_second_order_term(dev):
    """The second-order part alone (order 2 minus order 1) against the
    oracle's: it is small next to the first order, so it gets its own check."""
    import pmwd_amd as pa
    shape = (32, 32, 32)
    out = {}
    for order in (1, 2):
        kw = dict(mesh_shape=1, a_start=1 / 2, lpt_order=order)
        conf, oconf = pa.Configuration(2.0, shape, **kw), OConf(2.0, shape, **kw)
        cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
        ocosmo = oc.boltzmann(oc.SimpleLCDM(oconf), oconf)
        white = np.random.default_rng(5).standard_normal(shape).astype(np.float32)
        with torch.no_grad():
            ptcl, _ = pa.lpt(pa.linear_modes(torch.as_tensor(white, device=dev), cosmo, conf),
                             cosmo, conf)
        _, od, ov = oic.lpt(oic.linear_modes(white, ocosmo, oconf), ocosmo, oconf)
        out[order] = (ptcl.disp.cpu().numpy().astype(np.float64), od.astype(np.float64),
                      ptcl.vel.cpu().numpy().astype(np.float64), ov.astype(np.float64))
    d2 = out[2][0] - out[1][0]
    od2 = out[2][1] - out[1][1]
    v2 = out[2][2] - out[1][2]
    ov2 = out[2][3] - out[1][3]
    assert np.linalg.norm(od2) > 1e-3 * np.linalg.norm(out[2][1])  # the term is there
    assert rel_l2(d2, od2) < 1e-3 and rel_l2(v2, ov2) < 1e-3, (rel_l2(d2, od2), rel_l2(v2, ov2))


@pytest.mark.parametrize('shape,a', [((32, 32, 32), None), ((12, 10, 8), 0.5), ((9, 8, 10), None)])
def test_fused_linear_modes(dev, shape, a):
    """linear_modes without gradients (one r2c, one pmx_linear_modes pass)
    against the differentiable tensor path and the oracle (modes.py:67-117)."""
    import pmwd_amd as pa
    conf = pa.Configuration(3.0, shape, mesh_shape=1)
    oconf = OConf(3.0, shape, mesh_shape=1)
    cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
    ocosmo = oc.boltzmann(oc.SimpleLCDM(oconf), oconf)
    white = np.random.default_rng(3).standard_normal(shape).astype(np.float32)
    w = torch.as_tensor(white, device=dev)
    with torch.no_grad():
        fused = pa.linear_modes(w, cosmo, conf, a=a).cpu().numpy()
        fused_real = pa.linear_modes(w, cosmo, conf, a=a, real=True).cpu().numpy()
    tens = pa.linear_modes(w.clone().requires_grad_(True), cosmo, conf, a=a)
    tens = tens.detach().cpu().numpy()
    want = oic.linear_modes(white, ocosmo, oconf, a=a)
    assert rel_l2(fused, tens) < 1e-7, rel_l2(fused, tens)
    assert rel_l2(fused, want) < 1e-6, rel_l2(fused, want)
    want_real = oic.linear_modes(white, ocosmo, oconf, a=a, real=True)
    assert rel_l2(fused_real, want_real) < 1e-6
