"""The HIP path against the committed golden fixtures (tests/golden)."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize('case', ['cic_even', 'cic_odd'])
def test_cic_fixture(dev, case):
    import pmwd_amd as pa
    from pmwd_amd.pm_util import cell_keys, sort_perm
    z = np.load(os.path.join(GOLDEN, f'{case}.npz'))
    shape = tuple(int(s) for s in z['ptcl_shape'])
    ratio = int(z['mesh_shape'][0]) // shape[0]
    cell = float(z['cell_size'])
    conf = pa.Configuration(cell * ratio, shape, mesh_shape=ratio)
    assert conf.mesh_shape == tuple(int(s) for s in z['mesh_shape'])
    P = lambda a: torch.as_tensor(np.ascontiguousarray(a), device=dev)
    ptcl = pa.Particles(conf, P(z['pmid']), P(z['disp']))
    keys = cell_keys(ptcl, conf)
    np.testing.assert_array_equal(keys.cpu().numpy().astype(np.uint32), z['keys'])
    np.testing.assert_array_equal(sort_perm(keys).cpu().numpy().astype(np.uint32), z['perm'])
    dens = pa.scatter(ptcl, conf).cpu().numpy()
    assert rel_l2(dens, z['dens']) < 1e-6
    g = pa.gather(ptcl, conf, P(z['mesh'])).cpu().numpy()
    np.testing.assert_array_equal(g, z['gather'])
    # adjoints through autograd of the drop-in operators
    d = P(z['disp']).requires_grad_(True)
    m = P(z['mesh']).requires_grad_(True)
    out = pa.gather(pa.Particles(conf, P(z['pmid']), d), conf, m)
    gd, gm = torch.autograd.grad(out, (d, m), P(z['val_cot']))
    np.testing.assert_allclose(gd.cpu().numpy(), z['gather_adj_disp'], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(gm.cpu().numpy(), z['gather_adj_mesh'], rtol=1e-5, atol=1e-5)
    d = P(z['disp']).requires_grad_(True)
    val = torch.full((len(z['pmid']),), float(z['val']), device=dev, requires_grad=True)
    dens = pa.scatter(pa.Particles(conf, P(z['pmid']), d), conf, val=val)
    gd, gv = torch.autograd.grad(dens, (d, val), P(z['mesh_cot']))
    np.testing.assert_allclose(gd.cpu().numpy(), z['scatter_adj_disp'], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(gv.cpu().numpy(), z['scatter_adj_val'], rtol=1e-5, atol=1e-5)


def test_run_fixture(dev):
    """White noise -> linear modes -> 2LPT -> 4 KDK steps (16^3) against the
    oracle's float32 run; the float64 error of both is printed beside."""
    import pmwd_amd as pa
    z = np.load(os.path.join(GOLDEN, 'run16.npz'))
    conf = pa.Configuration(4.0, (16, 16, 16), mesh_shape=1, a_start=1 / 16, a_stop=5 / 16,
                            a_nbody_maxstep=1 / 16)
    assert len(conf.a_nbody) == 5
    cosmo = pa.boltzmann(pa.SimpleLCDM(conf), conf)
    white = torch.as_tensor(z['white'], device=dev)
    # the device generator reproduces the fixture's noise (seed 0) within an
    # ulp, at most a few elements differing
    dw = pa.white_noise(0, conf, real=True, device=dev).cpu().numpy()
    diff = np.abs(dw.astype(np.float64) - z['white'])
    assert np.all(diff <= np.spacing(np.abs(z['white']))) and np.count_nonzero(diff) <= 2
    modes = pa.linear_modes(white, cosmo, conf)
    ptcl, _ = pa.lpt(modes, cosmo, conf)
    np.testing.assert_array_equal(ptcl.pmid.cpu().numpy(), z['pmid'])
    assert rel_l2(ptcl.disp.cpu().numpy(), z['lpt_disp_f32']) < 1e-5
    assert rel_l2(ptcl.vel.cpu().numpy(), z['lpt_vel_f32']) < 1e-5
    out, _ = pa.nbody(ptcl, None, cosmo, conf)
    res = {}
    for k in ('disp', 'vel', 'acc'):
        got = getattr(out, k).cpu().numpy()
        res[k] = (rel_l2(got, z[f'{k}_f32']), rel_l2(got, z[f'{k}_f64']),
                  rel_l2(z[f'{k}_f32'], z[f'{k}_f64']))
    print({k: 'vs f32 %.2e, vs f64 %.2e (oracle f32 vs f64 %.2e)' % v for k, v in res.items()})
    assert res['disp'][0] < 1e-5 and res['vel'][0] < 1e-4 and res['acc'][0] < 1e-4
    # the float64 error is that of a float32 run, not larger
    for k, (e32, e64, o64) in res.items():
        assert e64 < 3 * o64 + 1e-6, k
