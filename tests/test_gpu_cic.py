"""GPU parity of the CIC kernels (keys, sort, scatter, gather, adjoints) with the oracle."""
import numpy as np
import pytest
import torch

from oracle import pm as opm
from oracle.conf import OConf

pytestmark = pytest.mark.gpu


def make_ptcl(ptcl_grid_shape, mesh_shape, cell, disp_std, seed=0):
    pmid, disp = opm.gen_grid(ptcl_grid_shape, mesh_shape, cell)
    rng = np.random.default_rng(seed)
    disp = (disp + disp_std * rng.standard_normal(disp.shape)).astype(np.float32)
    return pmid, disp


def conf_of(ptcl_grid_shape, mesh_ratio, spacing=1.0, **kw):
    import pmwd_amd as pa
    return pa.Configuration(spacing, ptcl_grid_shape, mesh_shape=mesh_ratio, **kw)


def to_ptcl(conf, pmid, disp, dev):
    import pmwd_amd as pa
    return pa.Particles(conf, torch.as_tensor(pmid, device=dev), torch.as_tensor(disp, device=dev))


SHAPES = [
    ((8, 8, 8), 1, 3.0),
    ((16, 16, 16), 2, 7.0),
    ((4, 9, 7), 1, 7.0),
    ((7, 8, 5), 2, 2.0),
    ((20, 12, 35), 2, 1.5),
    ((32, 32, 32), 1, 0.7),
]


@pytest.mark.parametrize('spacing', [2.0, 1.7])
@pytest.mark.parametrize('shape,ratio,std', SHAPES)
def test_cell_keys_bit_exact(dev, shape, ratio, std, spacing):
    from pmwd_amd.pm_util import cell_keys
    conf = conf_of(shape, ratio, spacing=spacing)
    pmid, disp = make_ptcl(shape, conf.mesh_shape, conf.cell_size, std * conf.cell_size, seed=1)
    ptcl = to_ptcl(conf, pmid, disp, dev)
    got = cell_keys(ptcl, conf).cpu().numpy()
    want = opm.cell_keys(pmid, disp, conf.cell_size, conf.mesh_shape)
    np.testing.assert_array_equal(got, want.astype(np.int64))


@pytest.mark.parametrize('spacing,ratio', [(1.0, 2), (1.0, 1), (2.0, 1), (1.5, 2), (0.7, 1)])
def test_cell_keys_at_cell_boundaries(dev, spacing, ratio):
    """disp at and one ulp either side of multiples of the cell size: the
    corner follows float32 disp/cell exactly (a product with the exact
    reciprocal when the cell is a power of two, a division otherwise)."""
    from pmwd_amd.pm_util import cell_keys
    shape = (8, 8, 8)
    conf = conf_of(shape, ratio, spacing=spacing)
    pmid, disp = make_ptcl(shape, conf.mesh_shape, conf.cell_size, 0.0)
    rng = np.random.default_rng(3)
    k = rng.integers(-40, 40, size=disp.shape).astype(np.float32)
    base = (k * np.float32(conf.cell_size)).astype(np.float32)
    step = rng.integers(-1, 2, size=disp.shape)
    d = np.where(step < 0, np.nextafter(base, np.float32(-np.inf)),
                 np.where(step > 0, np.nextafter(base, np.float32(np.inf)), base))
    d = d.astype(np.float32)
    got = cell_keys(to_ptcl(conf, pmid, d, dev), conf).cpu().numpy()
    want = opm.cell_keys(pmid, d, conf.cell_size, conf.mesh_shape)
    np.testing.assert_array_equal(got, want.astype(np.int64))


@pytest.mark.parametrize('n,kmax', [(1, 5), (1000, 7), (4096, 100), (70001, 1 << 20),
                                    (1 << 20, 3), (300000, (1 << 31) - 1)])
def test_sort_perm_stable(dev, n, kmax):
    from pmwd_amd.pm_util import sort_perm
    rng = np.random.default_rng(n)
    keys = rng.integers(0, kmax, size=n, dtype=np.int64)
    got = sort_perm(torch.as_tensor(keys, device=dev)).cpu().numpy()
    want = opm.sort_perm(keys).astype(np.int64)
    np.testing.assert_array_equal(got, want)


def test_sort_perm_of_cell_keys(dev):
    from pmwd_amd.pm_util import cell_keys, sort_perm
    conf = conf_of((32, 32, 32), 2)
    pmid, disp = make_ptcl((32, 32, 32), conf.mesh_shape, conf.cell_size, 4.0, seed=3)
    ptcl = to_ptcl(conf, pmid, disp, dev)
    got = sort_perm(cell_keys(ptcl, conf)).cpu().numpy()
    want = opm.sort_perm(opm.cell_keys(pmid, disp, conf.cell_size, conf.mesh_shape))
    np.testing.assert_array_equal(got, want.astype(np.int64))


def test_tile_keys(dev):
    from pmwd_amd.pm_util import tile_keys
    conf = conf_of((20, 12, 35), 2)
    pmid, disp = make_ptcl((20, 12, 35), conf.mesh_shape, conf.cell_size, 5.0, seed=4)
    got, tile = tile_keys(to_ptcl(conf, pmid, disp, dev), conf)
    want = opm.tile_keys(pmid, disp, conf.cell_size, conf.mesh_shape, tile)
    np.testing.assert_array_equal(got.cpu().numpy(), want.astype(np.int64))


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize('shape,ratio,std', SHAPES)
def test_scatter_mass(dev, shape, ratio, std):
    import pmwd_amd as pa
    conf = conf_of(shape, ratio)
    pmid, disp = make_ptcl(shape, conf.mesh_shape, conf.cell_size, std * conf.cell_size, seed=2)
    dens = pa.scatter(to_ptcl(conf, pmid, disp, dev), conf).cpu().numpy()
    want = opm.scatter(pmid, disp, conf.cell_size, conf.mesh_shape)
    assert dens.shape == want.shape
    assert rel_l2(dens, want) < 1e-6
    np.testing.assert_allclose(dens, want, rtol=1e-5, atol=1e-5)
    # mass conservation (pm_test.py:86-101): the mean density is 1
    assert abs(dens.astype(np.float64).sum() - conf.mesh_size) < 1e-4 * conf.mesh_size


@pytest.mark.parametrize('val', [1e-33, 3e-20, 1e30])
def test_scatter_fixed_point_extremes(dev, val):
    """The fixed-point tile accumulation at the ends of its scale range: a
    scale of 2^139 (val 1e-33) takes the double-precision conversion, 2^94
    and 2^-70 the single-precision one."""
    import pmwd_amd as pa
    conf = conf_of((16, 16, 16), 2)
    pmid, disp = make_ptcl((16, 16, 16), conf.mesh_shape, conf.cell_size, 3.0, seed=9)
    got = pa.scatter(to_ptcl(conf, pmid, disp, dev), conf, val=val).cpu().numpy()
    want = opm.scatter(pmid, disp, conf.cell_size, conf.mesh_shape, val=val)
    assert np.all(np.isfinite(got))
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=2e-6 * val)


@pytest.mark.parametrize('clustered', [False, True])
def test_scatter_sign_symmetry(dev, clustered):
    """A non-negative value accumulates in split 32-bit words (a carry word
    beside the low word), a negative one in 64-bit words; both hold the same
    integer sums, so scatter with -v is -(scatter with v) bit for bit.  The
    clustered case puts 32^3 particles within a few cells, so each cell takes
    thousands of contributions and the low words carry many times."""
    import pmwd_amd as pa
    conf = conf_of((32, 32, 32), 1)
    pmid, disp = make_ptcl((32, 32, 32), conf.mesh_shape, conf.cell_size, 1.3, seed=4)
    if clustered:
        rng = np.random.default_rng(8)
        target = np.array([5.3, 17.8, 30.1]) * conf.cell_size
        disp = (target - pmid.astype(np.float64) * conf.cell_size
                + 1.5 * conf.cell_size * rng.standard_normal(disp.shape)).astype(np.float32)
    ptcl = to_ptcl(conf, pmid, disp, dev)
    pos = pa.scatter(ptcl, conf, val=1.7).cpu().numpy()
    neg = pa.scatter(ptcl, conf, val=-1.7).cpu().numpy()
    np.testing.assert_array_equal(neg, -pos)
    want = opm.scatter(pmid, disp, conf.cell_size, conf.mesh_shape, val=1.7, dtype=np.float64)
    assert rel_l2(pos, want) < 1e-6


def test_scatter_accumulates_into_mesh(dev):
    import pmwd_amd as pa
    conf = conf_of((16, 16, 16), 1)
    pmid, disp = make_ptcl((16, 16, 16), conf.mesh_shape, conf.cell_size, 2.0, seed=5)
    base = np.random.default_rng(0).standard_normal(conf.mesh_shape).astype(np.float32)
    got = pa.scatter(to_ptcl(conf, pmid, disp, dev), conf,
                     mesh=torch.as_tensor(base, device=dev)).cpu().numpy()
    want = opm.scatter(pmid, disp, conf.cell_size, conf.mesh_shape, mesh=base)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize('pos', [(-3., 5., 7.), (3., -5., 7.), (3., 5., -7.), (1., 1., 1.)])
def test_scatter_centered_ptcl(dev, pos):
    """pm_test.py:43-71 restated: odd offsets on a 2^3 mesh of cell 2."""
    import pmwd_amd as pa
    n = 7
    conf = pa.Configuration(2.0, (1, 1, 1), mesh_shape=(2, 2, 2))
    assert conf.cell_size == 1.0
    conf = pa.Configuration(4.0, (1, 1, 1), mesh_shape=(2, 2, 2))
    assert conf.cell_size == 2.0
    pmid = torch.zeros((n, 3), dtype=torch.int16, device=dev)
    disp = torch.tensor([pos] * n, dtype=torch.float32, device=dev)
    mesh = pa.scatter(pa.Particles(conf, pmid, disp), conf, val=1.0)
    np.testing.assert_array_equal(mesh.cpu().numpy(), np.full((2, 2, 2), n * 2.0**-3, np.float32))


def test_gather_uniform(dev):
    """pm_test.py:103-118 restated: a mesh of ones gathers to 1.

    In float32 the eight rounded fractions sum to 1 within 2 ulp (exactly 1
    for particles on cell boundaries); the result equals the oracle bit for bit.
    """
    import pmwd_amd as pa
    conf = conf_of((5, 5, 5), 1, spacing=2.0)
    pmid, disp = make_ptcl((5, 5, 5), conf.mesh_shape, conf.cell_size, 7.0, seed=6)
    ones = np.ones(conf.mesh_shape, np.float32)
    val = pa.gather(to_ptcl(conf, pmid, disp, dev), conf,
                    torch.as_tensor(ones, device=dev)).cpu().numpy()
    np.testing.assert_allclose(val, 1.0, rtol=0, atol=2.5e-7)
    np.testing.assert_array_equal(val, opm.gather(pmid, disp, conf.cell_size, conf.mesh_shape, ones))
    on_grid = np.zeros_like(disp)
    val = pa.gather(to_ptcl(conf, pmid, on_grid, dev), conf, torch.as_tensor(ones, device=dev))
    np.testing.assert_array_equal(val.cpu().numpy(), np.ones(len(pmid), np.float32))


@pytest.mark.parametrize('shape,ratio,std', SHAPES)
def test_gather_bit_exact(dev, shape, ratio, std):
    import pmwd_amd as pa
    conf = conf_of(shape, ratio)
    pmid, disp = make_ptcl(shape, conf.mesh_shape, conf.cell_size, std * conf.cell_size, seed=7)
    mesh = np.random.default_rng(1).standard_normal(conf.mesh_shape).astype(np.float32)
    got = pa.gather(to_ptcl(conf, pmid, disp, dev), conf, torch.as_tensor(mesh, device=dev))
    want = opm.gather(pmid, disp, conf.cell_size, conf.mesh_shape, mesh)
    np.testing.assert_array_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize('offset,out_cell,chan', [
    (0.0, None, (2,)), (0.37, None, ()), ((0.5, -1.25, 2.0), None, (3,)),
    (0.0, 0.5, ()), (0.3, 0.75, (2,)), (0.0, 1.5, ()),
])
def test_generic_scatter_gather(dev, offset, out_cell, chan):
    import pmwd_amd as pa
    conf = conf_of((8, 9, 6), 1)
    pmid, disp = make_ptcl((8, 9, 6), conf.mesh_shape, conf.cell_size, 3.0, seed=8)
    ptcl = to_ptcl(conf, pmid, disp, dev)
    out_shape = conf.mesh_shape if out_cell is None else tuple(
        int(np.ceil(s * conf.cell_size / out_cell)) for s in conf.mesh_shape)
    rng = np.random.default_rng(2)
    val = rng.standard_normal((len(pmid),) + chan).astype(np.float32) if chan else None
    mesh0 = np.zeros(out_shape + chan, np.float32)
    got = pa.scatter(ptcl, conf, mesh=torch.as_tensor(mesh0, device=dev),
                     val=None if val is None else torch.as_tensor(val, device=dev),
                     offset=offset, cell_size=out_cell).cpu().numpy()
    want = opm.scatter(pmid, disp, conf.cell_size, conf.mesh_shape, mesh=mesh0, val=val,
                       offset=offset, out_cell_size=out_cell)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)
    m = rng.standard_normal(out_shape + chan).astype(np.float32)
    v0 = np.zeros((len(pmid),) + chan, np.float32) if chan else 0
    g = pa.gather(ptcl, conf, torch.as_tensor(m, device=dev),
                  val=torch.as_tensor(v0, device=dev) if chan else 0,
                  offset=offset, cell_size=out_cell).cpu().numpy()
    w = opm.gather(pmid, disp, conf.cell_size, conf.mesh_shape, m, val=v0, offset=offset,
                   out_cell_size=out_cell)
    np.testing.assert_allclose(g, w, rtol=1e-6, atol=1e-6)


def test_scatter_gather_vjp(dev):
    import pmwd_amd as pa
    conf = conf_of((9, 8, 7), 1)
    pmid, disp = make_ptcl((9, 8, 7), conf.mesh_shape, conf.cell_size, 2.0, seed=9)
    rng = np.random.default_rng(3)
    mc = rng.standard_normal(conf.mesh_shape).astype(np.float32)
    d = torch.as_tensor(disp, device=dev).requires_grad_(True)
    ptcl = pa.Particles(conf, torch.as_tensor(pmid, device=dev), d)
    dens = pa.scatter(ptcl, conf)
    (dc,) = torch.autograd.grad(dens, d, torch.as_tensor(mc, device=dev))
    want, _ = opm.scatter_adj(pmid, disp, conf.cell_size, conf.mesh_shape, mc)
    np.testing.assert_allclose(dc.cpu().numpy(), want, rtol=1e-5, atol=1e-5)

    mesh = torch.as_tensor(rng.standard_normal(conf.mesh_shape).astype(np.float32),
                           device=dev).requires_grad_(True)
    vc = rng.standard_normal(len(pmid)).astype(np.float32)
    d2 = torch.as_tensor(disp, device=dev).requires_grad_(True)
    out = pa.gather(pa.Particles(conf, torch.as_tensor(pmid, device=dev), d2), conf, mesh)
    gd, gm = torch.autograd.grad(out, (d2, mesh), torch.as_tensor(vc, device=dev))
    wd, wm, _ = opm.gather_adj(pmid, disp, conf.cell_size, conf.mesh_shape,
                               mesh.detach().cpu().numpy(), vc)
    np.testing.assert_allclose(gd.cpu().numpy(), wd, rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(gm.cpu().numpy(), wm, rtol=1e-5, atol=1e-5)


def test_empty_particles(dev):
    import pmwd_amd as pa
    conf = conf_of((8, 8, 8), 1)
    ptcl = pa.Particles(conf, torch.zeros((0, 3), dtype=torch.int16, device=dev),
                        torch.zeros((0, 3), device=dev))
    m = pa.scatter(ptcl, conf, val=1.0)
    assert float(m.abs().sum()) == 0.0
    v = pa.gather(ptcl, conf, torch.ones(conf.mesh_shape, device=dev))
    assert v.shape == (0,)
