"""Particle-mesh operators restated from pmwd in numpy (oracle side).

Follows pmwd/pm_util.py (enmesh 33-156, fftfreq 159-199, fftfwd/fftinv
236-344), pmwd/scatter.py (8-150), pmwd/gather.py (8-144), pmwd/gravity.py
(9-72) and pmwd/particles.py (gen_grid 109-144, raveled_id 156-182).

Arithmetic mirrors the reference dtypes: int16 indices with floor-mod wrap,
float32 division ``disp / cell``, float32 multilinear fractions multiplied as
``(f0*f1)*f2``.  Two choices are the oracle's own and are stated where they
are made: scatter accumulates with ``np.bincount`` (float64 sums, then cast),
and the 8-neighbour gather sum uses the fixed tree order
``((v0+v1)+(v2+v3))+((v4+v5)+(v6+v7))``.  FFTs are scipy/pocketfft in the
input precision.
"""
import numpy as np
import scipy.fft as sfft


def neighbors(dim, dtype=np.int16):
    """Neighbour offsets, bit j>>d of neighbour j on axis d (pm_util.py:95-97)."""
    return ((np.arange(2**dim)[:, None] >> np.arange(dim)) & 1).astype(dtype)


def enmesh(i1, d1, a1, s1, b12, a2, s2, grad=False):
    """Multilinear indices and fractions (pm_util.py:33-156), numpy restatement."""
    i1 = np.asarray(i1)
    d1 = np.asarray(d1)
    idt = i1.dtype
    fdt = d1.dtype
    dim = i1.shape[1]
    nb = neighbors(dim, idt)

    if a2 is not None:
        a1 = np.float64(a1)
        b12 = np.float64(b12)
        a2 = np.float64(a2)
        P = i1 * a1 + d1 - b12
        P = P[:, None]
        i2 = P + nb * a2
        if s1 is not None:
            L = np.asarray(s1, dtype=idt) * a1
            i2 = np.mod(i2, L)
        i2 = np.floor_divide(i2, a2)
        d2 = P - i2 * a2
        if s1 is not None:
            d2 = d2 - np.rint(d2 / L) * L
        i2 = i2.astype(idt)
        d2 = d2.astype(fdt)
        d2 = d2 / fdt.type(a2)
    else:
        a1 = fdt.type(a1)
        i12, d12 = np.divmod(np.float64(b12), np.float64(a1))
        i1 = i1 - idt.type(i12)
        d1 = d1 - fdt.type(d12)
        i1 = i1[:, None]
        d1 = d1[:, None] / a1
        i2 = np.floor(d1).astype(idt)
        i2 = (i2 + nb).astype(idt)
        d2 = d1 - i2.astype(fdt)
        i2 = (i2 + i1).astype(idt)
        if s1 is not None:
            i2 = np.mod(i2, np.asarray(s1, dtype=idt)).astype(idt)

    f2 = (1 - np.abs(d2)).astype(fdt)

    if s1 is None and s2 is not None:
        i2 = np.where(i2 < 0, np.asarray(s2, dtype=idt), i2).astype(idt)

    if grad:
        sign = np.sign(-d2).astype(fdt)
        f2g = []
        for i in range(dim):
            not_i = tuple(range(i + 1, dim)) + tuple(range(0, i))
            p = f2[..., not_i[0]] if not_i else np.ones_like(f2[..., 0])
            for j in not_i[1:]:
                p = p * f2[..., j]
            f2g.append(sign[..., i] * p)
        f2g = np.stack(f2g, axis=-1)
    w = f2[..., 0]
    for d in range(1, dim):
        w = w * f2[..., d]
    if grad:
        return i2, w, f2g
    return i2, w


def raveled(ind, shape):
    """C-order raveled index with strides (particles.py:156-182)."""
    out = np.zeros(ind.shape[:-1], dtype=np.int64)
    for d, s in enumerate(shape):
        out = out * s + ind[..., d].astype(np.int64)
    return out


def cell_keys(pmid, disp, cell_size, mesh_shape):
    """Raveled C-order key of each particle's lower CIC corner (neighbour 0).

    Equals ``raveled_id`` of ``enmesh(...)[0][:, 0]`` (pm_util.py:119-137 with
    particles.py:156-182 strides), as uint32.
    """
    ind, _ = enmesh(pmid, disp, cell_size, mesh_shape, 0, None, mesh_shape)
    return raveled(ind[:, 0], mesh_shape).astype(np.uint32)


def sort_perm(keys):
    """Stable ascending sort permutation of the cell keys."""
    return np.argsort(np.asarray(keys), kind='stable').astype(np.uint32)


def tile_keys(pmid, disp, cell_size, mesh_shape, tile):
    """Raveled index of the mesh tile holding each particle's lower corner.

    Tiles are ``tile``-shaped blocks of cells in C order, ``ceil(mesh/tile)``
    per axis.  This is the binning key of the HIP scatter.
    """
    ind, _ = enmesh(pmid, disp, cell_size, mesh_shape, 0, None, mesh_shape)
    lo = ind[:, 0].astype(np.int64)
    ntile = [-(-m // t) for m, t in zip(mesh_shape, tile)]
    out = np.zeros(len(lo), dtype=np.int64)
    for d in range(len(mesh_shape)):
        out = out * ntile[d] + lo[:, d] // tile[d]
    return out.astype(np.uint32)


def _tree8(v):
    """Sum over axis 1 (length 2^dim) in the pairwise tree order documented above."""
    n = v.shape[1]
    if n == 1:
        return v[:, 0]
    if n == 2:
        return v[:, 0] + v[:, 1]
    if n == 4:
        return (v[:, 0] + v[:, 1]) + (v[:, 2] + v[:, 3])
    return ((v[:, 0] + v[:, 1]) + (v[:, 2] + v[:, 3])) + ((v[:, 4] + v[:, 5]) + (v[:, 6] + v[:, 7]))


def _valid(ind, shape):
    ok = np.ones(ind.shape[:-1], dtype=bool)
    for d, s in enumerate(shape):
        ok &= (ind[..., d] >= 0) & (ind[..., d] < s)
    return ok


def _chunks(n, chunk):
    if chunk is None or n == 0:
        yield 0, n
        return
    for s in range(0, n, chunk):
        yield s, min(n, s + chunk)


def scatter(pmid, disp, cell_size, mesh_shape, mesh=None, val=None, offset=0,
            out_cell_size=None, ptcl_num=None, dtype=np.float32, chunk=2**22):
    """CIC mass assignment (scatter.py:33-83).

    ``val`` defaults to ``mesh_size / ptcl_num`` so the mean density is 1.
    ``mesh`` defaults to zeros of ``mesh_shape + val.shape[1:]``.  Products
    ``val * frac`` are formed in ``dtype``; the sum over particles uses
    ``np.bincount`` (float64) and is cast back to ``dtype``.
    """
    pmid = np.asarray(pmid)
    disp = np.asarray(disp, dtype=dtype)
    n, dim = pmid.shape
    if ptcl_num is None:
        ptcl_num = n
    if val is None:
        val = int(np.prod(mesh_shape)) / ptcl_num
    val = np.asarray(val, dtype=dtype)
    if mesh is None:
        mesh = np.zeros(tuple(mesh_shape) + val.shape[1:], dtype=dtype)
    mesh = np.asarray(mesh, dtype=dtype)
    if mesh.shape[dim:] != val.shape[1:]:
        raise ValueError(f'channel shape mismatch: {mesh.shape[dim:]} != {val.shape[1:]}')
    sshape = mesh.shape[:dim]
    chan = mesh.shape[dim:]
    C = int(np.prod(chan)) if chan else 1
    M = int(np.prod(sshape))
    acc = mesh.reshape(M, C).astype(np.float64)
    for s, e in _chunks(n, chunk):
        ind, frac = enmesh(pmid[s:e], disp[s:e], cell_size, mesh_shape, offset,
                           out_cell_size, sshape)
        ok = _valid(ind, sshape)
        flat = raveled(np.where(ok[..., None], ind, 0), sshape)[ok]
        if val.ndim == 0:
            contrib = (val * frac)[ok][:, None]
        else:
            v = val[s:e].reshape(e - s, 1, C)
            contrib = (v * frac[..., None])[ok]
        for c in range(C):
            acc[:, c] += np.bincount(flat, weights=contrib[:, c].astype(np.float64), minlength=M)
    return acc.astype(dtype).reshape(mesh.shape)


def gather(pmid, disp, cell_size, mesh_shape, mesh, val=0, offset=0,
           out_cell_size=None, dtype=np.float32, chunk=2**22):
    """CIC readout (gather.py:33-77); out-of-bounds neighbours read 0."""
    pmid = np.asarray(pmid)
    disp = np.asarray(disp, dtype=dtype)
    n, dim = pmid.shape
    mesh = np.asarray(mesh, dtype=dtype)
    val = np.asarray(val, dtype=dtype)
    if mesh.shape[dim:] != val.shape[1:]:
        raise ValueError(f'channel shape mismatch: {mesh.shape[dim:]} != {val.shape[1:]}')
    sshape = mesh.shape[:dim]
    chan = mesh.shape[dim:]
    C = int(np.prod(chan)) if chan else 1
    flatmesh = mesh.reshape(-1, C)
    out = np.empty((n, C), dtype=dtype)
    for s, e in _chunks(n, chunk):
        ind, frac = enmesh(pmid[s:e], disp[s:e], cell_size, mesh_shape, offset,
                           out_cell_size, sshape)
        ok = _valid(ind, sshape)
        flat = raveled(np.where(ok[..., None], ind, 0), sshape)
        m = np.where(ok[..., None], flatmesh[flat], 0).astype(dtype)   # (n, 8, C)
        out[s:e] = _tree8(m * frac[..., None])
    out = out.reshape((n,) + chan)
    return (val + out).astype(dtype)


def scatter_adj(pmid, disp, cell_size, mesh_shape, mesh_cot, val=None, offset=0,
                out_cell_size=None, ptcl_num=None, dtype=np.float32):
    """VJP of scatter (scatter.py:86-148): returns (disp_cot, val_cot)."""
    pmid = np.asarray(pmid)
    disp = np.asarray(disp, dtype=dtype)
    n, dim = pmid.shape
    if ptcl_num is None:
        ptcl_num = n
    if val is None:
        val = int(np.prod(mesh_shape)) / ptcl_num
    val = np.asarray(val, dtype=dtype)
    mesh_cot = np.asarray(mesh_cot, dtype=dtype)
    sshape = mesh_cot.shape[:dim]
    chan = mesh_cot.shape[dim:]
    C = int(np.prod(chan)) if chan else 1
    ind, frac, fg = enmesh(pmid, disp, cell_size, mesh_shape, offset, out_cell_size,
                           sshape, grad=True)
    ok = _valid(ind, sshape)
    flat = raveled(np.where(ok[..., None], ind, 0), sshape)
    vc = np.where(ok[..., None], mesh_cot.reshape(-1, C)[flat], 0).astype(dtype)  # (n,8,C)
    v = val.reshape(-1, 1, C) if val.ndim else val
    dc = (vc * v).sum(axis=-1)                                   # (n,8)
    disp_cot = (dc[..., None] * fg).sum(axis=1)
    disp_cot = disp_cot / dtype(out_cell_size if out_cell_size is not None else cell_size)
    val_cot = (vc * frac[..., None]).sum(axis=1).reshape((n,) + chan)
    return disp_cot.astype(dtype), val_cot.astype(dtype)


def gather_adj(pmid, disp, cell_size, mesh_shape, mesh, val_cot, offset=0,
               out_cell_size=None, dtype=np.float32):
    """VJP of gather (gather.py:80-142): returns (disp_cot, mesh_cot, val_cot)."""
    pmid = np.asarray(pmid)
    disp = np.asarray(disp, dtype=dtype)
    n, dim = pmid.shape
    mesh = np.asarray(mesh, dtype=dtype)
    sshape = mesh.shape[:dim]
    chan = mesh.shape[dim:]
    C = int(np.prod(chan)) if chan else 1
    val_cot = np.asarray(val_cot, dtype=dtype)
    ind, frac, fg = enmesh(pmid, disp, cell_size, mesh_shape, offset, out_cell_size,
                           sshape, grad=True)
    ok = _valid(ind, sshape)
    flat = raveled(np.where(ok[..., None], ind, 0), sshape)
    mv = np.where(ok[..., None], mesh.reshape(-1, C)[flat], 0).astype(dtype)   # (n,8,C)
    vcot = val_cot.reshape(n, 1, C)
    dc = (vcot * mv).sum(axis=-1)
    disp_cot = (dc[..., None] * fg).sum(axis=1)
    disp_cot = disp_cot / dtype(out_cell_size if out_cell_size is not None else cell_size)
    contrib = (vcot * frac[..., None])[ok]                        # (k, C)
    M = int(np.prod(sshape))
    mc = np.zeros((M, C))
    for c in range(C):
        mc[:, c] = np.bincount(flat[ok], weights=contrib[:, c].astype(np.float64), minlength=M)
    return disp_cot.astype(dtype), mc.astype(dtype).reshape(mesh.shape), val_cot


# ---------------- FFT and Poisson (pm_util.py, gravity.py) ----------------

def _jax_fftfreq(n, rfft=False):
    """jnp.fft.fftfreq / rfftfreq: integer frequencies divided by n in
    float64 (numpy's multiplies by 1/n instead, which can differ in the last
    bit when n is not a power of two)."""
    if rfft:
        j = np.arange(n // 2 + 1, dtype=np.float64)
    else:
        j = np.concatenate([np.arange(0, (n - 1) // 2 + 1), np.arange(-(n // 2), 0)]).astype(np.float64)
    return j / n


def fftfreq(shape, spacing, dtype=np.float64):
    """Sparse angular wavevectors (pm_util.py:159-199), computed in float64."""
    period = 1.0 if spacing is None else 2 * np.pi / spacing
    kvec = []
    for s in shape[:-1]:
        kvec.append((_jax_fftfreq(s) * period).astype(dtype))
    kvec.append((_jax_fftfreq(shape[-1], rfft=True) * period).astype(dtype))
    return np.meshgrid(*kvec, sparse=True, indexing='ij')


def fftfwd(f, shape=None, axes=None, norm=None):
    """pm_util.py:236-289."""
    f = np.asarray(f)
    if np.iscomplexobj(f):
        raise ValueError('input field must be real')
    if norm in (None, 'backward', 'ortho', 'forward'):
        return sfft.rfftn(f, s=shape, axes=axes, norm=norm)
    d = f.ndim if shape is None else len(shape)
    if axes is not None:
        d = len(axes)
    out = sfft.rfftn(f, s=shape, axes=axes, norm='backward')
    return (out * out.real.dtype.type(norm**d)).astype(out.dtype)


def fftinv(f, shape=None, axes=None, norm=None):
    """pm_util.py:292-344."""
    f = np.asarray(f)
    if not np.iscomplexobj(f):
        raise ValueError('input field must be Hermitian complex')
    if norm in (None, 'backward', 'ortho', 'forward'):
        return sfft.irfftn(f, s=shape, axes=axes, norm=norm)
    d = f.ndim if shape is None else len(shape)
    if axes is not None:
        d = len(axes)
    out = sfft.irfftn(f, s=shape, axes=axes, norm='backward')
    return (out * out.dtype.type(norm**-d)).astype(out.dtype)


def laplace(kvec, src):
    """-src/k^2 with the k=0 mode zeroed (gravity.py:9-16); self-adjoint."""
    k2 = 0
    for k in kvec:
        k2 = k2 + k**2
    with np.errstate(divide='ignore', invalid='ignore'):
        re = np.where(k2 != 0, (-src.real) / k2, 0)
        im = np.where(k2 != 0, (-src.imag) / k2, 0)
    out = np.empty(np.broadcast_shapes(src.shape, np.shape(k2)), dtype=src.dtype)
    out.real = re
    out.imag = im
    return out


def nyquist_mask(k, spacing):
    """True where |k| is the Nyquist wavenumber (gravity.py:38-40)."""
    nyq = np.pi / spacing
    eps = nyq * np.finfo(k.dtype).eps
    return np.abs(np.abs(k) - k.dtype.type(nyq)) <= eps


def neg_grad(k, pot, spacing):
    """(-i k) pot with Nyquist planes zeroed (gravity.py:37-44).

    The complex product with ``(0, -k)`` is exact: re = k*im(pot),
    im = -(k*re(pot)).
    """
    km = np.where(nyquist_mask(k, spacing), k.dtype.type(0), k)
    out = np.empty(np.broadcast_shapes(pot.shape, np.shape(km)), dtype=pot.dtype)
    out.real = km * pot.imag
    out.imag = -(km * pot.real)
    return out


# ------- default-argument CIC shared by one force (scatter + 3 gathers) -------

def cic_default(pmid, disp, cell_size, mesh_shape, grad=False):
    """enmesh's default branch (pm_util.py:119-155 with offset 0, a2 None,
    periodic ``mesh_shape``) computed per axis and shared by the scatter and
    the gathers of one force.

    Same operations and dtypes as ``enmesh``: ``d = disp / cell`` in float32,
    ``floor`` cast to int16, ``d - float(i + bit)``, the int16 sum with
    ``pmid`` and the floor-mod wrap, ``1 - |.|`` and the products
    ``(f0*f1)*f2`` over neighbour bits ``j >> axis`` -- so every index and
    fraction equals ``enmesh``'s bit for bit (checked in the oracle tests);
    it only avoids materialising the (n, 8, 3) intermediates.  Returns the
    raveled C-order indices (n, 8) int64, the weights (n, 8) and, with
    ``grad``, the weight gradients (n, 8, 3) of ``enmesh(grad=True)``.
    """
    pmid = np.asarray(pmid)
    disp = np.asarray(disp)
    idt, fdt = pmid.dtype, disp.dtype
    dim = pmid.shape[1]
    a1 = fdt.type(cell_size)
    d1 = disp / a1
    lo = np.floor(d1).astype(idt)
    idx, frac, sign = [], [], []
    for ax in range(dim):
        s = idt.type(mesh_shape[ax])
        i_ax, f_ax, g_ax = [], [], []
        for bit in (0, 1):
            i2 = (lo[:, ax] + idt.type(bit)).astype(idt)
            d2 = d1[:, ax] - i2.astype(fdt)
            i2 = np.mod((i2 + pmid[:, ax]).astype(idt), s).astype(idt)
            i_ax.append(i2)
            f_ax.append((1 - np.abs(d2)).astype(fdt))
            g_ax.append(np.sign(-d2).astype(fdt))
        idx.append(i_ax)
        frac.append(f_ax)
        sign.append(g_ax)
    n = len(pmid)
    nb = neighbors(dim, np.int64)
    flat = np.empty((n, 2**dim), dtype=np.int64)
    w = np.empty((n, 2**dim), dtype=fdt)
    f2g = np.empty((n, 2**dim, dim), dtype=fdt) if grad else None
    for j in range(2**dim):
        b = nb[j]
        r = np.zeros(n, dtype=np.int64)
        for ax in range(dim):
            r = r * int(mesh_shape[ax]) + idx[ax][b[ax]].astype(np.int64)
        flat[:, j] = r
        wj = frac[0][b[0]]
        for ax in range(1, dim):
            wj = wj * frac[ax][b[ax]]
        w[:, j] = wj
        if grad:
            for i in range(dim):
                not_i = tuple(range(i + 1, dim)) + tuple(range(0, i))
                p = frac[not_i[0]][b[not_i[0]]] if not_i else np.ones(n, fdt)
                for k in not_i[1:]:
                    p = p * frac[k][b[k]]
                f2g[:, j, i] = sign[i][b[i]] * p
    return flat, w, f2g


def _pool_map(fn, n, chunk=1 << 18):
    """fn(s, e) over particle chunks on a thread pool (numpy releases the GIL
    in the array work); results in chunk order."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    spans = [(s, min(n, s + chunk)) for s in range(0, n, chunk)] or [(0, 0)]
    workers = max(1, min(len(spans), os.cpu_count() or 1, 16))
    if workers == 1:
        return [fn(s, e) for s, e in spans]
    with ThreadPoolExecutor(workers) as ex:
        return list(ex.map(lambda se: fn(*se), spans))


def _cic_chunks(pmid, disp, cell_size, mesh_shape, grad=False):
    return _pool_map(lambda s, e: (s, e) + cic_default(pmid[s:e], disp[s:e], cell_size,
                                                        mesh_shape, grad), len(pmid))


def _scatter_cic(cics, M, contrib_fn, dtype):
    """Sum of contrib_fn(s, e, w) (n_chunk, 8) at the shared indices: float64
    bincount per chunk, chunks added in order, cast to ``dtype``."""
    parts = _pool_map(lambda i, _e: np.bincount(
        cics[i][2].ravel(), weights=contrib_fn(*cics[i][:2], cics[i][3]).ravel().astype(np.float64),
        minlength=M), len(cics), chunk=1)
    acc = np.zeros(M)
    for p in parts:
        acc += p
    return acc.astype(dtype)


def _gather_cic(cics, flatmesh, n, dtype, C=1):
    out = np.empty((n, C), dtype=dtype)

    def one(i, _e):
        s, e, flat, w = cics[i][:4]
        m = flatmesh[flat].astype(dtype)          # (k, 8, C)
        out[s:e] = _tree8(m * w[..., None])
    _pool_map(one, len(cics), chunk=1)
    return out


def gravity_mesh(dens, omega_m, cell_size, mesh_shape, dtype=np.float32):
    """Three force meshes from a density mesh (gravity.py:51-67, before gather)."""
    kvec = fftfreq(mesh_shape, cell_size, dtype=dtype)
    coef = dtype(1.5) * dtype(omega_m)
    d = (np.asarray(dens, dtype=dtype) - dtype(1)) * coef
    dk = sfft.rfftn(d)
    pot = laplace(kvec, dk)
    out = []
    for k in kvec:
        g = sfft.irfftn(neg_grad(k, pot, cell_size), s=mesh_shape).astype(dtype)
        out.append(g)
    return out


def gravity(pmid, disp, omega_m, cell_size, mesh_shape, ptcl_num=None,
            dtype=np.float32, return_dens=False, _cics=None):
    """Accelerations (N, 3) (gravity.py:47-72).

    Equals ``scatter`` followed by three ``gather`` calls (default
    arguments) with the CIC geometry computed once (``cic_default``)."""
    pmid = np.asarray(pmid)
    disp = np.asarray(disp, dtype=dtype)
    n = len(pmid)
    if ptcl_num is None:
        ptcl_num = n
    M = int(np.prod(mesh_shape))
    val = np.asarray(M / ptcl_num, dtype=dtype)
    cics = _cics if _cics is not None else _cic_chunks(pmid, disp, cell_size, mesh_shape)
    dens = _scatter_cic(cics, M, lambda s, e, w: val * w, dtype).reshape(mesh_shape)
    g = gravity_mesh(dens, omega_m, cell_size, mesh_shape, dtype)
    acc = _gather_cic(cics, np.stack([gi.ravel() for gi in g], axis=-1), n, dtype, C=3)
    if return_dens:
        return acc, dens, g
    return acc


def gravity_vjp(pmid, disp, omega_m, cell_size, mesh_shape, acc_cot, ptcl_num=None,
                dtype=np.float32):
    """VJP of gravity w.r.t. disp and Omega_m (what jax.vjp of gravity.py:47-72 yields).

    Chain: gather adjoints of the three force meshes, the transpose of
    irfftn(-ik * laplace(rfftn(.))) per axis, the scale 1.5*Omega_m, then the
    scatter adjoint.  Returns (acc, disp_cot, omega_m_cot).  The CIC
    geometry is computed once and shared (``cic_default``); the arithmetic is
    that of ``gather_adj`` and ``scatter_adj``.
    """
    pmid = np.asarray(pmid)
    disp = np.asarray(disp, dtype=dtype)
    n = len(pmid)
    if ptcl_num is None:
        ptcl_num = n
    M = int(np.prod(mesh_shape))
    cics = _cic_chunks(pmid, disp, cell_size, mesh_shape, grad=True)
    acc, dens, g = gravity(pmid, disp, omega_m, cell_size, mesh_shape, ptcl_num, dtype,
                           return_dens=True, _cics=cics)
    acc_cot = np.asarray(acc_cot, dtype=dtype)
    kvec = fftfreq(mesh_shape, cell_size, dtype=dtype)
    inv_cell = dtype(cell_size)
    disp_cot = np.zeros_like(disp, dtype=dtype)
    gflat = [gi.ravel() for gi in g]
    dsum = None
    for d in range(3):
        # gather_adj of mesh g[d] with val_cot = acc_cot[:, d]
        dc = np.empty_like(disp)

        def one(i, _e, d=d):
            s, e, flat, w, fg = cics[i]
            vcot = acc_cot[s:e, d][:, None]                       # (k, 1)
            mv = gflat[d][flat].astype(dtype)                     # (k, 8)
            dck = vcot * mv
            dc[s:e] = (dck[..., None] * fg).sum(axis=1) / inv_cell
        _pool_map(one, len(cics), chunk=1)
        disp_cot = disp_cot + dc
        mc = _scatter_cic(cics, M, lambda s, e, w, d=d: acc_cot[s:e, d][:, None] * w,
                          dtype).reshape(mesh_shape)
        # transpose of irfftn(s=mesh): the adjoint of the c2r inverse is a
        # forward r2c scaled by 1/M with the non-self-conjugate modes doubled;
        # composed with the odd, anti-Hermitian -ik/k^2 it equals -L_d applied
        # to the cotangent mesh (L_d = irfftn . (-ik) . laplace . rfftn).
        mk = sfft.rfftn(mc)
        t = neg_grad(kvec[d], laplace(kvec, mk), cell_size)
        dsum = t if dsum is None else dsum + t
    coef = dtype(1.5) * dtype(omega_m)
    dens_cot = ((-sfft.irfftn(dsum, s=mesh_shape)).astype(dtype) * coef).ravel()
    val = np.asarray(M / ptcl_num, dtype=dtype)
    dc2 = np.empty_like(disp)

    def one2(i, _e):
        s, e, flat, w, fg = cics[i]
        vc = dens_cot[flat].astype(dtype)                         # (k, 8)
        dck = vc * val
        dc2[s:e] = (dck[..., None] * fg).sum(axis=1) / inv_cell
    _pool_map(one2, len(cics), chunk=1)
    disp_cot = disp_cot + dc2
    om_cot = float(np.sum(acc_cot.astype(np.float64) * acc.astype(np.float64))) / omega_m
    return acc, disp_cot.astype(dtype), om_cot


def gen_grid(ptcl_grid_shape, mesh_shape, cell_size, dtype=np.float32, pmid_dtype=np.int16):
    """Particles.gen_grid (particles.py:109-144): Lagrangian pmid and disp."""
    pm, dp = [], []
    for sp, sm in zip(ptcl_grid_shape, mesh_shape):
        p1 = np.rint(np.linspace(0, sm, num=sp, endpoint=False)).astype(pmid_dtype)
        pm.append(p1)
        d1 = np.arange(sp) * sm - p1.astype(np.int64) * sp
        d1 = (d1 * (cell_size / sp)).astype(dtype)
        dp.append(d1)
    pmid = np.stack(np.meshgrid(*pm, indexing='ij'), axis=-1).reshape(-1, len(mesh_shape))
    disp = np.stack(np.meshgrid(*dp, indexing='ij'), axis=-1).reshape(-1, len(mesh_shape))
    return pmid, disp


def raveled_id(pmid, mesh_shape, wrap=False):
    """Particles.raveled_id (particles.py:156-182): sum pmid_d * stride_d with
    C-order strides, in uint64 arithmetic (so negative ids wrap mod 2^64)."""
    pmid = np.asarray(pmid)
    if wrap:
        pmid = np.mod(pmid, np.asarray(mesh_shape, dtype=pmid.dtype))
    strides = np.cumprod((1,) + tuple(mesh_shape)[:0:-1])[::-1]
    out = np.zeros(len(pmid), dtype=np.uint64)
    for d, s in enumerate(strides):
        out = out + pmid[:, d].astype(np.int64).astype(np.uint64) * np.uint64(s)
    return out


def pos(pmid, disp, cell_size, box_size, dtype=np.float64, wrap=True):
    """Particles.pos (particles.py:184-209): pmid*cell + disp in ``dtype``,
    wrapped with a floor mod into the box."""
    p = np.asarray(pmid).astype(dtype) * dtype(cell_size) + np.asarray(disp).astype(dtype)
    if wrap:
        p = np.mod(p, np.asarray(box_size, dtype=dtype))
    return p


def from_pos(pos_, cell_size, mesh_shape, wrap=True, pmid_dtype=np.int16, float_dtype=np.float32):
    """Particles.from_pos (particles.py:81-107): pmid = rint(pos/cell),
    disp = pos - pmid*cell in the positions' dtype, then the casts and the
    int16 floor-mod wrap."""
    pos_ = np.asarray(pos_)
    pmid = np.rint(pos_ / cell_size)
    disp = pos_ - pmid * cell_size
    pmid = pmid.astype(pmid_dtype)
    disp = disp.astype(float_dtype)
    if wrap:
        pmid = np.mod(pmid, np.asarray(mesh_shape, dtype=pmid_dtype)).astype(pmid_dtype)
    return pmid, disp


# ---------- pieces of one force for the timed CPU baseline (bench.py) ----------

def force_particle_setup(pmid, disp, cell_size, mesh_shape, ptcl_num, dtype=np.float32):
    """The particle half of a force before the Poisson solve: the shared CIC
    geometry and the particles' contribution to the density (gravity's
    scatter, values mesh_size / ptcl_num).  Returns (cics, density part)."""
    pmid = np.asarray(pmid)
    disp = np.asarray(disp, dtype=dtype)
    M = int(np.prod(mesh_shape))
    val = np.asarray(M / ptcl_num, dtype=dtype)
    cics = _cic_chunks(pmid, disp, cell_size, mesh_shape)
    dens = _scatter_cic(cics, M, lambda s, e, w: val * w, dtype).reshape(mesh_shape)
    return cics, dens


def force_particle_readout(cics, g, n, dtype=np.float32):
    """The particle half of a force after the Poisson solve: the three
    readouts of gravity (gather.py:33-77) from the force meshes g."""
    return _gather_cic(cics, np.stack([gi.ravel() for gi in g], axis=-1), n, dtype, C=3)
