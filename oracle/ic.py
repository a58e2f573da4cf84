"""Initial conditions restated from pmwd (oracle side).

``white_noise`` (modes.py:12-49) draws from Philox4x32-10 (Salmon et al.
2011) with a float64 Box-Muller transform instead of ``jax.random``
(threefry), whose bit stream is not reproducible without jax; the mapping
seed -> modes is therefore parity unpinned against pmwd.  The generator is
pinned by the published Philox known-answer vectors and, through its 4x64
variant with the same round function, by numpy's independent Philox.  ``linear_modes`` follows modes.py:67-117 and ``lpt`` follows
lpt.py:13-76 and 136-212 (2LPT), in ``dtype`` arithmetic with FP64 power
spectrum and growth factors.
"""
import numpy as np
import scipy.fft as sfft

from . import cosmo as oc
from . import pm


# Philox constants (Salmon et al. 2011): round multipliers and Weyl key
# increments of the 4x32 and 4x64 variants
PHILOX = {32: ((0xD2511F53, 0xCD9E8D57), (0x9E3779B9, 0xBB67AE85)),
          64: ((0xD2E7470EE14C6C93, 0xCA5A826395121157), (0x9E3779B97F4A7C15, 0xBB67AE8584CAA73B))}


def philox(ctr, key, bits=32, rounds=10):
    """Philox4xW-R of one counter (4 ints) and key (2 ints), Python ints:
    round r (key bumped by the Weyl increments before every round but the
    first) maps (c0, c1, c2, c3) to (hi(M1 c2)^c1^k0, lo(M1 c2), hi(M0 c0)^c3^k1,
    lo(M0 c0))."""
    (m0, m1), (w0, w1) = PHILOX[bits]
    mask = (1 << bits) - 1
    c, k = list(ctr), list(key)
    for r in range(rounds):
        if r:
            k = [(k[0] + w0) & mask, (k[1] + w1) & mask]
        p0, p1 = m0 * c[0], m1 * c[2]
        c = [(p1 >> bits) ^ c[1] ^ k[0], p1 & mask, (p0 >> bits) ^ c[3] ^ k[1], p0 & mask]
    return c


def philox4x32_blocks(seed, blocks):
    """Philox4x32-10 words (len(blocks), 4) uint32 of the counters
    (b low, b high, 0, 0) under the key (seed low, seed high), vectorised in
    uint64 arithmetic (the same rounds as ``philox``)."""
    (m0, m1), (w0, w1) = PHILOX[32]
    b = np.asarray(blocks, dtype=np.uint64)
    mask = np.uint64(0xFFFFFFFF)
    c = [b & mask, b >> np.uint64(32), np.zeros_like(b), np.zeros_like(b)]
    seed = int(seed) % 2**64
    k0, k1 = seed & 0xFFFFFFFF, seed >> 32
    for r in range(10):
        if r:
            k0, k1 = (k0 + w0) & 0xFFFFFFFF, (k1 + w1) & 0xFFFFFFFF
        p0 = np.uint64(m0) * c[0]
        p1 = np.uint64(m1) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ np.uint64(k0), p1 & mask,
             (p0 >> np.uint64(32)) ^ c[3] ^ np.uint64(k1), p0 & mask]
    return np.stack(c, axis=-1).astype(np.uint32)


def normals(seed, index):
    """Standard normals number ``index`` (any int array) of the stream of
    ``seed``: element i is normal i % 4 of Philox block i // 4, from the
    float64 Box-Muller transform of the block's word pairs."""
    index = np.asarray(index, dtype=np.int64)
    w = philox4x32_blocks(seed, index // 4).astype(np.float64)
    u = (w + 0.5) * 2.0**-32
    j = index % 4
    ua = np.where(j < 2, u[..., 0], u[..., 2])
    ub = np.where(j < 2, u[..., 1], u[..., 3])
    r = np.sqrt(-2.0 * np.log(ua))
    t = 6.283185307179586 * ub
    z = np.where(j % 2 == 0, r * np.cos(t), r * np.sin(t))
    return z.astype(np.float32)


def white_noise(seed, conf, real=False, unit_abs=False):
    """modes.py:12-49 with the Philox generator (see module docstring)."""
    n = int(np.prod(conf.ptcl_grid_shape))
    modes = normals(seed, np.arange(n)).reshape(conf.ptcl_grid_shape)
    if real and not unit_abs:
        return modes
    modes = sfft.rfftn(modes, norm='ortho')
    if unit_abs:
        modes = modes / np.abs(modes)
    if real:
        modes = sfft.irfftn(modes, s=conf.ptcl_grid_shape, norm='ortho')
    return modes


def linear_modes(modes, cosmo, conf, a=None, real=False, dtype=np.float32):
    """sqrt(V Plin(k)) * modes (modes.py:67-117).

    |k| is formed in ``dtype`` from the ``dtype`` wavevectors, Plin in FP64,
    and the product is rounded back to ``dtype``.
    """
    kvec = pm.fftfreq(conf.ptcl_grid_shape, conf.ptcl_spacing, dtype=dtype)
    k2 = 0
    for k in kvec:
        k2 = k2 + k**2
    k = np.sqrt(k2)
    P = oc.linear_power(k.astype(np.float64), a, cosmo, conf)
    if not np.iscomplexobj(modes):
        modes = sfft.rfftn(np.asarray(modes, dtype=dtype), norm='ortho')
    fac = np.sqrt(P * conf.box_vol).astype(dtype)
    modes = (modes * fac).astype(np.result_type(dtype, np.complex64))
    if real:
        modes = pm.fftinv(modes, shape=conf.ptcl_grid_shape, norm=conf.ptcl_spacing)
    return modes


def _strain(kvec, i, j, pot, conf, dtype):
    """lpt.py:13-37."""
    k_i, k_j = kvec[i], kvec[j]
    if i != j:
        k_i = np.where(pm.nyquist_mask(k_i, conf.ptcl_spacing), dtype(0), k_i)
        k_j = np.where(pm.nyquist_mask(k_j, conf.ptcl_spacing), dtype(0), k_j)
    s = (-k_i * k_j) * pot
    return sfft.irfftn(s, s=conf.ptcl_grid_shape).astype(dtype)


def _L(kvec, pot, conf, dtype):
    """lpt.py:40-76 with pot_n = pot_m (second-order source)."""
    L = np.zeros(conf.ptcl_grid_shape, dtype=dtype)
    dim = conf.dim
    for i in range(dim):
        sm = _strain(kvec, i, i, pot, conf, dtype)
        for j in range(dim - 1, i, -1):
            L = L + sm * _strain(kvec, j, j, pot, conf, dtype)
    for i in range(dim - 1):
        for j in range(i + 1, dim):
            s = _strain(kvec, i, j, pot, conf, dtype)
            L = L - s * s
    return L


def lpt(modes, cosmo, conf, dtype=np.float32):
    """2LPT displacements and velocities on the particle grid (lpt.py:136-212).

    Returns (pmid, disp, vel) in particle order.
    """
    modes = np.asarray(modes) / dtype(conf.ptcl_cell_vol)
    kvec = pm.fftfreq(conf.ptcl_grid_shape, conf.ptcl_spacing, dtype=dtype)
    pots = []
    if conf.lpt_order > 0:
        pots.append(pm.laplace(kvec, modes))
    if conf.lpt_order > 1:
        src2 = sfft.rfftn(_L(kvec, pots[0], conf, dtype))
        pots.append(pm.laplace(kvec, src2))
    if conf.lpt_order > 2:
        raise NotImplementedError('lpt_order=3 is not implemented in pmwd either')
    pmid, disp = pm.gen_grid(conf.ptcl_grid_shape, conf.mesh_shape, conf.cell_size, dtype)
    vel = np.zeros_like(disp)
    a = conf.a_start
    for order in range(1, 1 + conf.lpt_order):
        D = oc.growth(a, cosmo, conf, order=order)
        dD = oc.growth(a, cosmo, conf, order=order, deriv=1)
        a2HDp = dtype(a**2 * np.sqrt(oc.E2(a, cosmo)) * dD)
        D = dtype(D)
        for i, k in enumerate(kvec):
            g = pm.neg_grad(k, pots[order - 1], conf.ptcl_spacing)
            g = sfft.irfftn(g, s=conf.ptcl_grid_shape).astype(dtype).ravel()
            disp[:, i] = disp[:, i] + D * g
            vel[:, i] = vel[:, i] + a2HDp * g
    return pmid, disp, vel
