"""Initial conditions restated from pmwd (oracle side).

``white_noise`` (modes.py:12-49) draws from ``numpy.random.default_rng(seed)``
instead of ``jax.random`` (threefry), whose bit stream is not reproducible
without jax: the same generator is used by ``pmwd_amd.modes.white_noise``, so
both sides see identical modes (the mapping seed -> modes is parity unpinned
against pmwd).  ``linear_modes`` follows modes.py:67-117 and ``lpt`` follows
lpt.py:13-76 and 136-212 (2LPT), in ``dtype`` arithmetic with FP64 power
spectrum and growth factors.
"""
import numpy as np
import scipy.fft as sfft

from . import cosmo as oc
from . import pm


def white_noise(seed, conf, real=False, unit_abs=False):
    """modes.py:12-49 with a numpy generator (see module docstring)."""
    modes = np.random.default_rng(seed).standard_normal(conf.ptcl_grid_shape, dtype=np.float32)
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
