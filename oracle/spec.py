"""Power spectrum restated from pmwd/spec_util.py (oracle side, test
infrastructure).

``getbins`` follows _getbins (spec_util.py:9-44) and ``powspec`` follows
powspec (spec_util.py:47-147): the unnormalised rfftn of the field(s), |F|^2
(or F G* for a cross spectrum) in the field's precision, the wavenumber
|k| of fftfreq(shape, None) in the same precision, the digitised bins, and
float64 bincounts of k N, P N and N with N = 2 except on the kz = 0 (and,
for an even last axis, Nyquist) planes.
"""
import math

import numpy as np
import scipy.fft as sfft

from . import pm


def getbins(grid_shape, bins, cut_nyq):
    """spec_util.py:9-44: (bnum, bcut, edges in cycles per cell, right)."""
    kfun = 1 / max(grid_shape)
    knyq = 0.5
    kmax = knyq * math.sqrt(3)
    if isinstance(bins, (int, float)):
        width = bins * kfun
        bnum = 1 + math.ceil(kmax / width)
        bcut = 1 + math.ceil(knyq / width) if cut_nyq else bnum
        edges = width * np.arange(bnum, dtype=np.float64)
        right = True
    elif isinstance(bins, complex):
        kmaxable = all(s % 2 == 0 for s in grid_shape)
        bnum = 1 + math.ceil(math.log2(kmax / kfun) / bins.imag) + kmaxable
        bcut = 1 + math.ceil(math.log2(knyq / kfun) / bins.imag) if cut_nyq else bnum
        edges = kfun * 2 ** (bins.imag * np.arange(bnum, dtype=np.float64))
        right = False
    elif isinstance(bins, tuple):
        if bins[0] != 0:
            raise ValueError(f'{bins=} must starts from 0')
        bnum = len(bins)
        if cut_nyq:
            for bcut, edge in enumerate(bins, start=1):
                if edge >= knyq:
                    break
        else:
            bcut = bnum
        edges = np.asarray(bins, dtype=np.float64)
        right = True
    else:
        raise ValueError(f'{bins=} not supported')
    return bnum, bcut, edges, right


def powspec(f, spacing, bins=1j / 3, g=None, deconv=None, cut_zero=True, cut_nyq=True):
    """spec_util.py:47-147 in numpy; returns (k, P, N, bins)."""
    f = np.asarray(f)
    if g is not None and f.shape != np.shape(g):
        raise ValueError(f'shape mismatch: {f.shape} != {np.shape(g)}')
    grid_shape = f.shape[-3:]
    bnum, bcut, edges, right = getbins(grid_shape, bins, cut_nyq)
    F = sfft.rfftn(f, axes=(-3, -2, -1))
    if g is None:
        P = F.real**2 + F.imag**2
    else:
        G = sfft.rfftn(np.asarray(g), axes=(-3, -2, -1))
        P = F * G.conj()
    if P.ndim > 3:
        P = P.sum(tuple(range(P.ndim - 3)))
    rdt = P.real.dtype.type
    kvec = pm.fftfreq(grid_shape, None, dtype=rdt)
    k = np.sqrt(sum(kk**2 for kk in kvec))
    if deconv is not None:
        for kk in kvec:
            P = P * np.sinc(kk)**-deconv
    N = np.full(P.shape, 2, dtype=np.uint32)
    N[..., 0] = 1
    if grid_shape[-1] % 2 == 0:
        N[..., -1] = 1
    k, P, N = k.ravel(), P.ravel(), N.ravel()
    b = np.digitize(k.astype(np.float64), edges, right=right)
    keep = b < bnum
    kw = (k * N).astype(np.float64)
    Nw = N.astype(np.float64)

    def count(w):
        return np.bincount(b[keep], weights=w[keep], minlength=bnum)[:bnum]
    if np.iscomplexobj(P):
        Pw = (P * N).astype(np.complex128)
        Ps = count(Pw.real) + 1j * count(Pw.imag)
    else:
        Ps = count((P * N).astype(np.float64))
    ks = count(kw)
    Ns = count(Nw)
    ks, Ps, Ns = ks[cut_zero:bcut], Ps[cut_zero:bcut], Ns[cut_zero:bcut]
    edges = edges[:bcut]
    ks = ks / Ns
    Ps = Ps / Ns
    ks = ks * (2 * math.pi / spacing)
    edges = edges * (2 * math.pi / spacing)
    Ps = Ps * (spacing**3 / math.prod(grid_shape))
    return ks, Ps, Ns, edges
