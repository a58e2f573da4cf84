"""Power spectrum (pmwd/spec_util.py:9-147) on the GPU.

The bins are set up on the host as ``_getbins`` does; the transform and the
binning of every mode run in the library (``pmx_powspec``: the HIP r2c, the
float32 |k| and P of each mode, float64 sums per bin).  Fields with leading
dimensions are transformed one 3D slice at a time and their binned power
added, which equals binning the summed power up to float32 rounding.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib


def _getbins(grid_shape, bins, cut_nyq):
    """spec_util.py:9-44: (bnum, bcut, edges, right), edges in cycles per cell."""
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
        kmaxable = all(s % 2 == 0 for s in grid_shape)   # extra bin just in case
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


def powspec(f, spacing, bins=1j / 3, g=None, deconv=None, cut_zero=True, cut_nyq=True,
            dtype=torch.float64, int_dtype=torch.int64):
    """Auto or cross power spectrum in 3D averaged in spherical bins
    (spec_util.py:47-147).

    ``f`` (and ``g``) are float32 device tensors whose last three axes are
    transformed; the others are reduced by the sum of the power.  Returns
    ``(k, P, N, bins)`` as device tensors: wavenumbers and bin edges in
    angular units of 1/spacing, the power (complex for a cross spectrum) in
    spacing^3, and the number of modes per bin.
    """
    f = torch.as_tensor(f)
    _lib.require_device(f, 'f')
    if g is not None:
        g = torch.as_tensor(g)
        if tuple(f.shape) != tuple(g.shape):
            raise ValueError(f'shape mismatch: {tuple(f.shape)} != {tuple(g.shape)}')
    grid_shape = tuple(f.shape[-3:])
    bnum, bcut, edges, right = _getbins(grid_shape, bins, cut_nyq)
    if deconv is not None and (int(deconv) != deconv or deconv < 0):
        raise NotImplementedError('deconv must be a non-negative integer')
    c = _lib.context_any_cell(grid_shape, f.device)
    h = c.bind()
    ff = f.to(torch.float32).reshape((-1,) + grid_shape)
    gg = g.to(torch.float32).reshape((-1,) + grid_shape) if g is not None else None
    e = np.ascontiguousarray(edges, dtype=np.float64)
    tot = np.zeros(4 * bnum)
    out = np.zeros(4 * bnum)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    for i in range(ff.shape[0]):
        fi = ff[i].contiguous()
        gi = gg[i].contiguous() if gg is not None else None
        _lib.check(_lib.lib().pmx_powspec(h, _lib.ptr(fi), _lib.ptr(gi), bnum, P(e), int(right),
                                          int(deconv or 0), P(out)))
        # leading axes sum the power (spec_util.py:114-115); the k sums and
        # mode counts are those of one field
        if i == 0:
            tot += out
        else:
            tot[bnum:3 * bnum] += out[bnum:3 * bnum]
    kN, Pr, Pi, N = tot.reshape(4, bnum)
    sl = slice(int(cut_zero), bcut)
    kN, Pr, Pi, N = kN[sl], Pr[sl], Pi[sl], N[sl]
    k = kN / N * (2 * math.pi / spacing)
    scale = spacing**3 / math.prod(grid_shape)
    Pw = (Pr + 1j * Pi) if g is not None else Pr
    Pw = Pw / N * scale
    dev = f.device
    cdt = torch.complex128 if dtype == torch.float64 else torch.complex64
    return (torch.as_tensor(k, dtype=dtype, device=dev),
            torch.as_tensor(Pw, dtype=cdt if g is not None else dtype, device=dev),
            torch.as_tensor(N, device=dev).to(int_dtype),
            torch.as_tensor(edges[:bcut] * (2 * math.pi / spacing), dtype=dtype, device=dev))
