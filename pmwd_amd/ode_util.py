"""Adaptive Dormand-Prince (RK5(4)) integration on the host in FP64.

Same step control as pmwd's vendored integrator (pmwd/ode_util.py:60-186:
Dopri5 tableau, RMS error ratio over ``atol + rtol*max(|y0|,|y1|)``, step
factor ``clip(0.9*ratio^-1/5, 0.2 or 1, 10)``, 4th-order dense output at the
requested times) with one addition: the error norm can be restricted to a
leading subset of the state, so that forward-sensitivity components ride
along on exactly the steps the plain system would take.
"""
import numpy as np

# Dormand & Prince (1980) coefficients
_C = np.array([0.0, 1 / 5, 3 / 10, 4 / 5, 8 / 9, 1.0, 1.0])
_A = [
    [],
    [1 / 5],
    [3 / 40, 9 / 40],
    [44 / 45, -56 / 15, 32 / 9],
    [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
    [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
    [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84],
]
_B5 = np.array([35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0])
# error weights of pmwd's integrator (ode_util.py:111-114), so that the
# accepted steps, and therefore the tables, are the same
_BERR = np.array([35 / 384 - 1951 / 21600, 0, 500 / 1113 - 22642 / 50085,
                  125 / 192 - 451 / 720, -2187 / 6784 - -12231 / 42400,
                  11 / 84 - 649 / 6300, -1. / 60.])
# dense output at the step midpoint (Shampine 1986)
_BMID = np.array([6025192743 / 30085553152, 0, 51252292925 / 65400821598,
                  -2691868925 / 45128329728, 187940372067 / 1594534317056,
                  -1776094331 / 19743644256, 11237099 / 235043384]) / 2


def _step(f, t, y, f0, h):
    K = np.zeros((7, len(y)))
    K[0] = f0
    for i in range(1, 7):
        yi = y + h * (np.asarray(_A[i]) @ K[:i])
        K[i] = f(yi, t + _C[i] * h)
    y1 = h * (_B5 @ K) + y
    return y1, K, h * (_BERR @ K)


def _quartic(y0, y1, K, h):
    """Coefficients (highest first) of the quartic through the step."""
    ym = y0 + h * (_BMID @ K)
    d0, d1 = h * K[0], h * K[-1]
    return np.stack([
        -2 * d0 + 2 * d1 - 8 * y0 - 8 * y1 + 16 * ym,
        5 * d0 - 3 * d1 + 18 * y0 + 14 * y1 - 32 * ym,
        -4 * d0 + d1 - 11 * y0 - 5 * y1 + 16 * ym,
        d0,
        y0,
    ])


def odeint(f, y0, ts, rtol, atol, dt0=None, nerr=None, max_steps=100000):
    """Solution of y' = f(y, t) at the increasing times ``ts``.

    ``nerr`` restricts the error control to ``y[:nerr]``.
    """
    y = np.asarray(y0, dtype=np.float64).copy()
    ne = len(y) if nerr is None else nerr
    fy = f(y, ts[0])
    h = dt0[0] if isinstance(dt0, tuple) else dt0
    if h is None:
        h = _initial_step(f, ts[0], y[:ne], fy[:ne], rtol, atol, y, fy)
    t = ts[0]
    t_last = ts[0]
    poly = np.stack([y] * 5)
    out = [y.copy()]
    nstep = 0
    for target in ts[1:]:
        while t < target and h > 0:
            nstep += 1
            if nstep > max_steps:
                raise RuntimeError('odeint: too many steps')
            y1, K, err = _step(f, t, y, fy, h)
            scale = atol + rtol * np.maximum(np.abs(y[:ne]), np.abs(y1[:ne]))
            ratio = np.sqrt(np.mean((err[:ne] / scale)**2))
            if ratio == 0:
                h_new = 10.0 * h
            else:
                lo = 1.0 if ratio < 1 else 0.2
                h_new = h * min(10.0, max(0.9 * ratio**-0.2, lo))
            if ratio <= 1:
                poly = _quartic(y, y1, K, h)
                t_last, t = t, t + h
                y, fy = y1, K[-1]
            h = h_new
        x = (target - t_last) / (t - t_last)
        val = poly[0]
        for c in poly[1:]:
            val = val * x + c
        out.append(val)
    return np.stack(out)


def _initial_step(f, t0, y0, f0, rtol, atol, yfull, ffull):
    """Hairer-Norsett-Wanner initial step (Solving ODEs I, II.4)."""
    scale = atol + np.abs(y0) * rtol
    d0 = np.linalg.norm(y0 / scale)
    d1 = np.linalg.norm(f0 / scale)
    h0 = 1e-6 if (d0 < 1e-5 or d1 < 1e-5) else 0.01 * d0 / d1
    f1 = f(yfull + h0 * ffull, t0 + h0)[:len(y0)]
    d2 = np.linalg.norm((f1 - f0) / scale) / h0
    if d1 <= 1e-15 and d2 <= 1e-15:
        h1 = max(1e-6, h0 * 1e-3)
    else:
        h1 = (0.01 / max(d1, d2))**(1 / 5)
    return min(100 * h0, h1)
