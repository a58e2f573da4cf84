"""KDK N-body integrator and its reverse-time adjoint (oracle side).

Restates pmwd/nbody.py: growth factors ``_G_D``/``_G_K`` (12-22),
``drift_factor``/``kick_factor`` (25-36), ``drift``/``kick`` (39-77),
``force`` (102-105), ``integrate`` (121-140), ``nbody`` (215-223) and the
adjoint ``drift_adj``/``kick_adj``/``force_adj``/``integrate_adj``/
``nbody_adj`` (49-118, 143-162, 226-260).

State arrays are float32 (``dtype``) with float64 time-step factors cast to
``dtype`` at use, as in the reference.  The cosmology cotangent is reduced to
Omega_m: the factors' Omega_m derivatives are taken by a central difference
of the FP64 growth solution (step 1e-5), and the force term uses the closed
form sum(vel_cot*acc)/Omega_m.
"""
import numpy as np

from . import cosmo as oc
from . import pm


def G_D(a, cosmo, conf):
    """nbody.py:12-14."""
    return a**2 * np.sqrt(oc.E2(a, cosmo)) * oc.growth(a, cosmo, conf, deriv=1)


def G_K(a, cosmo, conf):
    """nbody.py:17-22."""
    return a**3 * oc.E2(a, cosmo) * (oc.growth(a, cosmo, conf, deriv=2)
                                      + (2 + oc.H_deriv(a, cosmo)) * oc.growth(a, cosmo, conf, deriv=1))


def drift_factor(a_vel, a_prev, a_next, cosmo, conf):
    """nbody.py:25-29."""
    return (oc.growth(a_next, cosmo, conf) - oc.growth(a_prev, cosmo, conf)) / G_D(a_vel, cosmo, conf)


def kick_factor(a_acc, a_prev, a_next, cosmo, conf):
    """nbody.py:32-36."""
    return (G_D(a_next, cosmo, conf) - G_D(a_prev, cosmo, conf)) / G_K(a_acc, cosmo, conf)


def step_schedule(conf, reverse=False):
    """Per-step (kind, factor-args) list of one KDK step (integrate, nbody.py:121-140)."""
    a_nbody = conf.a_nbody[::-1] if reverse else conf.a_nbody
    steps = []
    for a_prev, a_next in zip(a_nbody[:-1], a_nbody[1:]):
        ops = []
        D = K = 0
        a_disp = a_vel = a_acc = a_prev
        for d, k in conf.symp_splits:
            if d != 0:
                D += d
                a_disp_next = a_prev * (1 - D) + a_next * D
                ops.append(('drift', a_vel, a_disp, a_disp_next))
                a_disp = a_disp_next
                ops.append(('force', a_disp))
                a_acc = a_disp
            if k != 0:
                K += k
                a_vel_next = a_prev * (1 - K) + a_next * K
                ops.append(('kick', a_acc, a_vel, a_vel_next))
                a_vel = a_vel_next
        steps.append(ops)
    return a_nbody, steps


def nbody(pmid, disp, vel, cosmo, conf, reverse=False, dtype=np.float32, n_steps=None):
    """Forward integration; returns (disp, vel, acc) (nbody.py:193-223)."""
    disp = np.asarray(disp, dtype=dtype).copy()
    vel = np.asarray(vel, dtype=dtype).copy()
    cell, ms, N = conf.cell_size, conf.mesh_shape, conf.ptcl_num
    a_nbody, steps = step_schedule(conf, reverse)
    acc = pm.gravity(pmid, disp, cosmo.Omega_m, cell, ms, N, dtype)
    for ops in steps[:n_steps]:
        for op in ops:
            if op[0] == 'drift':
                f = dtype(drift_factor(op[1], op[2], op[3], cosmo, conf))
                disp = disp + vel * f
            elif op[0] == 'kick':
                f = dtype(kick_factor(op[1], op[2], op[3], cosmo, conf))
                vel = vel + acc * f
            else:
                acc = pm.gravity(pmid, disp, cosmo.Omega_m, cell, ms, N, dtype)
    return disp, vel, acc


def _dfac_dOm(fn, args, cosmo, conf, h=1e-5):
    # tight ODE tolerance so the adaptive step choice does not add noise to
    # the central difference
    tight = conf.replace(growth_rtol=1e-12, growth_atol=1e-12)
    out = []
    for s in (1, -1):
        c = oc.growth_integ(cosmo.replace(Omega_m=cosmo.Omega_m + s * h), tight)
        out.append(fn(*args, c, tight))
    return (out[0] - out[1]) / (2 * h)


def nbody_adj(pmid, disp, vel, acc, disp_cot, vel_cot, cosmo, conf, reverse=False,
              dtype=np.float32):
    """Reverse-time adjoint (nbody.py:226-260, 49-118, 143-162).

    Takes the final state and its cotangent; returns the reconstructed initial
    state, the initial-state cotangents (disp_cot, vel_cot, acc_cot) and the
    Omega_m cotangent.
    """
    cell, ms, N, Om = conf.cell_size, conf.mesh_shape, conf.ptcl_num, cosmo.Omega_m
    disp = np.asarray(disp, dtype=dtype).copy()
    vel = np.asarray(vel, dtype=dtype).copy()
    disp_cot = np.asarray(disp_cot, dtype=dtype).copy()
    vel_cot = np.asarray(vel_cot, dtype=dtype).copy()
    a_nbody = conf.a_nbody[::-1] if reverse else conf.a_nbody

    def force_adj():
        a, dc, oc_ = pm.gravity_vjp(pmid, disp, Om, cell, ms, vel_cot, N, dtype)
        return a, dc, oc_

    acc, acc_cot, cf = force_adj()
    om_cot = 0.0
    for a_prev, a_next in zip(a_nbody[:0:-1], a_nbody[-2::-1]):
        D = K = 0
        a_disp = a_vel = a_acc = a_prev
        for d, k in reversed(conf.symp_splits):
            if k != 0:
                K += k
                a_vel_next = a_prev * (1 - K) + a_next * K
                args = (a_acc, a_vel, a_vel_next)
                f = kick_factor(*args, cosmo, conf)
                dfdO = _dfac_dOm(kick_factor, args, cosmo, conf)
                f32 = dtype(f)
                vel = vel + acc * f32
                disp_cot = disp_cot - acc_cot * f32
                om_cot -= dfdO * float(np.sum(vel_cot.astype(np.float64) * acc)) + cf * float(f32)
                a_vel = a_vel_next
            if d != 0:
                D += d
                a_disp_next = a_prev * (1 - D) + a_next * D
                args = (a_vel, a_disp, a_disp_next)
                f = drift_factor(*args, cosmo, conf)
                dfdO = _dfac_dOm(drift_factor, args, cosmo, conf)
                f32 = dtype(f)
                disp = disp + vel * f32
                vel_cot = vel_cot - disp_cot * f32
                om_cot -= dfdO * float(np.sum(disp_cot.astype(np.float64) * vel))
                a_disp = a_disp_next
                acc, acc_cot, cf = force_adj()
                a_acc = a_disp
    return (disp, vel, acc), (disp_cot, vel_cot, acc_cot), om_cot
