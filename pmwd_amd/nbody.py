"""N-body KDK integration and its reverse-time adjoint (pmwd/nbody.py).

Time-step factors are computed on the host in FP64 from the growth table
(nbody.py:12-36) and rounded to float32 at use, as pmwd does; the particle
state never leaves the GPU.  With the default ``symp_splits`` (KDK) the
steps between checkpoints are one ``pmx_nbody_steps`` call (particles kept
in tile order on the device; per step a scatter, the Poisson solve and one
fused gather + closing kick + next kick/drift + re-bin kernel); other
splittings run the ops one by one.

The adjoint reconstructs the trajectory backward in time from the final
state as pmwd's ``nbody_adj`` (nbody.py:226-260) does, optionally snapping
the reconstruction to sparse forward checkpoints (every ``checkpoint_every``
steps, or at given steps) to bound the float32 reversal drift;
``reversal_checkpoints`` finds the steps where that drift exceeds a
tolerance, so that checkpoints go only there.  The cosmology cotangent holds
Omega_m (time-step factor sensitivities from the growth ODE's forward
sensitivity, and the force term sum(vel_cot*acc)/Omega_m) and, when set,
Omega_k_, w_0_ and w_a_, which enter only the time-step factors.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .boltzmann import boltzmann, growth, growth_dOm
from .cosmology import E2, H_deriv, E2_dOm, H_deriv_dOm
from .gravity import omega_coef


# ------------------------------------------------------ FP64 factors ---

def _G_D(a, cosmo, conf):
    """Growth factor of ZA canonical velocity in [H_0] (nbody.py:12-14)."""
    return a**2 * np.sqrt(E2(a, cosmo)) * growth(a, cosmo, conf, deriv=1)


def _G_K(a, cosmo, conf):
    """Growth factor of ZA accelerations in [H_0^2] (nbody.py:17-22)."""
    return a**3 * E2(a, cosmo) * (growth(a, cosmo, conf, deriv=2)
                                  + (2 + H_deriv(a, cosmo)) * growth(a, cosmo, conf, deriv=1))


def _G_D_dOm(a, cosmo, conf):
    e2 = E2(a, cosmo)
    d1 = growth(a, cosmo, conf, deriv=1)
    return a**2 * (0.5 / np.sqrt(e2) * E2_dOm(a, cosmo) * d1
                   + np.sqrt(e2) * growth_dOm(a, cosmo, conf, deriv=1))


def _G_K_dOm(a, cosmo, conf):
    e2 = E2(a, cosmo)
    h = H_deriv(a, cosmo)
    d1 = growth(a, cosmo, conf, deriv=1)
    d2 = growth(a, cosmo, conf, deriv=2)
    dd1 = growth_dOm(a, cosmo, conf, deriv=1)
    dd2 = growth_dOm(a, cosmo, conf, deriv=2)
    return a**3 * (E2_dOm(a, cosmo) * (d2 + (2 + h) * d1)
                   + e2 * (dd2 + H_deriv_dOm(a, cosmo) * d1 + (2 + h) * dd1))


def drift_factor(a_vel, a_prev, a_next, cosmo, conf):
    """Drift time-step factor in [1/H_0] (nbody.py:25-29)."""
    return (growth(a_next, cosmo, conf) - growth(a_prev, cosmo, conf)) / _G_D(a_vel, cosmo, conf)


def kick_factor(a_acc, a_prev, a_next, cosmo, conf):
    """Kick time-step factor in [1/H_0] (nbody.py:32-36)."""
    return (_G_D(a_next, cosmo, conf) - _G_D(a_prev, cosmo, conf)) / _G_K(a_acc, cosmo, conf)


def drift_factor_dOm(a_vel, a_prev, a_next, cosmo, conf):
    f = drift_factor(a_vel, a_prev, a_next, cosmo, conf)
    gd = _G_D(a_vel, cosmo, conf)
    num = growth_dOm(a_next, cosmo, conf) - growth_dOm(a_prev, cosmo, conf)
    return num / gd - f * _G_D_dOm(a_vel, cosmo, conf) / gd


def kick_factor_dOm(a_acc, a_prev, a_next, cosmo, conf):
    f = kick_factor(a_acc, a_prev, a_next, cosmo, conf)
    gk = _G_K(a_acc, cosmo, conf)
    num = _G_D_dOm(a_next, cosmo, conf) - _G_D_dOm(a_prev, cosmo, conf)
    return num / gk - f * _G_K_dOm(a_acc, cosmo, conf) / gk


def _f32(x):
    return float(np.float32(x))


# parameters besides Omega_m that enter the step factors (through E2 and
# H_deriv, cosmology.py:185-242, and the growth table); the others (A_s, n_s,
# Omega_b, h) do not enter the N-body step
_DE_PARAMS = ('Omega_k_', 'w_0_', 'w_a_')
_DE_STEP = 1e-6


class _CosmoCot:
    """Cosmology cotangent of the reverse-time adjoint (nbody.py:49-99).

    Each kick and drift adjoint subtracts d(factor)/d(theta) times its dot
    product, and each kick adds the force term: acc is proportional to
    Omega_m, so the force VJP's Omega_m cotangent is sum(vel_cot*acc)/Omega_m
    (gravity.py:54), and no other parameter enters the force.  The factor
    derivatives are total derivatives, growth table included: Omega_m's from
    the growth ODE's forward sensitivities; Omega_k_, w_0_ and w_a_ (only
    those set, not None) by a central difference of the factors over growth
    tables integrated at rtol = atol = 1e-12.
    """

    def __init__(self, cosmo, conf):
        self.cosmo, self.conf = cosmo, conf
        self.cot = {'Omega_m': 0.0}
        self.pert = {}
        tight = conf.replace(growth_rtol=1e-12, growth_atol=1e-12)
        for p in _DE_PARAMS:
            v = getattr(cosmo, p)
            if v is None:
                continue
            cs = [boltzmann(cosmo.replace(**{p: v + s * _DE_STEP}, transfer=None, growth=None,
                                          growth_dOm=None, varlin=None),
                            tight, transfer=False, varlin=False) for s in (1, -1)]
            self.pert[p] = (cs, tight)
            self.cot[p] = 0.0

    def _d(self, fn, args, p):
        (cp, cm), tight = self.pert[p]
        return (fn(*args, cp, tight) - fn(*args, cm, tight)) / (2 * _DE_STEP)

    def kick(self, args, dot, cf, f):
        """kick_adj (nbody.py:80-99): dot = sum(vel_cot*acc), cf the force
        VJP's Omega_m cotangent, f the float32 factor."""
        self.cot['Omega_m'] -= kick_factor_dOm(*args, self.cosmo, self.conf) * dot + cf * f
        for p in self.pert:
            self.cot[p] -= self._d(kick_factor, args, p) * dot

    def drift(self, args, dot):
        """drift_adj (nbody.py:49-67): dot = sum(disp_cot*vel)."""
        self.cot['Omega_m'] -= drift_factor_dOm(*args, self.cosmo, self.conf) * dot
        for p in self.pert:
            self.cot[p] -= self._d(drift_factor, args, p) * dot


def _is_kdk(conf):
    return tuple(tuple(s) for s in conf.symp_splits) == ((0, 0.5), (1, 0.5))


def step_ops(a_prev, a_next, conf):
    """Op list of one step of ``integrate`` (nbody.py:121-140)."""
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
    return ops


# ------------------------------------------------------- forward loop ---

def _mass_val(conf):
    return float(np.float32(conf.mesh_size / conf.ptcl_num))


def _force(c, pmid, disp, acc, n, val, coef):
    _lib.check(_lib.lib().pmx_gravity(c.h, _lib.ptr(pmid), _lib.ptr(disp), n, val, coef,
                                      _lib.ptr(acc)))


_PIN_CHUNK = 1 << 30  # host checkpoints are pinned in chunks of at most 1 GiB


class _HostCopy:
    """A device tensor's copy in host memory for a reversal checkpoint:
    pinned chunks of at most 1 GiB filled by asynchronous copies on the
    current stream (pageable where pinning fails); ``to(device)`` brings it
    back."""

    def __init__(self, t):
        flat = t.detach().reshape(-1)
        self.shape, self.dtype = tuple(t.shape), t.dtype
        per = max(1, _PIN_CHUNK // t.element_size())
        self.chunks = []
        for i in range(0, flat.numel(), per):
            src = flat[i:i + per]
            try:
                h = torch.empty(src.numel(), dtype=t.dtype, pin_memory=True)
                h.copy_(src, non_blocking=True)
            except RuntimeError:
                h = src.cpu()
            self.chunks.append(h)

    @property
    def device(self):
        return torch.device('cpu')

    def copy_to(self, dst):
        """Copies the stored values into the device tensor dst (no device
        allocation)."""
        flat = dst.view(-1)
        o = 0
        for h in self.chunks:
            flat[o:o + h.numel()].copy_(h, non_blocking=h.is_pinned())
            o += h.numel()
        return dst

    def to(self, device):
        return self.copy_to(torch.empty(self.shape, dtype=self.dtype, device=device))


def _restore(dst, saved):
    """dst <- a stored checkpoint array (device tensor or _HostCopy)."""
    if isinstance(saved, _HostCopy):
        saved.copy_to(dst)
    else:
        dst.copy_(saved)


def _checkpoint_where(where, t, count):
    """'auto': the device when `count` copies of t (disp and vel each) take
    at most half the device's free memory, else host memory."""
    if where != 'auto':
        return where
    free, _ = torch.cuda.mem_get_info(t.device)
    free += torch.cuda.memory_reserved(t.device) - torch.cuda.memory_allocated(t.device)
    need = 2 * count * t.numel() * t.element_size()
    return t.device if need <= 0.5 * free else 'cpu'


def _store(t, where):
    """A copy of t for a reversal checkpoint: on the device, or in host
    memory when ``where`` is 'cpu'."""
    if where is None or torch.device(where) == t.device:
        return t.clone()
    return _HostCopy(t)


def checkpoint_steps(checkpoint_every, nsteps):
    """Forward step indices before which the state is stored: every k-th step
    for an int k (k, 2k, ...; none for 0), or the steps of a sequence; each
    in [1, nsteps - 1]."""
    if checkpoint_every is None or checkpoint_every is False:
        return frozenset()
    if isinstance(checkpoint_every, (int, np.integer)):
        k = int(checkpoint_every)
        if k < 0:
            raise ValueError(f'checkpoint_every={k} must not be negative')
        return frozenset(range(k, nsteps, k)) if k else frozenset()
    steps = frozenset(int(i) for i in checkpoint_every)
    bad = sorted(i for i in steps if not 0 < i < nsteps)
    if bad:
        raise ValueError(f'checkpoint steps {bad} outside [1, {nsteps - 1}]')
    return steps


def kdk_factor_table(a_prev, a_next, cosmo, conf):
    """float32 (k1, d, k2) rows of the KDK steps a_prev[i] -> a_next[i]
    (nbody.py:205-212), evaluated on the arrays at once: the same float64
    operations per step as ``_kdk_factors``."""
    a_prev = np.asarray(a_prev, dtype=np.float64)
    a_next = np.asarray(a_next, dtype=np.float64)
    a_half = a_prev * (1 - 0.5) + a_next * 0.5
    k1 = kick_factor(a_prev, a_prev, a_half, cosmo, conf)
    d = drift_factor(a_half, a_prev, a_next, cosmo, conf)
    k2 = kick_factor(a_next, a_half, a_next, cosmo, conf)
    return np.stack([k1, d, k2], axis=-1).astype(np.float32)


def _kdk_factors(a_prev, a_next, cosmo, conf):
    """float32 k1, d, k2 of one KDK step (nbody.py:205-212)."""
    a_half = a_prev * (1 - 0.5) + a_next * 0.5
    return (_f32(kick_factor(a_prev, a_prev, a_half, cosmo, conf)),
            _f32(drift_factor(a_half, a_prev, a_next, cosmo, conf)),
            _f32(kick_factor(a_next, a_half, a_next, cosmo, conf)))


def _run_forward(pmid, disp, vel, cosmo, conf, reverse=False, checkpoint_every=None,
                 checkpoint_device=None, pipeline=True, pctx=None):
    """Integrates in place; returns acc and the checkpoints {step: (disp, vel)}.

    KDK steps between checkpoints run as one ``pmx_nbody_steps`` call (the
    particles kept in tile order on the device between steps), or one
    ``pmx_nbody_step`` call per step when ``pipeline`` is False; both give
    the same bits.  ``pctx`` is the library context to run on (default: the
    one ``_lib.context`` gives this thread).
    """
    L = _lib.lib()
    n = pmid.shape[0]
    c = pctx or _lib.context(conf.mesh_shape, conf.cell_size, disp.device)
    h = c.bind()
    val, coef = _mass_val(conf), omega_coef(cosmo, conf)
    acc = torch.empty_like(disp)
    a_nbody = conf.a_nbody[::-1] if reverse else conf.a_nbody
    _force(c, pmid, disp, acc, n, val, coef)
    ckpt = {}
    steps = list(zip(a_nbody[:-1], a_nbody[1:]))
    cset = checkpoint_steps(checkpoint_every, len(steps))
    # 'auto' decides at the first checkpoint, once the pipeline holds its scratch
    ckpt_where = checkpoint_device if checkpoint_device != 'auto' else None
    n_ckpt = len(cset)
    if _is_kdk(conf):
        i = 0
        while i < len(steps):
            if i in cset:
                if checkpoint_device == 'auto' and not ckpt:
                    ckpt_where = _checkpoint_where('auto', disp, n_ckpt)
                ckpt[i] = (_store(disp, ckpt_where), _store(vel, ckpt_where))
            j = min([t for t in cset if t > i], default=len(steps))
            if not pipeline:
                j = i + 1
            seg = np.array(steps[i:j], dtype=np.float64).reshape(-1, 2)
            f = kdk_factor_table(seg[:, 0], seg[:, 1], cosmo, conf)
            k1, d, k2 = (np.ascontiguousarray(f[:, m]) for m in range(3))
            if j - i == 1:
                _lib.check(L.pmx_nbody_step(h, _lib.ptr(pmid), _lib.ptr(disp), _lib.ptr(vel),
                                            _lib.ptr(acc), n, val, coef, float(k1[0]),
                                            float(d[0]), float(k2[0])))
            else:
                P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
                _lib.check(L.pmx_nbody_steps(h, _lib.ptr(pmid), _lib.ptr(disp), _lib.ptr(vel),
                                             _lib.ptr(acc), n, val, coef, j - i, P(k1), P(d),
                                             P(k2)))
            i = j
        return acc, ckpt
    for i, (a_prev, a_next) in enumerate(steps):
        if i in cset:
            if checkpoint_device == 'auto' and not ckpt:
                ckpt_where = _checkpoint_where('auto', disp, n_ckpt)
            ckpt[i] = (_store(disp, ckpt_where), _store(vel, ckpt_where))
        for op in step_ops(a_prev, a_next, conf):
            if op[0] == 'drift':
                f = _f32(drift_factor(*op[1:], cosmo, conf))
                _lib.check(L.pmx_drift(h, n, _lib.ptr(disp), _lib.ptr(vel), f))
            elif op[0] == 'kick':
                f = _f32(kick_factor(*op[1:], cosmo, conf))
                _lib.check(L.pmx_kick(h, n, _lib.ptr(vel), _lib.ptr(acc), f))
            else:
                _force(c, pmid, disp, acc, n, val, coef)
    return acc, ckpt


# ------------------------------------------------------------ adjoint ---

def _adj_pipeline_fits(n, c, device, extra=0):
    """Whether the reverse pipeline's particle copies (about 132 B per record
    slot, 1.2 slots per particle) may fit beside ``extra`` bytes the caller
    still has to allocate: in free device memory, plus what the context
    already holds (the forward pipeline's copies and scratch, which the
    reverse pipeline reuses), plus torch's cached free blocks -- counted as
    they are when the contexts share torch's pool, else given back to the
    device first.  The allocation itself decides; this only skips a hopeless
    attempt."""
    need = 132 * 1.2 * n + extra + (2 << 30)
    free, _ = torch.cuda.mem_get_info(device)
    if _lib.torch_pool():
        cached = torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
        return need < free + cached + c.nbytes
    if need < free + c.nbytes:
        return True
    with _lib.capture_lock:   # a device-wide wait
        torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info(device)
    return need < free + c.nbytes


def _run_adjoint(pmid, disp, vel, disp_cot, vel_cot, cosmo, conf, reverse=False, ckpt=None,
                 fused=True, pipeline=None, cot_out=None, pctx=None, free_ckpt=False):
    """Reverse-time adjoint in place (nbody.py:143-162, 226-260).

    On return disp/vel hold the reconstructed initial state and
    disp_cot/vel_cot the initial-state cotangents; returns (acc, acc_cot,
    Omega_m cotangent).  KDK steps between checkpoints run as one
    ``pmx_nbody_steps_adj`` call (particles and cotangents kept in tile
    order on the device) when ``pipeline`` is True -- by default when the
    device has room, falling back to the per-step path when the pipeline's
    allocation fails (the library reports that before it touches any array)
    -- else as one ``pmx_nbody_step_adj`` call per step; both give the same
    arrays.  With ``fused`` False every splitting runs op by op.  With
    ``free_ckpt`` each checkpoint leaves ``ckpt`` once used, so that its
    memory returns during the reverse run.
    """
    L = _lib.lib()
    n = pmid.shape[0]
    c = pctx or _lib.context(conf.mesh_shape, conf.cell_size, disp.device)
    kdk = fused and _is_kdk(conf)
    auto = pipeline is None
    if auto:
        pipeline = kdk and _adj_pipeline_fits(n, c, disp.device, extra=24 * n)  # acc, acc_cot
    if not pipeline:
        # the reverse steps bin per step; the forward pipeline's particle
        # copy gives its memory back to the cotangents
        c.trim(pipeline_only=True)
    h = c.bind()
    val, coef = _mass_val(conf), omega_coef(cosmo, conf)
    Om = cosmo.Omega_m
    acc = torch.empty_like(disp)
    acc_cot = torch.empty_like(disp)
    dot = ctypes.c_double(0.0)

    def force_adj():
        _lib.check(L.pmx_gravity_vjp(h, _lib.ptr(pmid), _lib.ptr(disp), n, val, coef,
                                     _lib.ptr(vel_cot), _lib.ptr(acc), _lib.ptr(acc_cot),
                                     ctypes.byref(dot)))
        return dot.value / Om

    a_nbody = conf.a_nbody[::-1] if reverse else conf.a_nbody
    nsteps = len(a_nbody) - 1
    cf = force_adj()
    dot_force = dot.value
    cc = _CosmoCot(cosmo, conf)
    dots = (ctypes.c_double * 2)()
    kp = 0.0  # closing kick of the previous KDK step, applied by the next call
    rev = list(zip(a_nbody[:0:-1], a_nbody[-2::-1]))
    if kdk and pipeline:
        s = 0
        while s < nsteps:
            # one library call per run of steps ending at a checkpoint (snapped)
            e = s
            while e < nsteps - 1 and not (ckpt and nsteps - 1 - e in ckpt):
                e += 1
            args = []
            for a_prev, a_next in rev[s:e + 1]:
                a_half = a_prev * (1 - 0.5) + a_next * 0.5
                args.append(((a_prev, a_prev, a_half), (a_half, a_prev, a_next),
                             (a_next, a_half, a_next)))
            f = np.array([[_f32(kick_factor(*ka, cosmo, conf)),
                           _f32(drift_factor(*da, cosmo, conf)),
                           _f32(kick_factor(*kb, cosmo, conf))] for ka, da, kb in args],
                         dtype=np.float32)
            fa, fd, fb = (np.ascontiguousarray(f[:, m]) for m in range(3))
            step_e = nsteps - 1 - e
            snap = ckpt[step_e][0].to(disp.device) if ckpt and step_e in ckpt else None
            k = e - s + 1
            dd = (ctypes.c_double * (2 * k))()
            P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
            rc = L.pmx_nbody_steps_adj(
                h, _lib.ptr(pmid), _lib.ptr(disp), _lib.ptr(vel), _lib.ptr(acc),
                _lib.ptr(disp_cot), _lib.ptr(vel_cot), _lib.ptr(acc_cot), n, val, coef, kp, k,
                P(fa), P(fd), P(fb), _lib.ptr(snap), dd)
            if rc == _lib.ENOMEM and auto and s == 0:
                # no room after all; nothing has changed yet
                c.trim(pipeline_only=True)
                pipeline = False
                break
            _lib.check(rc)
            for j, (ka_args, d_args, kb_args) in enumerate(args):
                cc.kick(ka_args, dot_force, cf, float(fa[j]))
                cc.drift(d_args, dd[2 * j])
                dot_force = dd[2 * j + 1]
                cf = dot_force / Om
                cc.kick(kb_args, dot_force, cf, float(fb[j]))
            kp = float(fb[-1])
            if ckpt and step_e in ckpt:
                _kick_adj(L, h, n, vel, acc, disp_cot, acc_cot, vel_cot, kp)
                kp = 0.0
                _restore(vel, (ckpt.pop(step_e) if free_ckpt else ckpt[step_e])[1])
            del snap
            s = e + 1
    if kdk and pipeline:
        if kp != 0.0:
            _kick_adj(L, h, n, vel, acc, disp_cot, acc_cot, vel_cot, kp)
        if cot_out is not None:
            cot_out.update(cc.cot)
        return acc, acc_cot, cc.cot['Omega_m']
    for s, (a_prev, a_next) in enumerate(rev):
        step = nsteps - 1 - s  # index of the forward step being undone
        if kdk:
            # kick_adj, drift_adj, force_adj in one library call, the closing
            # kick_adj deferred into the next call; the kick adjoints'
            # sum(vel_cot*acc) is the dot of the force VJP that produced acc
            a_half = a_prev * (1 - 0.5) + a_next * 0.5
            ka_args = (a_prev, a_prev, a_half)
            d_args = (a_half, a_prev, a_next)
            kb_args = (a_next, a_half, a_next)
            fa = _f32(kick_factor(*ka_args, cosmo, conf))
            fd = _f32(drift_factor(*d_args, cosmo, conf))
            fb = _f32(kick_factor(*kb_args, cosmo, conf))
            # the step's drift result is replaced by the stored state: copied
            # into disp first, which the step then reads as its snapshot (each
            # particle's value is read before its new one is written)
            snap = None
            if ckpt and step in ckpt:
                _restore(disp, ckpt[step][0])
                snap = disp
            _lib.check(L.pmx_nbody_step_adj(
                h, _lib.ptr(pmid), _lib.ptr(disp), _lib.ptr(vel), _lib.ptr(acc),
                _lib.ptr(disp_cot), _lib.ptr(vel_cot), _lib.ptr(acc_cot), n, val, coef,
                kp, fa, fd, _lib.ptr(snap), dots))
            cc.kick(ka_args, dot_force, cf, fa)
            cc.drift(d_args, dots[0])
            dot_force = dots[1]
            cf = dot_force / Om
            cc.kick(kb_args, dot_force, cf, fb)
            kp = fb
            if ckpt and step in ckpt:
                _kick_adj(L, h, n, vel, acc, disp_cot, acc_cot, vel_cot, kp)
                kp = 0.0
                _restore(vel, (ckpt.pop(step) if free_ckpt else ckpt[step])[1])
            del snap
            continue
        D = K = 0
        a_disp = a_vel = a_acc = a_prev
        for d, k in reversed(conf.symp_splits):
            if k != 0:
                K += k
                a_vel_next = a_prev * (1 - K) + a_next * K
                args = (a_acc, a_vel, a_vel_next)
                f = _f32(kick_factor(*args, cosmo, conf))
                _lib.check(L.pmx_kick_adj(h, n, _lib.ptr(vel), _lib.ptr(acc), _lib.ptr(disp_cot),
                                          _lib.ptr(acc_cot), _lib.ptr(vel_cot), f,
                                          ctypes.byref(dot)))
                cc.kick(args, dot.value, cf, f)
                a_vel = a_vel_next
            if d != 0:
                D += d
                a_disp_next = a_prev * (1 - D) + a_next * D
                args = (a_vel, a_disp, a_disp_next)
                f = _f32(drift_factor(*args, cosmo, conf))
                _lib.check(L.pmx_drift_adj(h, n, _lib.ptr(disp), _lib.ptr(vel),
                                           _lib.ptr(disp_cot), _lib.ptr(vel_cot), f,
                                           ctypes.byref(dot)))
                cc.drift(args, dot.value)
                a_disp = a_disp_next
                if ckpt and step in ckpt and D == 1:
                    disp.copy_(ckpt[step][0].to(disp.device))
                cf = force_adj()
                a_acc = a_disp
        if ckpt and step in ckpt:
            vel.copy_(ckpt[step][1].to(vel.device))
    if kp != 0.0:
        _kick_adj(L, h, n, vel, acc, disp_cot, acc_cot, vel_cot, kp)
    if cot_out is not None:
        cot_out.update(cc.cot)
    return acc, acc_cot, cc.cot['Omega_m']


def _kick_adj(L, h, n, vel, acc, disp_cot, acc_cot, vel_cot, k):
    _lib.check(L.pmx_kick_adj(h, n, _lib.ptr(vel), _lib.ptr(acc), _lib.ptr(disp_cot),
                              _lib.ptr(acc_cot), _lib.ptr(vel_cot), k, None))


class _NBodyFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, disp, vel, om_t, pmid, cosmo, conf, reverse, checkpoint_every,
                checkpoint_device):
        disp_f = disp.detach().clone().contiguous()
        vel_f = vel.detach().clone().contiguous()
        # the backward runs on the same library context: autograd may call it
        # on another thread, where the caller's context provider is not in scope
        ctx.pmx = _lib.context(conf.mesh_shape, conf.cell_size, disp.device)
        acc, ckpt = _run_forward(pmid, disp_f, vel_f, cosmo, conf, reverse, checkpoint_every,
                                 checkpoint_device, pctx=ctx.pmx)
        ctx.save_for_backward(disp_f, vel_f)
        ctx.pmid, ctx.cosmo, ctx.conf, ctx.reverse, ctx.ckpt = pmid, cosmo, conf, reverse, ckpt
        ctx.mark_non_differentiable(acc)
        return disp_f, vel_f, acc

    @staticmethod
    def backward(ctx, disp_cot, vel_cot, acc_cot):
        disp_f, vel_f = ctx.saved_tensors
        conf = ctx.conf
        pipeline = None
        if _is_kdk(conf):
            n = ctx.pmid.shape[0]
            c = ctx.pmx
            # four working copies plus acc and acc_cot: 72 B per particle
            if not _adj_pipeline_fits(n, c, disp_f.device, extra=72 * n):
                # per-step reverse calls: the forward pipeline's copy goes back
                # to the pool before the working copies are allocated
                c.trim(pipeline_only=True)
                pipeline = False
        disp = disp_f.clone()
        vel = vel_f.clone()
        dc = (disp_cot if disp_cot is not None else torch.zeros_like(disp)).clone().contiguous()
        vc = (vel_cot if vel_cot is not None else torch.zeros_like(vel)).clone().contiguous()
        # the checkpoints are not needed again (no double backward)
        _, _, om_cot = _run_adjoint(ctx.pmid, disp, vel, dc, vc, ctx.cosmo, conf,
                                    ctx.reverse, ctx.ckpt, pipeline=pipeline, pctx=ctx.pmx,
                                    free_ckpt=True)
        om_grad = None
        if ctx.needs_input_grad[2]:
            om_grad = torch.tensor(om_cot, dtype=torch.float64)
        return dc, vc, om_grad, None, None, None, None, None, None


def nbody(ptcl, obsvbl, cosmo, conf, reverse=False, checkpoint_every=None, omega_m=None,
          checkpoint_device=None):
    """N-body time integration (nbody.py:215-223).

    Returns ``(ptcl, obsvbl)`` with the final ``disp``, ``vel`` and ``acc``.
    Differentiable w.r.t. ``ptcl.disp``/``ptcl.vel`` (reverse-time adjoint)
    and, when ``omega_m`` is a 0-d float64 tensor equal to
    ``cosmo.Omega_m`` that requires grad, w.r.t. Omega_m.  With
    ``checkpoint_every`` the adjoint restarts from states stored every that
    many steps (an int) or before the given forward steps (a sequence, as
    ``reversal_checkpoints`` returns), kept on the device (None), in host
    memory ('cpu': pinned 1 GiB chunks, copied asynchronously) or where they
    fit ('auto': the device when, at the first checkpoint, they take at most
    half of its free memory).  The backward frees each checkpoint once it
    has used it.
    """
    if cosmo.growth is None:
        raise ValueError('Growth table is empty. Call growth_integ or boltzmann first.')
    _lib.require_device(ptcl.disp, 'ptcl.disp')
    vel = ptcl.vel if ptcl.vel is not None else torch.zeros_like(ptcl.disp)
    pmid = ptcl.pmid.contiguous()
    need_grad = torch.is_grad_enabled() and (ptcl.disp.requires_grad or vel.requires_grad or
                                             (omega_m is not None and omega_m.requires_grad))
    if need_grad:
        om_t = omega_m if omega_m is not None else torch.tensor(cosmo.Omega_m, dtype=torch.float64)
        disp, vel, acc = _NBodyFn.apply(ptcl.disp, vel, om_t, pmid, cosmo, conf, reverse,
                                        checkpoint_every, checkpoint_device)
    else:
        disp = ptcl.disp.detach().clone().contiguous()
        vel = vel.detach().clone().contiguous()
        acc, _ = _run_forward(pmid, disp, vel, cosmo, conf, reverse, None)
    return ptcl.replace(disp=disp, vel=vel, acc=acc), obsvbl


def nbody_adj(ptcl, ptcl_cot, cosmo, conf, reverse=False):
    """Functional reverse-time adjoint from a final state (nbody.py:251-260).

    ``ptcl`` is the final state, ``ptcl_cot`` its cotangent (disp, vel).
    Returns ``(ptcl0, ptcl_cot0, cosmo_cot)``: the reconstructed initial
    state, the initial cotangents (with ``acc`` holding the last force VJP,
    as pmwd returns it) and a ``Cosmology`` cotangent.  Its Omega_m entry,
    and the Omega_k_, w_0_ and w_a_ entries of the parameters that are set,
    are total derivatives (the growth table's dependence included); A_s_1e9,
    n_s, Omega_b and h do not enter the N-body step and are 0; unset
    parameters stay None, as in pmwd's cotangent pytree.
    """
    pmid = ptcl.pmid.contiguous()
    disp = ptcl.disp.detach().clone().contiguous()
    vel = ptcl.vel.detach().clone().contiguous()
    dc = ptcl_cot.disp.detach().clone().contiguous()
    vc = ptcl_cot.vel.detach().clone().contiguous()
    cot = {}
    acc, acc_cot, om_cot = _run_adjoint(pmid, disp, vel, dc, vc, cosmo, conf, reverse,
                                        cot_out=cot)
    de = {p: (cot[p] if getattr(cosmo, p) is not None else None) for p in _DE_PARAMS}
    cosmo_cot = cosmo.replace(A_s_1e9=0.0, n_s=0.0, Omega_m=om_cot, Omega_b=0.0, h=0.0,
                              transfer=None, growth=None, growth_dOm=None, varlin=None, **de)
    return (ptcl.replace(disp=disp, vel=vel, acc=acc),
            ptcl_cot.replace(disp=dc, vel=vc, acc=acc_cot), cosmo_cot)


def reversal_checkpoints(ptcl, cosmo, conf, tol=1e-4, candidates=1, checkpoint_device='auto'):
    """The forward steps at which float32 reversal needs a checkpoint:
    those where the reconstructed state drifts from the forward one by more
    than ``tol``, relative: the larger of ||disp_rev - disp|| / ||disp|| and
    the same for vel (L2 over all particles).

    Runs ``nbody`` forward from ``ptcl`` storing the state before every
    ``candidates``-th step (an int) or before the given steps (a sequence),
    then integrates back from the final state one KDK step at a time, as the
    adjoint reconstructs the trajectory.  Walking back, at each stored step
    it measures the drift; where it exceeds ``tol`` the step is marked and
    the reconstruction continues from the stored state, as the adjoint would
    with a checkpoint there.  Returns ``(steps, drift)``: the marked steps,
    ascending, to pass as ``nbody``'s ``checkpoint_every``, and {step:
    relative drift} at every candidate.  No pmwd counterpart: pmwd reverses
    without checkpoints.
    """
    if not _is_kdk(conf):
        raise NotImplementedError('reversal_checkpoints needs the KDK splitting')
    _lib.require_device(ptcl.disp, 'ptcl.disp')
    L = _lib.lib()
    pmid = ptcl.pmid.contiguous()
    n = pmid.shape[0]
    a = conf.a_nbody
    nsteps = len(a) - 1
    with torch.no_grad():
        disp = ptcl.disp.detach().clone().contiguous()
        vel = (ptcl.vel if ptcl.vel is not None else torch.zeros_like(disp)).detach().clone()
        vel = vel.contiguous()
        c = _lib.context(conf.mesh_shape, conf.cell_size, disp.device)
        acc, ckpt = _run_forward(pmid, disp, vel, cosmo, conf, False, candidates,
                                 checkpoint_device, pctx=c)
        h = c.bind()
        val, coef = _mass_val(conf), omega_coef(cosmo, conf)
        f = kdk_factor_table(a[1:][::-1], a[:-1][::-1], cosmo, conf)  # reverse steps
        ref = torch.empty_like(disp)
        marks, drift = [], {}
        for r, s in enumerate(range(nsteps - 1, -1, -1)):   # undoes forward step s
            k1, d, k2 = (float(x) for x in f[r])
            _lib.check(L.pmx_nbody_step(h, _lib.ptr(pmid), _lib.ptr(disp), _lib.ptr(vel),
                                        _lib.ptr(acc), n, val, coef, k1, d, k2))
            if s in ckpt:
                sd, sv = ckpt.pop(s)
                rel = []
                for now, saved in ((disp, sd), (vel, sv)):
                    _restore(ref, saved)
                    e = float(torch.linalg.vector_norm((now - ref).double()))
                    r = float(torch.linalg.vector_norm(ref.double()))
                    rel.append(e / r if r > 0 else e)
                drift[s] = max(rel)
                if drift[s] > tol:
                    marks.append(s)
                    _restore(disp, sd)
                    _restore(vel, sv)
                    _force(c, pmid, disp, acc, n, val, coef)
                del sd, sv
    return sorted(marks), drift
