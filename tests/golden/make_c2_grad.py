"""Writes tests/golden/c2_grad.npz: float64 oracle derivatives of the two
objectives of the 128^3-particle / 256^3-mesh configuration's gradient
tests, by central differences of the CPU oracle (float64 throughout, where
differences are exact to far below the GPU's float32 rounding).

    python tests/golden/make_c2_grad.py          (about an hour on 8 cores)

* the 2LPT objective J = <a, disp> + <b, vel> of the initial conditions
  (white noise seed 0 -> linear_modes -> lpt at a = 1/64, pmwd/modes.py,
  lpt.py), a and b standard normals (numpy default_rng(11), shape (N,3)):
  dJ/dOmega_m and dJ/dsigma8 (A_s following sigma8 as Cosmology.from_sigma8)
  and dJ along three directions of the white modes (default_rng(12));
* the density loss mean((rho - rho_tgt)^2) after the 63 KDK steps to a = 1
  (nbody.py, scatter.py), rho_tgt from white noise seed 1 at sigma8 = 0.85:
  its derivatives with respect to Omega_m and sigma8 and along three
  directions of the white modes (default_rng(7)).

Steps: 1e-6 in Omega_m and sigma8 and 1e-6 times the direction for the
N-body loss (small enough that few particles change CIC cell between the
two sides; the growth ODE is solved to 1e-12 so that it adds no noise),
1e-4 / 1e-2 for the smooth 2LPT objective (quadratic in the modes and
sigma8).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import cosmo as oc, ic as oic, nbody as onb, pm as opm  # noqa: E402
from oracle.conf import OConf  # noqa: E402

F64 = np.float64
# the growth ODE solved to 1e-12, so that its adaptive steps add no noise
# to differences over 1e-6 in Omega_m
CONF = OConf(2.0, (128,) * 3, mesh_shape=2, a_start=1 / 64, a_nbody_maxstep=1 / 64,
             growth_rtol=1e-12, growth_atol=1e-12)
COSMO = dict(n_s=0.96, Omega_b=0.05, h=0.7)


def cosmo(Om=0.3, s8=0.8):
    return oc.boltzmann(oc.from_sigma8(CONF, s8, Omega_m=Om, **COSMO), CONF)


def ics(white, c):
    return oic.lpt(oic.linear_modes(white, c, CONF, dtype=F64), c, CONF, dtype=F64)


def density(white, c):
    pmid, d0, v0 = ics(white, c)
    d1, _, _ = onb.nbody(pmid, d0, v0, c, CONF, dtype=F64)
    return opm.scatter(pmid, d1, CONF.cell_size, CONF.mesh_shape, dtype=F64)


def central(f, h):
    return (f(h) - f(-h)) / (2 * h)


def log(*a):
    print(f'[{time.strftime("%H:%M:%S")}]', *a, flush=True)


def ic_objective():
    """(J, white noise): the 2LPT objective <a, disp> + <b, vel> as a
    function of the white modes, Omega_m and sigma8."""
    n = CONF.ptcl_num
    w0 = oic.white_noise(0, CONF, real=True).astype(F64)
    rng = np.random.default_rng(11)
    a, b = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))

    def J(white=w0, Om=0.3, s8=0.8):
        _, d0, v0 = ics(white, cosmo(Om, s8))
        return float((a * d0).sum() + (b * v0).sum())
    return J, w0


def main():
    J, w0 = ic_objective()
    out = {'ic_J': J()}
    out['ic_dOm'] = central(lambda h: J(Om=0.3 + h), 1e-4)
    out['ic_ds8'] = central(lambda h: J(s8=0.8 + h), 1e-2)
    rng = np.random.default_rng(12)
    dirs = []
    for _ in range(3):
        u = rng.standard_normal(CONF.ptcl_grid_shape).astype(np.float32).astype(F64)
        dirs.append(central(lambda e: J(white=w0 + e * u), 1e-2))
    out['ic_dirs'] = np.array(dirs)
    log('2LPT objective', out)

    tgt = density(oic.white_noise(1, CONF, real=True).astype(F64), cosmo(s8=0.85))
    log('target done')

    def loss(white=w0, Om=0.3, s8=0.8):
        r = float(((density(white, cosmo(Om, s8)) - tgt)**2).mean())
        log('loss', Om, s8, r)
        return r

    out['loss'] = loss()
    out['loss_dOm'] = central(lambda h: loss(Om=0.3 + h), 1e-6)
    out['loss_ds8'] = central(lambda h: loss(s8=0.8 + h), 1e-6)
    rng = np.random.default_rng(7)
    dirs = []
    for _ in range(3):
        u = rng.standard_normal(CONF.ptcl_grid_shape).astype(np.float32).astype(F64)
        dirs.append(central(lambda e: loss(white=w0 + e * u), 1e-6))
    out['loss_dirs'] = np.array(dirs)
    log('N-body loss', out)
    np.savez(os.path.join(HERE, 'c2_grad.npz'), **out)


if __name__ == '__main__':
    main()
