"""Writes the golden fixtures of tests/golden from the CPU oracle.

    python tests/golden/make_golden.py

Fixtures (inputs and expected outputs, numpy .npz):

* cic_even.npz, cic_odd.npz -- CIC cases (8^3 particles on a 16^3 mesh,
  and 7x9x5 particles on their own odd mesh): pmid, disp, cell keys, stable
  sort permutation, enmesh indices and weights, scatter, gather of a random
  mesh, scatter and gather adjoints (pmwd/pm_util.py, scatter.py, gather.py);
* run16.npz -- white noise (seed 0), linear modes, 2LPT at a=1/16 and four
  KDK steps for 16^3 particles on a 16^3 mesh in float32 and float64
  (pmwd/modes.py, lpt.py, nbody.py);
* tables.npz -- growth, transfer and sigma^2 tables of SimpleLCDM and the
  growth table of Einstein-de Sitter (pmwd/boltzmann.py).

The oracle is deterministic, so rerunning this script reproduces the files
bit for bit (tests/test_golden.py checks that).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import cosmo as oc, ic as oic, nbody as onb, pm as opm  # noqa: E402
from oracle.conf import OConf  # noqa: E402


def cic_case(ptcl_shape, ratio, cell, std, seed):
    conf = OConf(cell * ratio, ptcl_shape, mesh_shape=ratio)
    ms = conf.mesh_shape
    pmid, disp = opm.gen_grid(ptcl_shape, ms, conf.cell_size)
    rng = np.random.default_rng(seed)
    disp = (disp + std * rng.standard_normal(disp.shape)).astype(np.float32)
    n = len(pmid)
    ind, frac = opm.enmesh(pmid, disp, conf.cell_size, ms, 0, None, ms)
    keys = opm.cell_keys(pmid, disp, conf.cell_size, ms)
    mesh = rng.standard_normal(ms).astype(np.float32)
    mesh_cot = rng.standard_normal(ms).astype(np.float32)
    val_cot = rng.standard_normal(n).astype(np.float32)
    val = np.float32(conf.mesh_size / n)
    dens = opm.scatter(pmid, disp, conf.cell_size, ms, ptcl_num=n)
    g = opm.gather(pmid, disp, conf.cell_size, ms, mesh)
    sa_disp, sa_val = opm.scatter_adj(pmid, disp, conf.cell_size, ms, mesh_cot, val=val)
    ga_disp, ga_mesh, _ = opm.gather_adj(pmid, disp, conf.cell_size, ms, mesh, val_cot)
    return dict(ptcl_shape=np.array(ptcl_shape), mesh_shape=np.array(ms),
                cell_size=np.float64(conf.cell_size), pmid=pmid, disp=disp, keys=keys,
                perm=opm.sort_perm(keys), ind=ind, frac=frac, dens=dens, mesh=mesh, gather=g,
                mesh_cot=mesh_cot, val=val, scatter_adj_disp=sa_disp, scatter_adj_val=sa_val,
                val_cot=val_cot, gather_adj_disp=ga_disp, gather_adj_mesh=ga_mesh)


def run_case():
    conf = OConf(4.0, (16, 16, 16), mesh_shape=1, a_start=1 / 16, a_nbody_maxstep=1 / 16)
    cosmo = oc.boltzmann(oc.SimpleLCDM(conf), conf)
    white = oic.white_noise(0, conf, real=True)
    out = dict(white=white)
    for name, dt in (('f32', np.float32), ('f64', np.float64)):
        modes = oic.linear_modes(white, cosmo, conf, dtype=dt)
        pmid, disp, vel = oic.lpt(modes, cosmo, conf, dtype=dt)
        out[f'lpt_disp_{name}'] = disp
        out[f'lpt_vel_{name}'] = vel
        d, v, a = onb.nbody(pmid, disp, vel, cosmo, conf, dtype=dt, n_steps=4)
        out[f'disp_{name}'] = d
        out[f'vel_{name}'] = v
        out[f'acc_{name}'] = a
    out['pmid'] = pmid
    return out


def tables_case():
    conf = OConf(4.0, (16, 16, 16), mesh_shape=1)
    lcdm = oc.boltzmann(oc.SimpleLCDM(conf), conf)
    eds = oc.boltzmann(oc.SimpleLCDM(conf, Omega_m=1.0, Omega_b=0.05), conf)
    return dict(growth_a=conf.growth_a, transfer_k=conf.transfer_k, varlin_R=conf.varlin_R,
                lcdm_growth=lcdm.growth, lcdm_transfer=lcdm.transfer, lcdm_varlin=lcdm.varlin,
                eds_growth=eds.growth)


CASES = {
    'cic_even': lambda: cic_case((8, 8, 8), 2, 1.0, 2.0, 1),
    'cic_odd': lambda: cic_case((7, 9, 5), 1, 1.0, 3.0, 2),
    'run16': run_case,
    'tables': tables_case,
}


def main():
    for name, fn in CASES.items():
        path = os.path.join(HERE, f'{name}.npz')
        np.savez_compressed(path, **fn())
        print(path, os.path.getsize(path))


if __name__ == '__main__':
    main()
