"""pmwd_amd: pmwd's particle-mesh hot path on MI355X (gfx950) HIP kernels.

The API mirrors pmwd (eelregit/pmwd, pmwd/__init__.py): Configuration,
Cosmology, boltzmann tables, white_noise/linear_modes, lpt, nbody, scatter,
gather, gravity.  Device arrays are torch tensors; the compute runs in
``libpmx.so`` through its C ABI (include/pmx.h).
"""
from .configuration import Configuration
from .cosmology import Cosmology, SimpleLCDM, Planck18, E2, H_deriv, Omega_m_a
from .boltzmann import (transfer_integ, transfer_fit, transfer, growth_integ, growth,
                        varlin_integ, varlin, boltzmann, linear_power)
from .particles import Particles, ptcl_enmesh, ptcl_pos, ptcl_rpos, ptcl_rsd, ptcl_los
from .scatter import scatter
from .gather import gather
from .gravity import laplace, neg_grad, gravity
from .modes import white_noise, linear_modes
from .lpt import lpt
from .nbody import nbody, nbody_adj, reversal_checkpoints
from .spec_util import powspec

__version__ = '0.1.0'
