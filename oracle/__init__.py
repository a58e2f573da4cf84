"""CPU oracle for the pmwd particle-mesh hot path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference algorithms of
eelregit/pmwd (``pmwd/*.py``) in numpy/scipy so that the HIP implementation in
``pmwd_amd`` can be checked against them.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline -- never as part
of the product path.

Modules
-------
conf   : restated ``Configuration`` derived quantities (configuration.py)
cosmo  : E2, H_deriv, Eisenstein-Hu transfer, Dormand-Prince growth ODE,
         linear power, top-hat variance, sigma8 (cosmology.py, boltzmann.py,
         ode_util.py)
pm     : enmesh, cell keys, scatter/gather and their adjoints, fftfreq,
         laplace, neg_grad, gravity and its VJP (pm_util.py, scatter.py,
         gather.py, gravity.py)
nbody  : drift/kick factors, KDK integrator, reverse-time adjoint (nbody.py)
ic     : white noise, linear_modes, 2LPT (modes.py, lpt.py)
"""
