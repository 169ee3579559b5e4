"""femflow_amd: FEMFlow's explicit MLS-MPM substep, rebuilt for MI355X (gfx950).

The compute path is the HIP library ``_build/libmpm355.so`` (C ABI in
``include/mpm355.h``); the Python modules mirror the reference's interface:

* ``femflow_amd.solvers.mpm``: ``solve_mls_mpm_3d``, ``make_mls_mpm_coefficients``,
  ``three_d.{p2g,grid_op,g2p}``, ``two_d.{p2g,grid_op,g2p}``, ``Particle``.
* ``femflow_amd.simulation.mpm.MPMSimulation``: the rollout driver.
* ``femflow_amd.engine.Engine``: the particle-resident batched engine.
"""
__version__ = "0.1.0"
