"""MI355X (gfx950) implementation of the larnd-sim-jax LUT simulation chain.

Entry points mirror the reference: ``sim.simulate_wfs`` / ``sim.simulate_stochastic``
(src/larndsim/sim_jax.py), ``strategies.HipLUTSimulation`` (optimize/strategies.py),
``consts`` (src/larndsim/consts_jax.py).  Compute runs in the HIP library ``liblsim.so``.
"""
from . import consts  # noqa: F401

__all__ = ["consts"]
