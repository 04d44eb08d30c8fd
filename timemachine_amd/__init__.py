"""MI355X-native (gfx950) MD hot path with the operator API of timemachine.

Subpackages:
  lib          drop-in ``custom_ops`` / ``potentials`` / integrator factories
  testsystems  seeded synthetic workloads (water box, ligand in water, jobs)
"""

__version__ = "0.1.0"
