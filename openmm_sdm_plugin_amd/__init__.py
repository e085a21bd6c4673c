"""MI355X-native SDM alchemical force + Langevin step (HIP, gfx950).

Public surface: LangevinIntegratorSDM / Context (the reference plugin's API),
SDMUtils (its restraint helper), SDMForce, ReplicaLadder, the synthetic-system
generators in `systems`, and the parameter-gradient autograd function in `adjoint`.
"""
from .integrator import Context, LangevinIntegratorSDM, ReplicaLadder, SDMForce  # noqa: F401
from .restraints import SDMUtils  # noqa: F401
from .systems import SDMParams, SDMSystem  # noqa: F401

__all__ = ["Context", "LangevinIntegratorSDM", "ReplicaLadder", "SDMForce", "SDMParams",
           "SDMSystem", "SDMUtils"]
