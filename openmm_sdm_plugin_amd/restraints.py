"""Restraint forces of the reference's SDMUtils (python/SDMUtils.py), for SDMSystem.

`SDMUtils(system)` mirrors the reference class: `addRestraintForce` (SDMUtils.py:32-158)
and `addAlignmentForce` (SDMUtils.py:161-257) record the restraint on the system
(`system.restraints`), and a context created from the system evaluates them on the GPU
as force-group-1 terms (at x, into the bonded forces and Ub).  The reference takes
OpenMM unit quantities; here every value is a plain number in OpenMM's internal units:
nm, rad, kJ/mol (force constants kJ/mol/nm^2 and kJ/mol/rad^2), as the reference
converts them before building its forces.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Sequence

import numpy as np

_DEG = math.pi / 180.0


@dataclasses.dataclass
class CentroidRestraint:
    """One addRestraintForce term: a CustomCentroidBondForce of 2 (or 8) groups."""

    lig_atoms: np.ndarray
    rcpt_atoms: np.ndarray
    kfcm: float
    tolcm: float
    offset: np.ndarray
    do_angles: bool = False
    rcpt_ref: Sequence[int] = (0, 0, 0)
    lig_ref: Sequence[int] = (0, 0, 0)
    kf: Sequence[float] = (0.0, 0.0, 0.0)  # theta, phi1, phi2
    lo: Sequence[float] = (0.0, 0.0, 0.0)  # window: center - tolerance
    hi: Sequence[float] = (0.0, 0.0, 0.0)  # window: center + tolerance
    control: float = 1.0  # SDMRestraintControlParameter
    lig_weights: Optional[np.ndarray] = None  # None: particle masses
    rcpt_weights: Optional[np.ndarray] = None


@dataclasses.dataclass
class AlignmentRestraint:
    """One addAlignmentForce term set (displacement, theta and psi restraints)."""

    liga_ref: Sequence[int]
    ligb_ref: Sequence[int]
    kfdispl: float
    ktheta: float
    kpsi: float
    offset: np.ndarray


class SDMUtils:
    """Mirror of SDMUtils (SDMUtils.py:3-257) over an SDMSystem."""

    def __init__(self, system):
        self.system = system
        self.RestraintControlParameterName = "SDMRestraintControlParameter"
        self.force = None
        self.LinearMethod = 0
        self.QuadraticMethod = 1
        self.ILogisticMethod = 2
        self.NoSoftCoreMethod = 0
        self.TanhSoftCoreMethod = 1
        self.RationalSoftCoreMethod = 2
        if getattr(system, "restraints", None) is None:
            system.restraints = []

    def getControlParameterName(self):
        return self.RestraintControlParameterName

    def addRestraintForce(self, lig_cm_particles=None, rcpt_cm_particles=None, kfcm=0.0,
                          tolcm=0.0, lig_ref_particles=None, rcpt_ref_particles=None,
                          angle_center=0.0, kfangle=0.0, angletol=10.0 * _DEG,
                          dihedral1center=0.0, kfdihedral1=0.0, dihedral1tol=10.0 * _DEG,
                          dihedral2center=0.0, kfdihedral2=0.0, dihedral2tol=10.0 * _DEG,
                          offset=(0.0, 0.0, 0.0)):
        """Centroid distance restraint, plus the angle and two dihedral restraints when
        both reference-atom lists are given (SDMUtils.py:32-158).  Does nothing when
        either centroid group is missing or empty, as the reference."""
        if not (lig_cm_particles is not None and rcpt_cm_particles is not None):
            return
        if not (len(lig_cm_particles) > 0 and len(rcpt_cm_particles) > 0):
            return
        do_angles = False
        if lig_ref_particles is not None and rcpt_ref_particles is not None:
            if not (len(lig_ref_particles) == 3 and len(rcpt_ref_particles) == 3):
                raise ValueError("Invalid lists of reference atoms")
            do_angles = True
        r = CentroidRestraint(
            lig_atoms=np.asarray(lig_cm_particles, dtype=np.int32),
            rcpt_atoms=np.asarray(rcpt_cm_particles, dtype=np.int32),
            kfcm=float(kfcm), tolcm=float(tolcm),
            offset=np.asarray(offset, dtype=np.float64).reshape(3))
        if do_angles:
            r.do_angles = True
            r.rcpt_ref = tuple(int(a) for a in rcpt_ref_particles)
            r.lig_ref = tuple(int(a) for a in lig_ref_particles)
            r.kf = (float(kfangle), float(kfdihedral1), float(kfdihedral2))
            r.lo = (angle_center - angletol, dihedral1center - dihedral1tol,
                    dihedral2center - dihedral2tol)
            r.hi = (angle_center + angletol, dihedral1center + dihedral1tol,
                    dihedral2center + dihedral2tol)
        self.system.restraints.append(r)
        self.force = r

    def addAlignmentForce(self, liga_ref_particles=None, ligb_ref_particles=None, kfdispl=0.0,
                          ktheta=0.0, kpsi=0.0, offset=(0.0, 0.0, 0.0)):
        """Keeps ligand b aligned with ligand a displaced by offset (SDMUtils.py:161-257):
        reference atoms 0 near each other, the 0->1 axes parallel (theta) and the
        0->2 directions about that axis parallel (psi, symmetrized)."""
        if not (liga_ref_particles is not None and ligb_ref_particles is not None and
                len(liga_ref_particles) == 3 and len(ligb_ref_particles) == 3):
            raise ValueError("Invalid lists of reference atoms")
        self.system.restraints.append(AlignmentRestraint(
            liga_ref=tuple(int(a) for a in liga_ref_particles),
            ligb_ref=tuple(int(a) for a in ligb_ref_particles),
            kfdispl=float(kfdispl), ktheta=float(ktheta), kpsi=float(kpsi),
            offset=np.asarray(offset, dtype=np.float64).reshape(3)))
