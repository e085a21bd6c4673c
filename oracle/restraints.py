"""FP64 restatement of the SDMUtils restraint energies (python/SDMUtils.py:32-257).

TEST INFRASTRUCTURE ONLY: imported by tests/.  The product package never imports it.

Each energy is written term by term from the reference's Lepton expressions:
  addRestraintForce (SDMUtils.py:61-81, a CustomCentroidBondForce):
    C*(kfcm/2)*step(d12-tolcm)*(d12-tolcm)^2, d12 = |g1 - off - g2| (centres weighted by
    mass, the force's default); with reference atoms, for theta = angle(g3,g6,g7),
    phi1 = dihedral(g4,g3,g6,g7), phi2 = dihedral(g3,g6,g7,g8):
    C*(kf/2)*(step(dm)*max(0,db)^2 + step(-dm)*max(0,-da)^2), db = wrap(t - b),
    da = wrap(t - a), dm = wrap(t - (a+b)/2), wrap(x) = x - P*floor(x/P + 0.5), P = pi
    for theta and 2 pi for the dihedrals, step(0) = 1.
  addAlignmentForce (SDMUtils.py:176-255, three CustomCompoundBondForces):
    (kfdispl/2)*|b0 - off - a0|^2; (ktheta/2)*(1 - cos t) with cos t between b1-b0 and
    a1-a0; psi bonds [b0,b1,b2,a0,a2] and [a0,a1,a2,b0,b2] each with parameter kpsi/2 in
    (kpsi/2)*(1 - cos p), cos p between the components of x3-x1 and x5-x4 perpendicular
    to x2-x1.
angle() and dihedral() follow OpenMM's definitions (third-party; angle at the middle
point in [0, pi]; dihedral with the IUPAC sign, which OpenMM's reference bonded code
uses).  No periodic images (the reference's forces do not set periodic boundaries).
Forces are central differences of these energies (h = 1e-6 nm), so they do not share
the HIP kernel's analytic gradients.
"""
from __future__ import annotations

import math

import numpy as np


def _wrap(x, period):
    return x - period * math.floor(x / period + 0.5)


def _step(x):
    return 1.0 if x >= 0.0 else 0.0


def angle(p1, p2, p3):
    u, v = p1 - p2, p3 - p2
    return math.atan2(float(np.linalg.norm(np.cross(u, v))), float(u @ v))


def dihedral(p1, p2, p3, p4):
    b1, b2, b3 = p2 - p1, p3 - p2, p4 - p3
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    return math.atan2(float(np.linalg.norm(b2) * (b1 @ n2)), float(n1 @ n2))


def flat_bottom(t, kf, a, b, period):
    db, da = _wrap(t - b, period), _wrap(t - a, period)
    dm = _wrap(t - 0.5 * (a + b), period)
    return 0.5 * kf * (_step(dm) * max(0.0, db) ** 2 + _step(-dm) * max(0.0, -da) ** 2)


def centre(x, atoms, weights):
    w = np.asarray(weights, dtype=np.float64)
    return (w[:, None] * x[np.asarray(atoms)]).sum(0) / w.sum()


def centroid_energy(x, r, mass):
    """One CentroidRestraint (openmm_sdm_plugin_amd.restraints) at positions x (n,3)."""
    wl = mass[r.lig_atoms] if r.lig_weights is None else r.lig_weights
    wr = mass[r.rcpt_atoms] if r.rcpt_weights is None else r.rcpt_weights
    g1, g2 = centre(x, r.lig_atoms, wl), centre(x, r.rcpt_atoms, wr)
    d12 = float(np.linalg.norm(g1 - np.asarray(r.offset) - g2))
    e = r.control * (r.kfcm / 2) * _step(d12 - r.tolcm) * (d12 - r.tolcm) ** 2
    if r.do_angles:
        g3, g4, _g5 = (x[a] for a in r.rcpt_ref)
        g6, g7, g8 = (x[a] for a in r.lig_ref)
        t = angle(g3, g6, g7)
        p1 = dihedral(g4, g3, g6, g7)
        p2 = dihedral(g3, g6, g7, g8)
        e += r.control * flat_bottom(t, r.kf[0], r.lo[0], r.hi[0], math.pi)
        e += r.control * flat_bottom(p1, r.kf[1], r.lo[1], r.hi[1], 2 * math.pi)
        e += r.control * flat_bottom(p2, r.kf[2], r.lo[2], r.hi[2], 2 * math.pi)
    return e


def _cos_psi(x1, x2, x3, x4, x5):
    d1 = x2 - x1
    n1 = d1 / np.linalg.norm(d1)
    d0 = x3 - x1
    v = d0 - (d0 @ n1) * n1
    d3 = x5 - x4
    w = d3 - (d3 @ n1) * n1
    return float((v @ w) / (np.linalg.norm(v) * np.linalg.norm(w)))


def alignment_energy(x, r):
    a0, a1, a2 = (x[i] for i in r.liga_ref)
    b0, b1, b2 = (x[i] for i in r.ligb_ref)
    dd = b0 - np.asarray(r.offset) - a0
    e = 0.5 * r.kfdispl * float(dd @ dd)
    d1, d2 = b1 - b0, a1 - a0
    cost = float((d1 @ d2) / (np.linalg.norm(d1) * np.linalg.norm(d2)))
    e += 0.5 * r.ktheta * (1.0 - cost)
    kp = 0.5 * r.kpsi  # the per-bond parameter of both psi bonds
    e += 0.5 * kp * (1.0 - _cos_psi(b0, b1, b2, a0, a2))
    e += 0.5 * kp * (1.0 - _cos_psi(a0, a1, a2, b0, b2))
    return e


def energy(x, restraints, mass):
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    e = 0.0
    for r in restraints:
        e += centroid_energy(x, r, mass) if hasattr(r, "lig_atoms") else alignment_energy(x, r)
    return e


def atoms_of(restraints):
    out = set()
    for r in restraints:
        if hasattr(r, "lig_atoms"):
            out.update(int(a) for a in r.lig_atoms)
            out.update(int(a) for a in r.rcpt_atoms)
            if r.do_angles:
                out.update(int(a) for a in r.rcpt_ref)
                out.update(int(a) for a in r.lig_ref)
        else:
            out.update(int(a) for a in r.liga_ref)
            out.update(int(a) for a in r.ligb_ref)
    return sorted(out)


def energy_forces(x, restraints, mass, h=1e-6):
    """(E, F) with F = -dE/dx by central differences over the restraint atoms."""
    x = np.array(x, dtype=np.float64).reshape(-1, 3)
    e = energy(x, restraints, mass)
    f = np.zeros_like(x)
    for a in atoms_of(restraints):
        for d in range(3):
            keep = x[a, d]
            x[a, d] = keep + h
            ep = energy(x, restraints, mass)
            x[a, d] = keep - h
            em = energy(x, restraints, mass)
            x[a, d] = keep
            f[a, d] = -(ep - em) / (2 * h)
    return e, f
