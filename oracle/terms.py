"""FP64 restatement of the DMS force-field terms and of position constraints.

TEST INFRASTRUCTURE ONLY: imported by tests/.  The product package never imports it.

Energies, written from the forces the reference's DMS reader builds
(example/desmonddmsfile75.py):
  periodic torsion     k (1 + cos(n phi - phase))   PeriodicTorsionForce (:666-682)
  harmonic improper    k (phi - phi0)^2             CustomTorsionForce (:684-717)
  flat-bottom posre    hk (d - tol)^2 for d >= tol, d the minimum-image distance to x0
                       (periodicdistance, :986-990)
  harmonic posre       sum_d hk_d (x_d - x0_d)^2    (:936)
  1-4 exception        4 eps ((s/r)^12 - (s/r)^6) + k_e qq / r, no cutoff, no image
                       (OpenMM NonbondedForce exceptions, added at :818-838)
with phi the dihedral in OpenMM's (IUPAC) sign convention.  Forces are central
differences of these energies.

Constraints: the positions x' = x_u + M^-1 sum_c lambda_c (r_c(x_old) on atom i, minus on
atom j) that satisfy |x'_i - x'_j| = d_c for every constraint c, found by Newton's method
on the multipliers lambda (a different solver from the kernel's SHAKE sweeps; both solve
the same equations, whose solution near x_u is unique).
"""
from __future__ import annotations

import math

import numpy as np

KE = 138.935456  # kJ nm / (mol e^2)


def dihedral(p1, p2, p3, p4):
    b1, b2, b3 = p2 - p1, p3 - p2, p4 - p3
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    return math.atan2(float(np.linalg.norm(b2) * (b1 @ n2)), float(n1 @ n2))


def _min_image(d, box):
    """Minimum image; no image for a non-periodic system (box None or zeros)."""
    if box is None or not np.any(np.asarray(box)):
        return d
    return d - box * np.round(d / box)


def _terms(ff, box):
    """(group, atoms, energy of the term's atoms' positions) for every term."""
    for (a, b, c, d), n, ph, k in zip(ff.torsions, ff.torsion_n, ff.torsion_phase, ff.torsion_k):
        yield 1, (a, b, c, d), (lambda p, n=n, ph=ph, k=k:
                                k * (1.0 + math.cos(n * dihedral(*p) - ph)))
    for (a, b, c, d), p0, k in zip(ff.impropers, ff.improper_t0, ff.improper_k):
        yield 1, (a, b, c, d), (lambda p, p0=p0, k=k: k * (dihedral(*p) - p0) ** 2)
    for a, x0, hk, tol in zip(ff.posre_fb_atoms, ff.posre_fb_x0, ff.posre_fb_hk,
                              ff.posre_fb_tol):
        def fb(p, x0=x0, hk=hk, tol=tol):
            dist = float(np.linalg.norm(_min_image(p[0] - x0, box)))
            return hk * (dist - tol) ** 2 if dist >= tol else 0.0
        yield 1, (a,), fb
    for a, x0, hk in zip(ff.posre_h_atoms, ff.posre_h_x0, ff.posre_h_hk):
        yield 1, (a,), (lambda p, x0=x0, hk=hk: float(np.sum(hk * (p[0] - x0) ** 2)))
    for (i, j), qq, sg, eps in zip(ff.exception_pairs, ff.exception_qq, ff.exception_sigma,
                                   ff.exception_epsilon):
        def ex(p, qq=qq, sg=sg, eps=eps):
            r = float(np.linalg.norm(p[0] - p[1]))
            sr6 = (sg / r) ** 6
            return 4 * eps * sr6 * (sr6 - 1.0) + KE * qq / r
        yield 2, (i, j), ex


def energy(x, ff, box):
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    e = [0.0, 0.0, 0.0]
    for grp, atoms, fn in _terms(ff, box):
        e[grp] += fn([x[a] for a in atoms])
    return e[1], e[2]


def energy_forces(x, ff, box, h=1e-6):
    """(E1, E2, F1, F2): group-1 and group-2 energies and forces, each term's force by
    central differences over its own atoms."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    e = [0.0, 0.0, 0.0]
    f = [None, np.zeros_like(x), np.zeros_like(x)]
    for grp, atoms, fn in _terms(ff, box):
        p = [x[a].copy() for a in atoms]
        e[grp] += fn(p)
        for m, a in enumerate(atoms):
            for d in range(3):
                keep = p[m][d]
                p[m][d] = keep + h
                ep = fn(p)
                p[m][d] = keep - h
                em = fn(p)
                p[m][d] = keep
                f[grp][a, d] -= (ep - em) / (2 * h)
    return e[1], e[2], f[1], f[2]


def _constrain_cluster(x_old, x_unc, im, pairs, d2, tol, max_iter):
    ro = x_old[pairs[:, 0]] - x_old[pairs[:, 1]]  # (c, 3)
    nc = pairs.shape[0]
    lam = np.zeros(nc)

    def positions(lam):
        x = x_unc.copy()
        np.add.at(x, pairs[:, 0], (lam * im[pairs[:, 0]])[:, None] * ro)
        np.add.at(x, pairs[:, 1], -(lam * im[pairs[:, 1]])[:, None] * ro)
        return x
    for _ in range(max_iter):
        x = positions(lam)
        rn = x[pairs[:, 0]] - x[pairs[:, 1]]
        g = (rn * rn).sum(1) - d2
        if np.max(np.abs(g) / d2) < tol:
            return x
        # dg_c/dlambda_k = 2 rn_c . d(rn_c)/dlambda_k
        J = np.zeros((nc, nc))
        for k in range(nc):
            i, j = pairs[k]
            for c in range(nc):
                a, b = pairs[c]
                coef = (im[i] if a == i else 0.0) - (im[i] if b == i else 0.0) \
                    - (im[j] if a == j else 0.0) + (im[j] if b == j else 0.0)
                J[c, k] = 2.0 * coef * (rn[c] @ ro[k])
        lam = lam - np.linalg.solve(J, g)
    raise RuntimeError("constraint Newton iteration did not converge")


def constrain(x_old, x_unc, mass, pairs, lengths, tol=1e-14, max_iter=50):
    """Constrained positions (see the module docstring), cluster by cluster (connected
    atoms of the constraint graph); raises if Newton fails."""
    x_old = np.asarray(x_old, dtype=np.float64).reshape(-1, 3)
    x = np.array(x_unc, dtype=np.float64).reshape(-1, 3)
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    d2 = np.asarray(lengths, dtype=np.float64) ** 2
    im = 1.0 / np.asarray(mass, dtype=np.float64)
    parent = {}

    def find(a):
        while parent.setdefault(a, a) != a:
            a = parent[a]
        return a
    for i, j in pairs:
        parent[find(int(i))] = find(int(j))
    groups = {}
    for k, (i, _j) in enumerate(pairs):
        groups.setdefault(find(int(i)), []).append(k)
    for ks in groups.values():
        atoms = sorted(set(pairs[ks].reshape(-1).tolist()))
        loc = {a: n for n, a in enumerate(atoms)}
        lp = np.array([[loc[int(i)], loc[int(j)]] for i, j in pairs[ks]])
        x[atoms] = _constrain_cluster(x_old[atoms], x[atoms], im[atoms], lp, d2[ks], tol,
                                      max_iter)
    return x


def project_velocities(x, v, mass, pairs):
    """Velocities with the component along every constraint removed: the kinetic energy of
    ReferenceSDMKernels.cpp:108-135 is summed after applyToVelocities(positions, v, 1/m),
    which makes the relative velocity of each constrained pair perpendicular to its bond.
    Here solved exactly, cluster by cluster, as the mass-weighted projection
        v' = v - M^-1 A^T lam,   (A M^-1 A^T) lam = A v,
    with A the constraint rows (r_c on atom i, -r_c on atom j, r_c = x_i - x_j at the
    current positions).  Returns v' (the input is not changed)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    v = np.array(v, dtype=np.float64).reshape(-1, 3)
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    im = 1.0 / np.asarray(mass, dtype=np.float64)
    parent = {}

    def find(a):
        while parent.setdefault(a, a) != a:
            a = parent[a]
        return a
    for i, j in pairs:
        parent[find(int(i))] = find(int(j))
    groups = {}
    for k, (i, _j) in enumerate(pairs):
        groups.setdefault(find(int(i)), []).append(k)
    for ks in groups.values():
        atoms = sorted(set(pairs[ks].reshape(-1).tolist()))
        loc = {a: n for n, a in enumerate(atoms)}
        na, nc = len(atoms), len(ks)
        A = np.zeros((nc, 3 * na))
        for row, k in enumerate(ks):
            i, j = int(pairs[k, 0]), int(pairs[k, 1])
            rc = x[i] - x[j]
            A[row, 3 * loc[i]:3 * loc[i] + 3] = rc
            A[row, 3 * loc[j]:3 * loc[j] + 3] = -rc
        minv = np.repeat(im[atoms], 3)
        vc = v[atoms].reshape(-1)
        lam = np.linalg.solve((A * minv) @ A.T, A @ vc)
        v[atoms] = (vc - minv * (A.T @ lam)).reshape(-1, 3)
    return v


def kinetic_energy(x, v, mass, pairs=None):
    """0.5 sum m v^2 over atoms with mass, after project_velocities when there are
    constraints (ReferenceSDMKernels.cpp:108-135)."""
    m = np.asarray(mass, dtype=np.float64)
    if pairs is not None and len(pairs):
        v = project_velocities(x, v, m, pairs)
    v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    keep = m > 0
    return 0.5 * float(np.sum(m[keep, None] * v[keep] ** 2))
