"""FP64 restatement of virtual sites: their positions, and forces moved to their parents.

TEST INFRASTRUCTURE ONLY: imported by tests/.  The product package never imports it.

The reference places virtual sites after every step (ReferenceStochasticDynamicsSDM.cpp:264,
OpenCLSDMKernels.cpp:376-377) through OpenMM's VirtualSite classes.  OpenMM is not part of
the reference's sources; its README names OpenMM 7.2.2 and 7.3.1.  The DMS reader defines
the sites (example/desmonddmsfile75.py:852-915):
  virtual_lc2   TwoParticleAverageSite(p1, p2, 1-c1, c1)           (:877)
                x = w1 x1 + w2 x2
  virtual_lc3   ThreeParticleAverageSite(p1, p2, p3, 1-c1-c2, c1, c2)   (:892)
                x = w1 x1 + w2 x2 + w3 x3
  virtual_out3  OutOfPlaneSite(p1, p2, p3, c1, c2, c3)             (:907)
                x = x1 + w12 r12 + w13 r13 + wc (r12 x r13),  r1k = xk - x1
with coordinates as stored (no periodic image).  A force f on a site belongs to the
energy written in the parents' coordinates, so parent k receives J_k^T f, J_k = dx/dx_k
(the chain rule; OpenMM distributes forces this way, and the site keeps none).  Here each
J_k is formed as a 3x3 matrix (for the out-of-plane site from the skew matrices of r12 and
r13) and applied transposed; tests/test_oracle_vsites.py pins positions by closed forms and
the distribution by central differences of energies of the site position.

Sites are given as (kind, atoms, weights): kind 0 (two-particle average), 1 (three-particle
average), 2 (out of plane); atoms (v, 4) = site, p1, p2, p3 (p3 unused for kind 0); weights
(v, 3) = (w1, w2, w3) or (w12, w13, wc).
"""
from __future__ import annotations

import numpy as np

AVERAGE2, AVERAGE3, OUT_OF_PLANE = 0, 1, 2


def _skew(a):
    """[a]x: [a]x b = a x b."""
    return np.array([[0.0, -a[2], a[1]], [a[2], 0.0, -a[0]], [-a[1], a[0], 0.0]])


def jacobians(kind, x1, x2, x3, w):
    """d(site)/d(parent k) for the parents actually used, as 3x3 matrices."""
    eye = np.eye(3)
    if kind == AVERAGE2:
        return [w[0] * eye, w[1] * eye]
    if kind == AVERAGE3:
        return [w[0] * eye, w[1] * eye, w[2] * eye]
    r12, r13 = x2 - x1, x3 - x1
    # d(r12 x r13)/d r12 = -[r13]x, d(r12 x r13)/d r13 = [r12]x
    j2 = w[0] * eye - w[2] * _skew(r13)
    j3 = w[1] * eye + w[2] * _skew(r12)
    return [eye - j2 - j3, j2, j3]


def site_position(kind, x1, x2, x3, w):
    if kind == AVERAGE2:
        return w[0] * x1 + w[1] * x2
    if kind == AVERAGE3:
        return w[0] * x1 + w[1] * x2 + w[2] * x3
    r12, r13 = x2 - x1, x3 - x1
    return x1 + w[0] * r12 + w[1] * r13 + w[2] * np.cross(r12, r13)


def _parents(vs, k, x):
    a = np.asarray(vs.atoms).reshape(-1, 4)[k]
    kind = int(np.asarray(vs.kind)[k])
    p = [int(a[1]), int(a[2])] + ([int(a[3])] if kind != AVERAGE2 else [])
    xs = [np.asarray(x[i], dtype=np.float64) for i in p]
    if kind == AVERAGE2:
        xs.append(np.zeros(3))
    return int(a[0]), kind, p, xs, np.asarray(vs.weights, dtype=np.float64).reshape(-1, 3)[k]


def positions(x, vs):
    """x (n,3) with every site placed from its parents."""
    x = np.array(x, dtype=np.float64).reshape(-1, 3)
    for k in range(len(vs.kind)):
        s, kind, _, xs, w = _parents(vs, k, x)
        x[s] = site_position(kind, xs[0], xs[1], xs[2], w)
    return x


def distribute(F, x, vs):
    """F (n,3) with every site's force moved to its parents (J_k^T f) and the site's set to
    zero; x the positions the forces were evaluated at."""
    F = np.array(F, dtype=np.float64).reshape(-1, 3)
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    for k in range(len(vs.kind)):
        s, kind, p, xs, w = _parents(vs, k, x)
        f = F[s].copy()
        F[s] = 0.0
        for i, J in zip(p, jacobians(kind, xs[0], xs[1], xs[2], w)):
            F[i] += J.T @ f
    return F
