"""The oracle of the DMS force-field terms and of constraints (oracle/terms.py), pinned
by closed forms: a torsion at a known dihedral, the LJ minimum and Coulomb law of an
exception, the flat-bottom restraint across the periodic boundary, the diatomic
constraint (a quadratic in the multiplier), and the conservation of momentum."""
import math
import types

import numpy as np
import pytest

from oracle import terms as OT


def _ff(**kw):
    base = dict(torsions=np.zeros((0, 4), int), torsion_n=[], torsion_phase=[], torsion_k=[],
                impropers=np.zeros((0, 4), int), improper_t0=[], improper_k=[],
                posre_fb_atoms=[], posre_fb_x0=np.zeros((0, 3)), posre_fb_hk=[],
                posre_fb_tol=[], posre_h_atoms=[], posre_h_x0=np.zeros((0, 3)),
                posre_h_hk=np.zeros((0, 3)), exception_pairs=np.zeros((0, 2), int),
                exception_qq=[], exception_sigma=[], exception_epsilon=[])
    base.update(kw)
    return types.SimpleNamespace(**base)


BOX = np.array([3.0, 3.0, 3.0])
# a dihedral of +90 degrees (IUPAC): looking from atom 1 to atom 2, atom 0 at +x, 3 at +y
X90 = np.array([[1.0, 0, 0], [0, 0, 0], [0, 0, 1.0], [0, 1.0, 1.0]])


@pytest.mark.parametrize("n,phase", [(1, 0.0), (2, math.pi), (3, 0.3)])
def test_torsion_closed_form(n, phase):
    ff = _ff(torsions=np.array([[0, 1, 2, 3]]), torsion_n=[n], torsion_phase=[phase],
             torsion_k=[2.5])
    e1, e2 = OT.energy(X90, ff, BOX)
    assert e1 == pytest.approx(2.5 * (1 + math.cos(n * math.pi / 2 - phase)), rel=1e-13)
    assert e2 == 0.0


def test_improper_and_force_direction():
    ff = _ff(impropers=np.array([[0, 1, 2, 3]]), improper_t0=[1.2], improper_k=[7.0])
    e1, _ = OT.energy(X90, ff, BOX)
    assert e1 == pytest.approx(7.0 * (math.pi / 2 - 1.2) ** 2, rel=1e-13)
    _, _, f1, _ = OT.energy_forces(X90, ff, BOX)
    e_after, _ = OT.energy(X90 + 1e-5 * f1, ff, BOX)
    assert e_after < e1
    assert np.abs(f1.sum(0)).max() < 1e-6  # no net force


def test_positional_restraints_periodic():
    x = np.array([[0.05, 1.0, 1.0]])
    # x0 across the boundary: the minimum-image distance is 0.15, not 2.85
    ff = _ff(posre_fb_atoms=[0], posre_fb_x0=np.array([[2.9, 1.0, 1.0]]), posre_fb_hk=[40.0],
             posre_fb_tol=[0.1])
    assert OT.energy(x, ff, BOX)[0] == pytest.approx(40.0 * 0.05 ** 2, rel=1e-10)
    ff.posre_fb_tol = [0.2]
    assert OT.energy(x, ff, BOX)[0] == 0.0
    ff = _ff(posre_h_atoms=[0], posre_h_x0=np.array([[0.0, 1.2, 1.0]]),
             posre_h_hk=np.array([[10.0, 20.0, 30.0]]))
    assert OT.energy(x, ff, BOX)[0] == pytest.approx(10 * 0.05 ** 2 + 20 * 0.2 ** 2, rel=1e-12)


def test_exception_lj_minimum_and_coulomb():
    s, eps = 0.3, 0.5
    rmin = 2 ** (1 / 6) * s
    x = np.array([[0.0, 0, 0], [rmin, 0, 0]])
    ff = _ff(exception_pairs=np.array([[0, 1]]), exception_qq=[0.0], exception_sigma=[s],
             exception_epsilon=[eps])
    assert OT.energy(x, ff, BOX) == (0.0, pytest.approx(-eps, rel=1e-13))
    _, _, _, f2 = OT.energy_forces(x, ff, BOX)
    assert np.abs(f2).max() < 1e-6  # the LJ minimum
    ff.exception_qq, ff.exception_epsilon = [0.25], [0.0]
    x[1, 0] = 1.7  # beyond any cutoff: exceptions have none
    assert OT.energy(x, ff, BOX)[1] == pytest.approx(OT.KE * 0.25 / 1.7, rel=1e-13)


def test_constraint_diatomic_quadratic():
    rng = np.random.default_rng(2)
    x_old = np.array([[0.0, 0, 0], [0.1, 0.0, 0.0]])
    x_unc = x_old + rng.normal(scale=0.01, size=(2, 3))
    m = np.array([16.0, 1.0])
    xc = OT.constrain(x_old, x_unc, m, [[0, 1]], [0.1])
    # x'_0 = x_u0 + l r_o/m0, x'_1 = x_u1 - l r_o/m1 with |x'_0 - x'_1| = d: a quadratic in l
    ro, ru, mu = x_old[0] - x_old[1], x_unc[0] - x_unc[1], 1 / m[0] + 1 / m[1]
    a, b, c = mu * mu * (ro @ ro), 2 * mu * (ru @ ro), ru @ ru - 0.01
    roots = np.roots([a, b, c])
    lam = roots[np.argmin(np.abs(roots))].real
    assert np.allclose(xc[0], x_unc[0] + lam * ro / m[0], atol=1e-15)
    assert np.allclose(xc[1], x_unc[1] - lam * ro / m[1], atol=1e-15)


def test_constraint_water_momentum_and_lengths():
    rng = np.random.default_rng(4)
    x_old = np.array([[0.0, 0, 0], [0.09572, 0, 0], [-0.024, 0.0927, 0]])
    d = [0.09572, 0.09572, float(np.linalg.norm(x_old[1] - x_old[2]))]
    x_old[2] *= 0.09572 / np.linalg.norm(x_old[2])
    d[2] = float(np.linalg.norm(x_old[1] - x_old[2]))
    m = np.array([15.999, 1.008, 1.008])
    x_unc = x_old + rng.normal(scale=0.003, size=(3, 3))
    pairs = [[0, 1], [0, 2], [1, 2]]
    xc = OT.constrain(x_old, x_unc, m, pairs, d)
    for (i, j), L in zip(pairs, d):
        assert np.linalg.norm(xc[i] - xc[j]) == pytest.approx(L, rel=1e-13)
    # internal constraint forces: the mass-weighted displacement (momentum) is unchanged
    assert np.allclose((m[:, None] * xc).sum(0), (m[:, None] * x_unc).sum(0), atol=1e-15)
