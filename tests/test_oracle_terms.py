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


def _water_cluster(rng, n=3):
    """n rigid waters (O-H, O-H, H-H constraints) at random orientations, 0.4 nm apart."""
    x, pairs, m = [], [], []
    canon = np.array([[0.0, 0, 0], [0.09572, 0, 0], [-0.02399, 0.09266, 0]])
    for k in range(n):
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        x.extend(list(canon @ q.T + np.array([0.4 * k, 0, 0])))
        pairs += [[3 * k, 3 * k + 1], [3 * k, 3 * k + 2], [3 * k + 1, 3 * k + 2]]
        m += [15.9994, 1.008, 1.008]
    return np.array(x), np.array(pairs), np.array(m)


def test_velocity_projection_removes_bond_components():
    """After the projection every constrained pair's relative velocity is perpendicular
    to its bond, the momentum of every cluster is unchanged (the constraint impulses are
    internal), and a second projection changes nothing."""
    rng = np.random.default_rng(7)
    x, pairs, m = _water_cluster(rng)
    v = rng.normal(size=x.shape)
    vp = OT.project_velocities(x, v, m, pairs)
    for i, j in pairs:
        assert abs((x[i] - x[j]) @ (vp[i] - vp[j])) < 1e-13
    for k in range(3):
        sl = slice(3 * k, 3 * k + 3)
        assert np.allclose((m[sl, None] * vp[sl]).sum(0), (m[sl, None] * v[sl]).sum(0),
                           atol=1e-13)
    assert np.allclose(OT.project_velocities(x, vp, m, pairs), vp, atol=1e-13)


def test_velocity_projection_diatomic_closed_form():
    """Diatomic: only the relative velocity along the bond is removed, split between the
    atoms in inverse proportion to their masses (a mass-for-inverse-mass slip fails)."""
    x = np.array([[0.0, 0, 0], [0.06, 0.08, 0.0]])
    v = np.array([[1.0, -2.0, 0.5], [0.3, 0.4, -1.0]])
    m = np.array([12.0, 1.0])
    r = x[0] - x[1]
    u = r / np.linalg.norm(r)
    w = (v[0] - v[1]) @ u  # relative speed along the bond
    mu0, mu1 = (1 / m[0]) / (1 / m[0] + 1 / m[1]), (1 / m[1]) / (1 / m[0] + 1 / m[1])
    want = np.array([v[0] - mu0 * w * u, v[1] + mu1 * w * u])
    assert np.allclose(OT.project_velocities(x, v, m, [[0, 1]]), want, atol=1e-15)


def test_kinetic_energy_rigid_rotor():
    """A rigid diatomic spinning about its centre with a bond-stretching velocity added:
    the projected kinetic energy is the rotational one, 0.5 I w^2, and without constraints
    the stretching energy is kept."""
    m = np.array([3.0, 1.0])
    d = 0.2
    x = np.array([[-0.25 * d, 0, 0], [0.75 * d, 0, 0]])  # centre of mass at the origin
    w = 5.0
    v_rot = np.array([[0, w * x[0, 0], 0], [0, w * x[1, 0], 0]])
    v = v_rot + np.array([[-0.7, 0, 0], [2.1, 0, 0]])  # zero-momentum stretch
    inertia = float(np.sum(m * x[:, 0] ** 2))
    assert OT.kinetic_energy(x, v, m, [[0, 1]]) == pytest.approx(0.5 * inertia * w * w,
                                                                 rel=1e-13)
    assert OT.kinetic_energy(x, v, m) == pytest.approx(
        0.5 * float(np.sum(m[:, None] * v ** 2)), rel=1e-15)
