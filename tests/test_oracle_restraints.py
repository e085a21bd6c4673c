"""The restraint oracle (oracle/restraints.py) pinned by closed forms and invariances.

Each check is chosen so that a plausible slip in the restatement fails it: a centre that
ignores the masses, a window edge on the wrong side, a dihedral with the opposite
handedness, a missing wrap at +-pi, the psi terms' factor of 4, a force with the wrong
sign.  The references' expressions are in SDMUtils.py:61-81 and :176-255."""
import math

import numpy as np
import pytest

from openmm_sdm_plugin_amd import restraints as R
from oracle import restraints as OR


def _centroid(lig, rcpt, **kw):
    base = dict(lig_atoms=np.array(lig), rcpt_atoms=np.array(rcpt), kfcm=0.0, tolcm=0.0,
                offset=np.zeros(3))
    base.update(kw)
    return R.CentroidRestraint(**base)


def test_distance_term_closed_form_and_mass_weighting():
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [3.0, 0.5, 0]])
    mass = np.array([1.0, 3.0, 2.0])
    r = _centroid([0, 1], [2], kfcm=400.0, tolcm=0.5)
    c1 = np.array([0.75, 0, 0])  # mass-weighted, not the midpoint 0.5
    d = float(np.linalg.norm(c1 - x[2]))
    assert OR.energy(x, [r], mass) == pytest.approx(200.0 * (d - 0.5) ** 2, rel=1e-14)
    # explicit equal weights give the geometric centre instead
    r.lig_weights = np.array([1.0, 1.0])
    d2 = float(np.linalg.norm(np.array([0.5, 0, 0]) - x[2]))
    assert OR.energy(x, [r], mass) == pytest.approx(200.0 * (d2 - 0.5) ** 2, rel=1e-14)
    # inside the tolerance: zero; the offset is subtracted from the ligand centre
    r.tolcm = 5.0
    assert OR.energy(x, [r], mass) == 0.0
    r2 = _centroid([0], [2], kfcm=10.0, offset=x[0] - x[2])  # g1 - offset = g2
    assert OR.energy(x, [r2], mass) == 0.0
    r2.offset = x[2] - x[0]  # the wrong sign doubles the separation: (10/2) * (2 d)^2
    assert OR.energy(x, [r2], mass) == pytest.approx(5.0 * 4 * 9.25, rel=1e-14)


def _angle_points(theta):
    return np.array([1.0, 0, 0]), np.zeros(3), np.array([math.cos(theta), math.sin(theta), 0])


def test_angle_window_and_wrap():
    p1, p2, p3 = _angle_points(2.0)
    assert OR.angle(p1, p2, p3) == pytest.approx(2.0, abs=1e-15)
    # window [1.3, 1.7]: 0.3 above its upper edge
    assert OR.flat_bottom(2.0, 50.0, 1.3, 1.7, math.pi) == pytest.approx(25.0 * 0.09, rel=1e-12)
    assert OR.flat_bottom(1.5, 50.0, 1.3, 1.7, math.pi) == 0.0
    assert OR.flat_bottom(1.0, 50.0, 1.3, 1.7, math.pi) == pytest.approx(25.0 * 0.09, rel=1e-12)
    # a window across 0 ([-0.1, 0.1]) holds small angles; 0.3 is 0.2 past its edge
    assert OR.flat_bottom(0.05, 8.0, -0.1, 0.1, math.pi) == 0.0
    assert OR.flat_bottom(0.3, 8.0, -0.1, 0.1, math.pi) == pytest.approx(4.0 * 0.04, rel=1e-12)


def test_dihedral_iupac_sign_and_wrap():
    a, b, c = np.array([1.0, 0, 0]), np.zeros(3), np.array([0, 0, 1.0])
    d = np.array([0, 1.0, 1.0])
    # looking from b to c, a at +x and d at +y: clockwise, +90 degrees (IUPAC)
    assert OR.dihedral(a, b, c, d) == pytest.approx(math.pi / 2, abs=1e-15)
    d2 = np.array([0, -1.0, 1.0])
    assert OR.dihedral(a, b, c, d2) == pytest.approx(-math.pi / 2, abs=1e-15)
    assert abs(OR.dihedral(a, b, c, np.array([-1.0, 0, 1.0]))) == pytest.approx(math.pi)
    deg = math.pi / 180
    # window 150..190 degrees contains -175 (= 185) degrees through the wrap
    assert OR.flat_bottom(-175 * deg, 3.0, 150 * deg, 190 * deg, 2 * math.pi) == 0.0
    # 90 degrees against a window 50..70 degrees: 20 degrees past its edge
    assert OR.flat_bottom(90 * deg, 3.0, 50 * deg, 70 * deg, 2 * math.pi) == pytest.approx(
        1.5 * (20 * deg) ** 2, rel=1e-12)


def _align_system(twist):
    """b: axis along z from the origin, b2 on +x; a: the same, displaced by (2,0,0), a2
    rotated by `twist` about z."""
    b = np.array([[0.0, 0, 0], [0, 0, 1.0], [1.0, 0, 0]])
    off = np.array([2.0, 0, 0])
    a = b + off
    a[2] = off + np.array([math.cos(twist), math.sin(twist), 0])
    return np.vstack([a, b]), off


def test_alignment_terms_closed_forms():
    x, off = _align_system(0.0)
    r = R.AlignmentRestraint(liga_ref=(0, 1, 2), ligb_ref=(3, 4, 5), kfdispl=0.0, ktheta=0.0,
                             kpsi=0.0, offset=-off)
    mass = np.ones(6)
    # aligned: every term is zero
    r.kfdispl, r.ktheta, r.kpsi = 10.0, 20.0, 30.0
    assert OR.energy(x, [r], mass) == pytest.approx(0.0, abs=1e-12)
    # a 90 degree twist about the common axis: each psi bond (kpsi/2)/2 * (1 - 0)
    x, _ = _align_system(math.pi / 2)
    r.kfdispl, r.ktheta = 0.0, 0.0
    assert OR.energy(x, [r], mass) == pytest.approx(30.0 / 2, rel=1e-12)
    # antiparallel axes: ktheta
    x, _ = _align_system(0.0)
    x[1] = x[0] - np.array([0, 0, 1.0])
    r.kpsi, r.ktheta = 0.0, 20.0
    assert OR.energy(x, [r], mass) == pytest.approx(20.0, rel=1e-12)
    # displacement: (k/2)|b0 - offset - a0|^2 with b0 moved by 0.3
    x, _ = _align_system(0.0)
    x[3] += np.array([0.3, 0, 0])
    r.ktheta, r.kfdispl = 0.0, 10.0
    assert OR.energy(x, [r], mass) == pytest.approx(5.0 * 0.09, rel=1e-12)


def _random_set(rng):
    x = rng.normal(scale=0.6, size=(14, 3))
    mass = rng.uniform(1, 16, size=14)
    c = _centroid(list(range(0, 5)), list(range(5, 11)), kfcm=300.0, tolcm=0.05,
                  offset=np.zeros(3), do_angles=True, rcpt_ref=(5, 6, 7), lig_ref=(0, 1, 2),
                  kf=(40.0, 30.0, 20.0))
    t = OR.angle(x[5], x[0], x[1])
    p1 = OR.dihedral(x[6], x[5], x[0], x[1])
    p2 = OR.dihedral(x[5], x[0], x[1], x[2])
    # windows that leave every term active
    c.lo = (t + 0.3, p1 + 0.4, p2 - 0.9)
    c.hi = (t + 0.5, p1 + 0.6, p2 - 0.5)
    a = R.AlignmentRestraint(liga_ref=(0, 1, 2), ligb_ref=(11, 12, 13), kfdispl=50.0,
                             ktheta=25.0, kpsi=35.0, offset=np.zeros(3))
    return x, mass, [c, a]


def test_invariances_and_force_balance():
    rng = np.random.default_rng(11)
    x, mass, rs = _random_set(rng)
    e = OR.energy(x, rs, mass)
    assert e > 1.0
    assert OR.energy(x + np.array([0.7, -1.1, 2.3]), rs, mass) == pytest.approx(e, rel=1e-12)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]  # a proper rotation keeps the dihedral handedness
    assert OR.energy(x @ q.T, rs, mass) == pytest.approx(e, rel=1e-10)
    _, f = OR.energy_forces(x, rs, mass)
    # translation invariance: no net force; rotation invariance: no net torque
    assert np.abs(f.sum(0)).max() < 1e-6 * np.abs(f).max()
    assert np.abs(np.cross(x, f).sum(0)).max() < 1e-6 * np.abs(f).max()


def test_force_is_minus_gradient_sign():
    rng = np.random.default_rng(5)
    x, mass, rs = _random_set(rng)
    e, f = OR.energy_forces(x, rs, mass)
    # a small step along the force lowers the energy by about |F|^2 * step
    s = 1e-6
    e2 = OR.energy(x + s * f, rs, mass)
    assert e2 < e
    assert (e - e2) == pytest.approx(s * float((f * f).sum()), rel=1e-3)
