"""The virtual-site oracle (oracle/vsites.py), pinned by closed forms (the TIP4P-style
charge site on the water bisector, the right-handed out-of-plane site), by central
differences (the distributed forces are minus the gradient of an energy of the site
position written in the parents' coordinates), and by conservation: the parents' forces
sum to the site's force and carry its torque."""
import types

import numpy as np
import pytest

from oracle import vsites as OV


def _vs(kind, atoms, weights):
    return types.SimpleNamespace(kind=np.array(kind, dtype=np.int32),
                                 atoms=np.array(atoms, dtype=np.int32).reshape(-1, 4),
                                 weights=np.array(weights, dtype=np.float64).reshape(-1, 3))


def test_water_charge_site_on_the_bisector():
    """x_M = (1-2a) x_O + a x_H1 + a x_H2 lies on the H-O-H bisector at
    2 a r_OH cos(theta / 2) from the oxygen."""
    r, th, a = 0.09572, np.radians(104.52), 0.1
    o = np.array([1.0, 1.0, 1.0])
    h1 = o + r * np.array([np.sin(th / 2), np.cos(th / 2), 0.0])
    h2 = o + r * np.array([-np.sin(th / 2), np.cos(th / 2), 0.0])
    x = np.array([o, h1, h2, np.zeros(3)])
    y = OV.positions(x, _vs([OV.AVERAGE3], [3, 0, 1, 2], [1 - 2 * a, a, a]))[3]
    d = y - o
    assert abs(d[0]) < 1e-15 and abs(d[2]) < 1e-15
    assert d[1] == pytest.approx(2 * a * r * np.cos(th / 2), rel=1e-13)


def test_out_of_plane_site_is_right_handed():
    """x1 at the origin, r12 = a e_x, r13 = b e_y: the site is (w12 a, w13 b, wc a b), on
    the +z side for wc > 0 (r12 x r13 = a b e_z)."""
    a, b = 0.3, 0.2
    x = np.array([[0, 0, 0], [a, 0, 0], [0, b, 0], [9, 9, 9]], dtype=float)
    y = OV.positions(x, _vs([OV.OUT_OF_PLANE], [3, 0, 1, 2], [0.25, -0.5, 2.0]))[3]
    assert np.allclose(y, [0.25 * a, -0.5 * b, 2.0 * a * b], rtol=0, atol=1e-15)


def test_two_particle_site_between_its_parents():
    x = np.array([[0.1, 0.2, 0.3], [1.1, -0.8, 0.3], [0, 0, 0]])
    y = OV.positions(x, _vs([OV.AVERAGE2], [2, 0, 1, -1], [0.75, 0.25, 0.0]))[2]
    # a quarter of the way from p1 to p2
    assert np.allclose(y, x[0] + 0.25 * (x[1] - x[0]), rtol=0, atol=1e-15)


def _energy(y):
    """an energy of the site position: a harmonic tether and a Coulomb-like term"""
    c, q = np.array([0.7, -0.2, 0.4]), np.array([-0.3, 0.5, 1.1])
    return 3.0 * np.sum((y - c) ** 2) + 2.0 / np.linalg.norm(y - q)


def _grad(y, h=1e-6):
    g = np.zeros(3)
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        g[d] = (_energy(y + e) - _energy(y - e)) / (2 * h)
    return g


@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("kind", [OV.AVERAGE2, OV.AVERAGE3, OV.OUT_OF_PLANE])
def test_distribution_is_minus_the_gradient_in_parent_coordinates(kind, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.5, 0.5, (4, 3))
    w = rng.uniform(-1.0, 1.0, 3)
    if kind == OV.AVERAGE2:
        w[1] = 1 - w[0]
    if kind == OV.AVERAGE3:
        w[2] = 1 - w[0] - w[1]
    vs = _vs([kind], [3, 0, 1, 2], w)
    xs = OV.positions(x, vs)
    F = np.zeros((4, 3))
    F[3] = -_grad(xs[3])
    Fd = OV.distribute(F, xs, vs)
    assert np.all(Fd[3] == 0.0)
    # the same energy as a function of the parents, differentiated numerically
    h = 1e-6
    for i in range(3):
        for d in range(3):
            xp, xm = xs.copy(), xs.copy()
            xp[i, d] += h
            xm[i, d] -= h
            ep = _energy(OV.positions(xp, vs)[3])
            em = _energy(OV.positions(xm, vs)[3])
            assert Fd[i, d] == pytest.approx(-(ep - em) / (2 * h), rel=1e-6, abs=1e-8), (i, d)


@pytest.mark.parametrize("kind", [OV.AVERAGE3, OV.OUT_OF_PLANE])
def test_distribution_keeps_force_and_torque(kind):
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, (4, 3))
    w = np.array([0.2, 0.3, 0.5]) if kind == OV.AVERAGE3 else np.array([0.4, -0.7, 1.3])
    vs = _vs([kind], [3, 0, 1, 2], w)
    xs = OV.positions(x, vs)
    f = rng.normal(size=3)
    F = np.zeros((4, 3))
    F[3] = f
    Fd = OV.distribute(F, xs, vs)
    assert np.allclose(Fd[:3].sum(0), f, rtol=0, atol=1e-14)
    torque = sum(np.cross(xs[i], Fd[i]) for i in range(3))
    assert np.allclose(torque, np.cross(xs[3], f), rtol=0, atol=1e-13)
