"""Pins of the FP64 oracle (CPU).  The reference ships no golden vectors or tests, so each
oracle function is pinned by what the mathematics fixes: published known answers
(Random123 Philox vectors), closed forms at special points, asymptotes, finite
differences, conservation laws and symmetry invariants.  Each check is chosen so that a
plausible slip (a dropped term, a wrong sign, umax vs umax-ub, a wrong mixing rule) fails.
"""
import math

import numpy as np
import pytest

from openmm_sdm_plugin_amd import systems
from oracle import oracle as O

KE = 138.935456


# ---------------------------------------------------------------- Philox ----
def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                            0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                    [0xA4093822, 0x299F31D0]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_gauss3_moments_and_determinism():
    g = np.array([O.gauss3(99, 1, s, a) for s in range(200) for a in range(50)]).ravel()
    assert abs(g.mean()) < 0.03 and abs(g.var() - 1.0) < 0.04
    assert O.gauss3(5, 2, 7, 11) == O.gauss3(5, 2, 7, 11)
    assert O.gauss3(5, 2, 7, 11) != O.gauss3(5, 3, 7, 11)


# ------------------------------------------------------------- soft core ----
@pytest.mark.parametrize("method", [0, 1, 2])
def test_softcore_identity_below_ubcore(method):
    for u in (-50.0, 0.0, 9.999):
        up, fp = O.softcore(method, u, 200.0, 0.0625, 10.0)
        assert up == u and fp == 1.0


def test_softcore_rational_special_points():
    umax, a, ub = 200.0, 0.0625, 100.0
    up, fp = O.softcore(2, ub + 1e-10, umax, a, ub)
    assert abs(up - ub) < 1e-8 and abs(fp - 1.0) < 1e-6
    # asymptote is umax (not umax + ub), approached from below (slowly for a = 1/16)
    ups = [O.softcore(2, u, umax, a, ub)[0] for u in (1e3, 1e6, 1e12, 1e30, 1e60)]
    assert all(x < y for x, y in zip(ups, ups[1:])) and ups[-1] < umax
    assert umax - ups[-1] < 1e-3 and O.softcore(2, 1e30, umax, a, ub)[1] < 1e-20
    # closed form at g = 1: zeta = 5
    u = ub + a * (umax - ub)
    z = 5.0 ** a
    up, fp = O.softcore(2, u, umax, a, ub)
    assert math.isclose(up, (umax - ub) * (z - 1) / (z + 1) + ub, rel_tol=1e-14)
    assert math.isclose(fp, 4 * 3 / 5 * z / (1 + z) ** 2, rel_tol=1e-14)


def test_softcore_tanh_closed_form():
    umax, ub = 50.0, 10.0
    for u in (12.0, 60.0, 500.0):
        t = math.tanh((u - ub) / umax)
        up, fp = O.softcore(1, u, umax, 0.0625, ub)
        assert math.isclose(up, umax * t + ub, rel_tol=1e-14)
        assert math.isclose(fp, 1 - t * t, rel_tol=1e-12, abs_tol=1e-300)
    up, _ = O.softcore(1, 1e6, umax, 0.0625, ub)
    assert math.isclose(up, umax + ub)  # the reference's tanh form saturates at umax+ub


@pytest.mark.parametrize("method", [0, 1, 2])
def test_softcore_derivative_by_fd(method):
    for u in (15.0, 80.0, 140.0, 400.0, 3000.0):
        h = 1e-5 * max(1.0, abs(u))
        up1, _ = O.softcore(method, u + h, 200.0, 0.0625, 10.0)
        up0, _ = O.softcore(method, u - h, 200.0, 0.0625, 10.0)
        _, fp = O.softcore(method, u, 200.0, 0.0625, 10.0)
        assert math.isclose((up1 - up0) / (2 * h), fp, rel_tol=1e-6, abs_tol=1e-10)


def test_softcore_unknown_method_raises():
    with pytest.raises(ValueError):
        O.softcore(7, 500.0, 200.0, 0.0625, 10.0)


# ------------------------------------------------------------------ bias ----
def _params(**kw):
    p = systems.SDMParams(softcore_method=0)
    for k, v in kw.items():
        setattr(p, k, v)
    return O.make_params(p)


def test_bias_linear():
    sc = O.execute_scalars(_params(bias_method=0, lambda_=0.3), -100.0, -60.0, 5.0, 0.0)
    assert sc.bind_e == 40.0 and sc.bfp == 0.3 and math.isclose(sc.ebias, 12.0)
    assert math.isclose(sc.pot_energy, -100.0 + 12.0 + 5.0)


def test_bias_quadratic():
    sc = O.execute_scalars(_params(bias_method=1, gamma=0.02, wbcoeff=0.5, w0coeff=3.0),
                           0.0, 10.0, 0.0, 0.0)
    assert math.isclose(sc.ebias, 0.5 * 0.02 * 100 + 5.0 + 3.0)
    assert math.isclose(sc.bfp, 0.02 * 10 + 0.5)


def test_bias_ilogistic_equal_lambdas_is_linear():
    for u in (-30.0, 5.0, 250.0):
        sc = O.execute_scalars(_params(bias_method=2, lambda1=0.4, lambda2=0.4, alpha=0.1, u0=20.0,
                                       w0coeff=1.5), 0.0, u, 0.0, 0.0)
        assert math.isclose(sc.bfp, 0.4, rel_tol=1e-14)
        assert math.isclose(sc.ebias, 0.4 * u + 1.5, rel_tol=1e-14)


def test_bias_ilogistic_limits_and_derivative():
    kw = dict(bias_method=2, lambda1=0.1, lambda2=0.7, alpha=0.2, u0=30.0, w0coeff=0.0)
    lo = O.execute_scalars(_params(**kw), 0.0, -400.0, 0.0, 0.0)  # far below u0: W' -> lambda1
    hi = O.execute_scalars(_params(**kw), 0.0, 900.0, 0.0, 0.0)  # far above: W' -> lambda2
    assert math.isclose(lo.bfp, 0.1, abs_tol=1e-12) and math.isclose(hi.bfp, 0.7, abs_tol=1e-12)
    mid = O.execute_scalars(_params(**kw), 0.0, 30.0, 0.0, 0.0)  # at u0: W' = (l1+l2)/2
    assert math.isclose(mid.bfp, 0.4, rel_tol=1e-14)
    for u in (-20.0, 25.0, 31.0, 80.0):
        h = 1e-5
        w1 = O.execute_scalars(_params(**kw), 0.0, u + h, 0.0, 0.0).ebias
        w0 = O.execute_scalars(_params(**kw), 0.0, u - h, 0.0, 0.0).ebias
        d = O.execute_scalars(_params(**kw), 0.0, u, 0.0, 0.0).bfp
        assert math.isclose((w1 - w0) / (2 * h), d, rel_tol=1e-7)


def test_bias_ilogistic_alpha_zero_quirk():
    # alpha <= 0 skips the log term but W' stays (l2-l1)/(1+e^0)+l1 = (l1+l2)/2
    sc = O.execute_scalars(_params(bias_method=2, lambda1=0.2, lambda2=0.6, alpha=0.0, u0=0.0,
                                   w0coeff=2.0), 0.0, 50.0, 0.0, 0.0)
    assert math.isclose(sc.bfp, 0.4) and math.isclose(sc.ebias, 0.6 * 50 + 2.0)


def test_sp_is_chain_rule_product():
    p = _params(bias_method=2, softcore_method=2, lambda1=0.2, lambda2=0.5, alpha=0.05, u0=120.0,
                umax=200.0, acore=0.0625, ubcore=100.0)
    sc = O.execute_scalars(p, 0.0, 350.0, 0.0, 0.0)
    assert math.isclose(sc.sp, sc.bfp * sc.fp, rel_tol=1e-15)
    # dW/du by FD through the soft core
    h = 1e-4
    wp = O.execute_scalars(_params(bias_method=2, softcore_method=2, lambda1=0.2, lambda2=0.5,
                                   alpha=0.05, u0=120.0, umax=200.0, acore=0.0625, ubcore=100.0),
                           0.0, 350.0 + h, 0.0, 0.0).ebias
    wm = O.execute_scalars(_params(bias_method=2, softcore_method=2, lambda1=0.2, lambda2=0.5,
                                   alpha=0.05, u0=120.0, umax=200.0, acore=0.0625, ubcore=100.0),
                           0.0, 350.0 - h, 0.0, 0.0).ebias
    assert math.isclose((wp - wm) / (2 * h), sc.sp, rel_tol=1e-6)


def test_nonequilibrium_schedule_and_work_by_fd():
    kw = dict(bias_method=2, alpha=0.1, nonequilibrium=1, noneq_tmax=2.0, step_size=0.002,
              m_l1=0.3, m_l2=0.8, m_u0=15.0, m_w0=2.0, b_l1=0.05, b_l2=0.1, b_u0=40.0, b_w0=1.0)
    p = _params(**kw)
    t = 0.5
    sc = O.execute_scalars(p, 0.0, 70.0, 0.0, t)
    lam = t / 2.0
    assert math.isclose(p.lambda_, lam)
    assert math.isclose(p.lambda1, 0.3 * lam + 0.05) and math.isclose(p.lambda2, 0.8 * lam + 0.1)
    assert math.isclose(p.u0, 15 * lam + 40) and math.isclose(p.w0coeff, 2 * lam + 1)

    def W_of(l):
        q = _params(**kw)
        q.nonequilibrium = 0
        q.lambda1, q.lambda2, q.u0, q.w0coeff = 0.3 * l + 0.05, 0.8 * l + 0.1, 15 * l + 40, 2 * l + 1
        return O.execute_scalars(q, 0.0, 70.0, 0.0, 0.0).ebias

    h = 1e-6
    dWdl = (W_of(lam + h) - W_of(lam - h)) / (2 * h)
    assert math.isclose(p.work, (0.002 / 2.0) * dWdl, rel_tol=1e-6)
    assert sc.bind_e == 70.0


# ------------------------------------------------------------- nonbonded ----
def _two_atom(q1, q2, s1, s2, e1, e2, r, rc=1.0, eps_rf=78.5, box=10.0, lj_rule=0):
    s = systems.make_small()
    sys2 = s.with_coordinates(np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]]))
    sys2.charge = np.array([q1, q2])
    sys2.sigma = np.array([s1, s2])
    sys2.epsilon = np.array([e1, e2])
    sys2.mass = np.array([1.0, 1.0])
    sys2.exclusions = np.zeros((0, 2), np.int32)
    sys2.bonds = np.zeros((0, 2), np.int32)
    sys2.angles = np.zeros((0, 3), np.int32)
    sys2.bond_r0 = sys2.bond_k = sys2.angle_t0 = sys2.angle_k = np.zeros(0)
    sys2.displacement = np.zeros((2, 3))
    sys2.box = np.array([box, box, box])
    sys2.cutoff = rc
    sys2.eps_rf = eps_rf
    sys2.lj_rule = lj_rule
    topo = O.Topology(sys2)
    E, F, n = O.nonbonded(topo, sys2.positions, "brute")
    return E, F, n


def test_lj_minimum_and_zero_crossing():
    sig, eps = 0.3, 0.5
    rmin = 2 ** (1 / 6) * sig
    E, F, _ = _two_atom(0, 0, sig, sig, eps, eps, rmin)
    assert math.isclose(E, -eps, rel_tol=1e-12) and np.max(np.abs(F)) < 1e-9
    E, F, _ = _two_atom(0, 0, sig, sig, eps, eps, sig)
    assert abs(E) < 1e-12
    assert math.isclose(F[1, 0], 24 * eps / sig, rel_tol=1e-12)  # repulsive, pushes atom 1 to +x


def test_lorentz_berthelot_mixing():
    s1, s2, e1, e2 = 0.25, 0.35, 0.2, 0.8
    E, F, _ = _two_atom(0, 0, s1, s2, e1, e2, 2 ** (1 / 6) * 0.5 * (s1 + s2))
    assert math.isclose(E, -math.sqrt(e1 * e2), rel_tol=1e-12) and np.max(np.abs(F)) < 1e-9


def test_geometric_mixing():
    """OPLS (the reference's CustomNonbondedForce for OPLS, desmonddmsfile75.py:781):
    sigma12 = sqrt(s1 s2), so the minimum -sqrt(e1 e2) sits at 2^(1/6) sqrt(s1 s2), not at
    the Lorentz-Berthelot distance."""
    s1, s2, e1, e2 = 0.25, 0.35, 0.2, 0.8
    rg = 2 ** (1 / 6) * math.sqrt(s1 * s2)
    E, F, _ = _two_atom(0, 0, s1, s2, e1, e2, rg, lj_rule=1)
    assert math.isclose(E, -math.sqrt(e1 * e2), rel_tol=1e-12) and np.max(np.abs(F)) < 1e-9
    E, _, _ = _two_atom(0, 0, s1, s2, e1, e2, 2 ** (1 / 6) * 0.5 * (s1 + s2), lj_rule=1)
    assert E > -math.sqrt(e1 * e2) * (1 - 1e-4)
    # equal sigmas: both rules agree
    Ea, _, _ = _two_atom(0, 0, 0.3, 0.3, e1, e2, 0.41, lj_rule=0)
    Eg, _, _ = _two_atom(0, 0, 0.3, 0.3, e1, e2, 0.41, lj_rule=1)
    assert math.isclose(Ea, Eg, rel_tol=1e-14)


def test_reaction_field_special_cases():
    rc = 1.0
    # eps_rf = 1: k_rf = 0, c_rf = 1/rc -> shifted Coulomb
    E, F, _ = _two_atom(0.5, -0.8, 0.1, 0.1, 0.0, 0.0, 0.5, rc=rc, eps_rf=1.0)
    assert math.isclose(E, KE * 0.5 * -0.8 * (1 / 0.5 - 1 / rc), rel_tol=1e-12)
    assert math.isclose(F[1, 0], KE * 0.5 * -0.8 / 0.25, rel_tol=1e-12)
    # energy vanishes at the cutoff for any eps_rf; zero beyond
    E, F, n = _two_atom(1.0, 1.0, 0.1, 0.1, 0.0, 0.0, rc * (1 - 1e-9), rc=rc, eps_rf=78.5)
    assert abs(E) < 1e-6 and n == 1
    E, F, n = _two_atom(1.0, 1.0, 0.1, 0.1, 0.0, 0.0, rc * (1 + 1e-9), rc=rc)
    assert E == 0.0 and n == 0
    # conducting boundary (eps_rf -> inf): k_rf -> 1/(2 rc^3), c_rf -> 3/(2 rc)
    E, _, _ = _two_atom(1.0, 1.0, 0.1, 0.1, 0.0, 0.0, 0.5, rc=rc, eps_rf=1e12)
    assert math.isclose(E, KE * (1 / 0.5 + 0.25 / 2 - 1.5), rel_tol=1e-9)


def test_pair_force_is_minus_dE_dr():
    for r in (0.28, 0.4, 0.7, 0.95):
        h = 1e-7
        e1, _ = O.pair_energy(r + h, 0.3, 0.3, 0.4, 1.0, 78.5)
        e0, _ = O.pair_energy(r - h, 0.3, 0.3, 0.4, 1.0, 78.5)
        _, fr = O.pair_energy(r, 0.3, 0.3, 0.4, 1.0, 78.5)
        assert math.isclose(-(e1 - e0) / (2 * h), fr * r, rel_tol=1e-6)


@pytest.fixture(scope="module")
def small():
    s = systems.make_small(n_water=120, box=2.0, cutoff=0.9, frame=3)
    return s, O.Topology(s)


def test_brute_equals_cells_and_newton3(small):
    s, topo = small
    Eb, Fb, nb = O.nonbonded(topo, s.positions, "brute")
    Ec, Fc, nc = O.nonbonded(topo, s.positions, "cells")
    assert nb == nc and math.isclose(Eb, Ec, rel_tol=1e-12)
    assert np.max(np.abs(Fb - Fc)) < 1e-9 * np.max(np.abs(Fb))
    assert np.max(np.abs(Fb.sum(axis=0))) < 1e-8 * np.max(np.abs(Fb))


def test_nonbonded_force_by_fd(small):
    s, topo = small
    _, F, _ = O.nonbonded(topo, s.positions)
    rng = np.random.default_rng(0)
    for a in rng.choice(s.n, 6, replace=False):
        for k in range(3):
            x = s.positions.copy()
            h = 1e-6
            x[a, k] += h
            ep, _, _ = O.nonbonded(topo, x)
            x[a, k] -= 2 * h
            em, _, _ = O.nonbonded(topo, x)
            assert math.isclose(-(ep - em) / (2 * h), F[a, k], rel_tol=1e-5, abs_tol=1e-4)


def test_exclusions_remove_exactly_those_pairs(small):
    s, topo = small
    E, _, _ = O.nonbonded(topo, s.positions)
    s_all = s.with_coordinates(s.positions)
    s_all.exclusions = np.zeros((0, 2), np.int32)
    E_all, _, _ = O.nonbonded(O.Topology(s_all), s.positions)
    excl_sum = 0.0
    for a, b in s.exclusions:
        d = s.positions[a] - s.positions[b]
        d -= s.box * np.round(d / s.box)
        r = np.linalg.norm(d)
        if r < s.cutoff:
            e, _ = O.pair_energy(r, s.charge[a] * s.charge[b], 0.5 * (s.sigma[a] + s.sigma[b]),
                                 math.sqrt(s.epsilon[a] * s.epsilon[b]), s.cutoff, s.eps_rf)
            excl_sum += e
    assert math.isclose(E, E_all - excl_sum, rel_tol=1e-10, abs_tol=1e-8)


def test_periodic_translation_invariance(small):
    s, topo = small
    E, F, _ = O.nonbonded(topo, s.positions)
    x = s.positions + np.array([0.37, -1.1, 2.9])
    E2, F2, _ = O.nonbonded(topo, x)
    assert math.isclose(E, E2, rel_tol=1e-11)
    x = s.positions.copy()
    x[5] += s.box * np.array([1, -2, 3])
    E3, _, _ = O.nonbonded(topo, x)
    assert math.isclose(E, E3, rel_tol=1e-11)


# ------------------------------------------------------------ dual state ----
def test_zero_and_box_vector_displacement(small):
    s, topo = small
    U0, u, _, dF, _, _ = O.sdm_direct(topo, s.positions, np.zeros((s.n, 3)))
    assert u == 0.0 and np.all(dF == 0.0)
    d = np.zeros((s.n, 3))
    d[s.ligand] = (0.0, s.box[1], 0.0)
    U0, u, _, dF, _, _ = O.sdm_direct(topo, s.positions, d)
    E1, E2, *_ = O.reference_eval(topo, s.positions, d)
    assert abs(u) < 1e-9 and abs(E2 - E1) < 1e-9 * abs(E1)


def test_uniform_displacement_gives_zero_u(small):
    s, topo = small
    d = np.tile([0.4, -0.2, 1.3], (s.n, 1))
    _, u, _, _, _, _ = O.sdm_direct(topo, s.positions, d)
    E1, E2, *_ = O.reference_eval(topo, s.positions, d)
    assert u == 0.0 and abs(E2 - E1) < 1e-9 * abs(E1)


@pytest.mark.parametrize("maker", [systems.make_small, systems.make_config1])
def test_direct_u_equals_two_sweep_difference(maker):
    s = maker()
    topo = O.Topology(s)
    U0, u, F0, dF, _, _ = O.sdm_direct(topo, s.positions)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, s.positions)
    assert math.isclose(U0, E1, rel_tol=1e-13)
    assert math.isclose(u, E2 - E1, rel_tol=1e-9, abs_tol=1e-9 * abs(E1))
    assert np.max(np.abs(F0 + dF - F2)) <= 1e-8 * max(1.0, np.max(np.abs(F2)))


def test_hybrid_force_is_minus_gradient_of_alchemical_potential():
    s = systems.make_small(n_water=40, n_lig=4, box=1.9, cutoff=0.8, frame=5, displ=(0.5, 0, 0))
    topo = O.Topology(s)
    p = systems.SDMParams(bias_method=2, softcore_method=2, lambda1=0.2, lambda2=0.6, alpha=0.02,
                          u0=150.0, umax=300.0, acore=0.0625, ubcore=20.0)
    op = O.make_params(p)

    def U_alch(x):
        E1, E2, Eb, *_ = O.reference_eval(topo, x)
        return O.execute_scalars(O.make_params(p), E1, E2, Eb, 0.0).pot_energy

    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, s.positions)
    sc = O.execute_scalars(op, E1, E2, Eb, 0.0)
    assert E2 - E1 > p.ubcore  # the soft core is active
    F = sc.sp * F2 + (1 - sc.sp) * F1 + Fb
    for a in list(s.ligand[:2]) + [int(s.ligand[-1]) + 1, int(s.ligand[-1]) + 7]:
        for k in range(3):
            x = s.positions.copy()
            h = 1e-6
            x[a, k] += h
            up = U_alch(x)
            x[a, k] -= 2 * h
            um = U_alch(x)
            assert math.isclose(-(up - um) / (2 * h), F[a, k], rel_tol=1e-4, abs_tol=1e-3)


def test_force_mix_linear_endpoints(small):
    s, topo = small
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, s.positions)
    for lam, ref in ((0.0, F1 + Fb), (1.0, F2 + Fb)):
        sc = O.execute_scalars(_params(bias_method=0, lambda_=lam), E1, E2, Eb, 0.0)
        F = np.empty_like(F1)
        O.lib().orc_force_mix(s.n, sc.sp, F1.ctypes.data_as(O._dp), F2.ctypes.data_as(O._dp),
                              Fb.ctypes.data_as(O._dp), F.ctypes.data_as(O._dp))
        assert np.array_equal(F, ref)


@pytest.mark.parametrize("lj_rule", [0, 1])
def test_du_dparams_by_fd(lj_rule):
    s = systems.make_small(n_water=60, n_lig=5, frame=4, displ=(0.7, 0.2, 0.0))
    s.lj_rule = lj_rule
    topo = O.Topology(s)
    sel = s.ligand[:3].astype(np.int32)
    dq, dsig, deps = O.du_dparams(topo, s.positions, sel)
    for k, a in enumerate(sel):
        for arr, grad in (("charge", dq), ("sigma", dsig), ("epsilon", deps)):
            h = 1e-4 * max(abs(getattr(s, arr)[a]), 0.1)  # central FD, h = 1e-4 relative
            vals = []
            for sgn in (1, -1):
                s2 = s.with_coordinates(s.positions)
                setattr(s2, arr, getattr(s, arr).copy())
                getattr(s2, arr)[a] += sgn * h
                vals.append(O.sdm_direct(O.Topology(s2), s.positions)[1])
            fd = (vals[0] - vals[1]) / (2 * h)
            assert math.isclose(fd, grad[k], rel_tol=1e-4, abs_tol=1e-6), (arr, k, fd, grad[k])


# --------------------------------------------------------------- bonded ----
def test_bonded_force_by_fd_and_rest_geometry():
    s = systems.make_small(n_water=10, n_lig=6, frame=2)
    topo = O.Topology(s)
    x = s.positions + np.random.default_rng(3).normal(scale=0.01, size=s.positions.shape)
    E, F = O.bonded(topo, x)
    for a in (0, 3, 7, 12):
        for k in range(3):
            h = 1e-6
            xp = x.copy()
            xp[a, k] += h
            xm = x.copy()
            xm[a, k] -= h
            fd = -(O.bonded(topo, xp)[0] - O.bonded(topo, xm)[0]) / (2 * h)
            assert math.isclose(fd, F[a, k], rel_tol=1e-5, abs_tol=1e-4)
    # generated geometry sits at the bond/angle rest values up to float32 rounding
    E0, _ = O.bonded(topo, s.positions)
    assert E0 < 1e-3


# ------------------------------------------------------------- Langevin ----
def test_langevin_zero_temperature_closed_form():
    rng = np.random.default_rng(1)
    n = 7
    x, v, F = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    m = rng.uniform(1, 20, n)
    dt, gam = 0.002, 3.0
    xi = rng.normal(size=(n, 3))
    vs = math.exp(-gam * dt)
    fs = (1 - vs) / gam
    vexp = vs * v + fs * F / m[:, None]
    for form in ("reference", "cl"):
        x1, v1 = O.langevin_update(x, v, F, m, dt, gam, 0.0, xi, form)
        assert np.allclose(v1, vexp, rtol=1e-12, atol=1e-14)
        assert np.allclose(x1, x + dt * vexp, rtol=1e-13, atol=1e-15)


def test_langevin_zero_friction_quirk():
    x, v, F = np.ones((2, 3)), np.ones((2, 3)), np.ones((2, 3))
    m = np.array([2.0, 4.0])
    xi = np.zeros((2, 3))
    _, v_ref = O.langevin_update(x, v, F, m, 0.001, 0.0, 300.0, xi, "reference")
    assert np.all(np.isnan(v_ref))  # Reference platform: tau = 0 gives NaN
    _, v_cl = O.langevin_update(x, v, F, m, 0.001, 0.0, 300.0, xi, "cl")
    assert np.allclose(v_cl, v + 0.001 * F / m[:, None], rtol=1e-12)


def test_langevin_equipartition():
    n, T, dt, gam = 400, 300.0, 0.002, 5.0
    m = np.full(n, 12.0)
    x, v, F = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    kes = []
    for step in range(600):
        xi = np.array([O.gauss3(17, 0, step, a) for a in range(n)])
        x, v = O.langevin_update(x, v, F, m, dt, gam, T, xi, "cl")
        if step > 300:
            kes.append(O.kinetic_energy(v, m))
    kT = 1.380658e-23 * 6.0221367e23 / 1000.0 * T
    assert math.isclose(np.mean(kes), 1.5 * n * kT, rel_tol=0.03)


def test_kinetic_energy_closed_form():
    v = np.array([[1.0, 2.0, 2.0], [0.0, 0.0, 3.0], [5.0, 5.0, 5.0]])
    m = np.array([2.0, 4.0, 0.0])
    assert math.isclose(O.kinetic_energy(v, m), 0.5 * (2 * 9 + 4 * 9))


def test_work_alpha_slope_is_dW_dalpha():
    """With only the alpha slope set, one switching step adds dt/tmax * m_alpha * dW/dalpha,
    dW/dalpha taken by central differences of the bias energy W(alpha) at fixed lambda."""
    from openmm_sdm_plugin_amd import systems as S
    p = S.params_config1()
    p.bias_method, p.softcore_method = 2, 0
    p.u0, p.w0coeff = 20.0, 0.0
    E1, E2, Eb = 0.0, 35.0, 0.0  # u = 35

    def W(alpha):
        q = O.make_params(p)
        q.nonequilibrium, q.alpha, q.lambda1, q.lambda2 = 0, alpha, 0.2, 0.7
        return O.execute_scalars(q, E1, E2, Eb, 0.0).ebias
    q = O.make_params(p)
    q.nonequilibrium, q.noneq_tmax, q.step_size = 1, 1.0, 0.002
    q.alpha, q.m_alpha = 0.05, 0.3
    q.m_l1 = q.m_l2 = q.m_u0 = q.m_w0 = 0.0
    q.b_l1, q.b_l2, q.b_u0, q.b_w0 = 0.2, 0.7, 20.0, 0.0
    q.work = 0.0
    O.execute_scalars(q, E1, E2, Eb, 0.4)  # lambda = 0.4: alpha = 0.05 + 0.3 * 0.4
    a, h = 0.17, 1e-6
    dwda = (W(a + h) - W(a - h)) / (2 * h)
    assert abs(dwda) > 1.0
    assert q.work == pytest.approx(0.002 * 0.3 * dwda, rel=1e-7)
    q.m_alpha, q.work = 0.0, 0.0  # the reference: no alpha term
    O.execute_scalars(q, E1, E2, Eb, 0.4)
    assert q.work == 0.0
