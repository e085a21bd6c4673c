"""GPU parity of the DMS force-field terms and of constraints against the oracle.

config 1 (3,000 atoms) with periodic torsions and harmonic impropers along the ligand
chain, flat-bottom and harmonic positional restraints, 1-4 exceptions on excluded ligand
pairs, and every water made rigid (two O-H and the H-H constraint).  The terms are FP64
on both sides; the exception and pair forces meet the FP32 pair forces in the state
accumulators, so the state forces keep the 1e-4 RMS bound of the pair path."""
import types

import numpy as np
import pytest

from openmm_sdm_plugin_amd import _native, systems
from oracle import oracle as O
from oracle import terms as OT

pytestmark = pytest.mark.gpu


def rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _terms(s, rng, constraints=True):
    x = np.asarray(s.positions)
    lig = [int(a) for a in s.ligand]
    tors = np.array([lig[k:k + 4] for k in range(0, 12, 2)])
    imps = np.array([lig[k:k + 4] for k in (14, 18)])
    ex = np.asarray(s.exclusions).reshape(-1, 2)
    ligset = set(lig)
    expairs = np.array([p for p in ex.tolist() if p[0] in ligset and p[1] in ligset and
                        np.linalg.norm(x[p[0]] - x[p[1]]) > 0.2][:12])
    fb = np.array([lig[0], lig[5], 100, 500, 901])
    ph = np.array([lig[7], 200, 1200])
    waters = [(h1, o, h2) for h1, o, h2 in np.asarray(s.angles).reshape(-1, 3).tolist()
              if o not in ligset]
    cons, clen = [], []
    if constraints:
        for h1, o, h2 in waters:
            for a, b in ((o, h1), (o, h2), (h1, h2)):
                cons.append((a, b))
                clen.append(float(np.linalg.norm(x[a] - x[b])))
    return types.SimpleNamespace(
        torsions=tors, torsion_n=rng.integers(1, 4, len(tors)).astype(np.int32),
        torsion_phase=rng.uniform(-3, 3, len(tors)), torsion_k=rng.uniform(1, 10, len(tors)),
        impropers=imps, improper_t0=rng.uniform(-1, 1, len(imps)),
        improper_k=rng.uniform(5, 50, len(imps)),
        posre_fb_atoms=fb, posre_fb_x0=x[fb] + rng.normal(scale=0.2, size=(len(fb), 3)),
        posre_fb_hk=rng.uniform(100, 500, len(fb)), posre_fb_tol=np.full(len(fb), 0.05),
        posre_h_atoms=ph, posre_h_x0=x[ph] + rng.normal(scale=0.1, size=(len(ph), 3)),
        posre_h_hk=rng.uniform(50, 200, (len(ph), 3)),
        exception_pairs=expairs, exception_qq=rng.uniform(-0.1, 0.1, len(expairs)),
        exception_sigma=np.full(len(expairs), 0.15),
        exception_epsilon=rng.uniform(0.1, 0.5, len(expairs)),
        constraints=np.array(cons, dtype=np.int32).reshape(-1, 2),
        constraint_r=np.array(clen))


def test_terms_components(hip_device):
    s = systems.make_config1(frame=1)
    ff = _terms(s, np.random.default_rng(7), constraints=False)
    print(f"case terms config1 n={s.n}", flush=True)
    ctx = _native.NativeContext(s, 1)
    ctx.set_forcefield_terms(ff)
    ctx.set_params(0, systems.params_config1())
    ctx.set_positions(0, s.positions)
    ctx.compute_components()
    e = ctx.energies()[0]
    topo = O.Topology(s)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, s.positions)
    e1, e2, f1, f2 = OT.energy_forces(s.positions, ff, np.asarray(s.box))
    assert abs(e1) > 1.0 and abs(e2) > 0.1  # every kind of term contributes
    assert abs(e["Ub"] - (Eb + e1)) <= 1e-9 * abs(Eb + e1)
    assert abs(e["U0"] - (E1 + e2)) <= 1e-5 * abs(E1), (e["U0"], E1 + e2)
    assert abs(e["U1"] - (E2 + e2)) <= 1e-5 * max(abs(E1), abs(E2))
    assert np.max(np.abs(ctx.forces(0, _native.FORCE_BONDED) - (Fb + f1))) < 1e-5
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE1), F1 + f2) < 1e-4
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE2), F2 + f2) < 1e-4
    # the exception forces alone, to a tight bound: difference of two runs
    ctx.set_forcefield_terms(None)
    ctx.compute_components()
    diff = ctx.forces(0, _native.FORCE_STATE1)
    ctx.set_forcefield_terms(ff)
    ctx.compute_components()
    diff = ctx.forces(0, _native.FORCE_STATE1) - diff
    assert np.max(np.abs(diff - f2)) < 1e-5


def test_constrained_step_matches_oracle(hip_device):
    s = systems.make_config1(frame=2)
    ff = _terms(s, np.random.default_rng(9))
    p = systems.params_config1()
    p.temperature = 0.0  # no noise: the oracle step below is deterministic
    ctx = _native.NativeContext(s, 1)
    ctx.set_forcefield_terms(ff)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.set_velocities(0, s.velocities)
    ctx.step(1)
    x_gpu, v_gpu = ctx.get_positions(0), ctx.get_velocities(0)
    topo = O.Topology(s)
    x0 = np.asarray(s.positions)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, x0)
    e1, e2, f1, f2 = OT.energy_forces(x0, ff, np.asarray(s.box))
    sc = O.execute_scalars(O.make_params(p), E1 + e2, E2 + e2, Eb + e1, 0.0)
    F = (1 - sc.sp) * (F1 + f2) + sc.sp * (F2 + f2) + Fb + f1
    xu, _ = O.langevin_update(x0, s.velocities, F, s.mass, p.step_size, p.friction, 0.0,
                              np.zeros_like(x0), form="cl")
    xc = OT.constrain(x0, xu, s.mass, ff.constraints, ff.constraint_r)
    vc = (xc - x0) / p.step_size
    assert np.max(np.abs(x_gpu - xc)) < 1e-7
    assert rms_rel(v_gpu, vc) < 1e-4
    assert ctx.energies()[0]["flags"] == 0


def test_constraints_hold_over_steps(hip_device):
    s = systems.make_config1(frame=0)
    ff = _terms(s, np.random.default_rng(3))
    ctx = _native.NativeContext(s, 2)
    ctx.set_forcefield_terms(ff, constraint_tol=1e-10)
    for r in range(2):
        ctx.set_params(r, systems.params_config1())
        ctx.set_positions(r, s.positions)
        ctx.set_velocities(r, s.velocities)
    ctx.step(40)
    for r in range(2):
        x = ctx.get_positions(r)
        d = np.linalg.norm(x[ff.constraints[:, 0]] - x[ff.constraints[:, 1]], axis=1)
        assert np.max(np.abs(d - ff.constraint_r) / ff.constraint_r) < 1e-8
        assert ctx.energies()[r]["flags"] == 0
    # an exception between atoms of different displacement is refused
    bad = _terms(s, np.random.default_rng(3), constraints=False)
    other = next(i for i in range(s.n) if i not in set(int(a) for a in s.ligand))
    bad.exception_pairs = np.array([[int(s.ligand[0]), other]])
    bad.exception_qq, bad.exception_sigma, bad.exception_epsilon = [0.1], [0.3], [0.2]
    with pytest.raises(_native.SDMError):
        ctx.set_forcefield_terms(bad)


def test_cm_motion_removal(hip_device):
    """With removal every step, a uniform drift is taken out before the update; the
    internal forces add no net momentum, so the momentum after a noise-free,
    friction-free step is zero (without removal it is M * drift)."""
    s = systems.make_config1(frame=1)
    p = systems.params_config1()
    p.temperature, p.friction = 0.0, 0.0
    m = np.asarray(s.mass)
    v = np.asarray(s.velocities) + np.array([0.7, -0.4, 0.2])
    out = []
    for every in (0, 1):
        ctx = _native.NativeContext(s, 1)
        ctx.set_cm_motion_removal(every)
        ctx.set_params(0, p)
        ctx.set_positions(0, s.positions)
        ctx.set_velocities(0, v)
        ctx.step(1)
        out.append((m[:, None] * ctx.get_velocities(0)).sum(0))
    scale = float((m[:, None] * np.abs(v)).sum())
    assert np.allclose(out[0], (m[:, None] * v).sum(0), atol=1e-6 * scale)
    assert np.max(np.abs(out[1])) < 1e-9 * scale


def test_cm_motion_removal_interval(hip_device):
    """With removal every 3 steps, the drift is taken out on the steps whose index (the
    step count before the step) is 0, 3, 6, ... as OpenMM's CMMotionRemover tests it at
    the start of the step: the drift is re-injected before each of 8 single steps and the
    momentum after the step is zero exactly on those."""
    s = systems.make_config1(frame=1)
    p = systems.params_config1()
    p.temperature, p.friction = 0.0, 0.0
    m = np.asarray(s.mass)
    drift = np.array([0.7, -0.4, 0.2])
    ctx = _native.NativeContext(s, 1)
    ctx.set_cm_motion_removal(3)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    scale = float(m.sum()) * float(np.abs(drift).max())
    removed = []
    for k in range(8):
        v = ctx.get_velocities(0)
        v = v - (m[:, None] * v).sum(0) / m.sum() + drift  # momentum M * drift
        ctx.set_velocities(0, v)
        ctx.step(1)
        assert ctx.energies()[0]["step"] == k + 1
        mom = (m[:, None] * ctx.get_velocities(0)).sum(0)
        removed.append(bool(np.max(np.abs(mom)) < 1e-6 * scale))
        if not removed[-1]:
            assert np.allclose(mom, m.sum() * drift, atol=1e-6 * scale)
    assert removed == [k % 3 == 0 for k in range(8)], removed
