"""Dynamics of the HIP path over many steps: a deterministic 50-step trajectory against
the FP64 oracle, and the Langevin thermostat's equipartition over 1,000 steps of config 1
(LangevinIntegratorSDM.cpp:153-183, ReferenceStochasticDynamicsSDM.cpp:131-266,
kinetic energy as ReferenceSDMKernels.cpp:108-135)."""
import types

import numpy as np
import pytest

from openmm_sdm_plugin_amd import _native, systems
from oracle import oracle as O

pytestmark = pytest.mark.gpu

KB = 1.380658e-23 * 6.0221367e23 / 1000.0  # kJ/mol/K (OpenCLSDMKernels.cpp:57-60)


def rms_rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))


def close(a, b, tol=1e-5, floor=1.0):
    return abs(a - b) <= tol * max(floor, abs(b))


def test_config1_zero_temperature_trajectory(hip_device):
    """50 deterministic steps (T = 0: no noise) of config 1 against the oracle's
    Reference-form steps: the pair forces of every step enter the trajectory.  The
    velocity change is compared at 1e-4 RMS relative (the force bound); positions within
    1e-4 nm, the drift that the FP32 pair coordinates (ulp ~2.4e-7 nm in a 3.1 nm box)
    leave after 50 steps (measured 1.5e-5 nm; a 0.1 kJ/mol/nm force error on a hydrogen
    moves it ~1e-4 nm in 50 steps)."""
    s = systems.make_config1(frame=4)
    p = systems.params_config1()
    p.temperature = 0.0
    p.lambda1, p.lambda2, p.alpha, p.u0 = 0.2, 0.5, 0.05, 60.0
    n = 50
    ctx = _native.NativeContext(s, 1)
    ctx.set_tuning(pair_group=0, pair_newton3=1, pair_split=4, use_graph=0)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.set_velocities(0, s.velocities)
    ctx.step(n)
    e = ctx.energies()[0]
    topo = O.Topology(s)
    x, v, t, st, rep = O.ref_step(topo, O.make_params(p), s.positions, s.velocities, 0.0, 0,
                                  p.seed, 0, n, use_cl_langevin=True)
    xg, vg = ctx.get_positions(0), ctx.get_velocities(0)
    dx = float(np.max(np.abs(xg - x)))
    dv = rms_rel(vg - s.velocities, v - s.velocities)
    print(f"T=0 trajectory: max |dx| {dx:.2e} nm, velocity-change rms rel {dv:.2e}",
          flush=True)
    assert e["flags"] == 0 and e["step"] == n
    assert dx < 1e-4, dx
    assert dv < 1e-4, dv
    assert close(e["bind_e"], rep.s.bind_e) and close(e["W"], rep.s.ebias)


def _rigid_water(s):
    """Constraint-only force-field terms: every water of a config made rigid (O-H, O-H,
    H-H at their built lengths)."""
    x = np.asarray(s.positions)
    lig = set(int(a) for a in s.ligand)
    cons, clen = [], []
    for h1, o, h2 in np.asarray(s.angles).reshape(-1, 3).tolist():
        if o in lig:
            continue
        for a, b in ((o, h1), (o, h2), (h1, h2)):
            cons.append((a, b))
            clen.append(float(np.linalg.norm(x[a] - x[b])))
    z4, z3, z0 = np.zeros((0, 4), np.int32), np.zeros((0, 3)), np.zeros(0)
    zi = np.zeros(0, np.int32)
    return types.SimpleNamespace(
        torsions=z4, torsion_n=zi, torsion_phase=z0, torsion_k=z0, impropers=z4,
        improper_t0=z0, improper_k=z0, posre_fb_atoms=zi, posre_fb_x0=z3, posre_fb_hk=z0,
        posre_fb_tol=z0, posre_h_atoms=zi, posre_h_x0=z3, posre_h_hk=z3,
        exception_pairs=np.zeros((0, 2), np.int32), exception_qq=z0, exception_sigma=z0,
        exception_epsilon=z0, constraints=np.array(cons, np.int32), constraint_r=np.array(clen))


def _run_ke(s, ff, friction, R=4, nsteps=1000, every=10):
    ctx = _native.NativeContext(s, R)
    if ff is not None:
        ctx.set_forcefield_terms(ff, constraint_tol=1e-10)
    for r in range(R):
        p = systems.params_config1()
        p.friction, p.seed = friction, 900 + r
        ctx.set_params(r, p)
        ctx.set_positions(r, s.positions)
        ctx.set_velocities(r, s.velocities)
    trace = []
    for _ in range(nsteps // every):
        ctx.step(every)
        trace.append(ctx.kinetic_energy())
    flags = [e["flags"] for e in ctx.energies()]
    return np.array(trace), flags, ctx


def test_config1_equipartition_rigid_water(hip_device):
    """1,000 steps at 300 K (1 fs) of config 1 with rigid waters, 4 replicas on their own
    noise: the mean kinetic energy over the last 500 steps is (3N - N_c)/2 kT within 3%.
    Friction 20/ps so the heat the lattice start releases is carried off within the run
    (at 1/ps the run is still cooling after 1,000 steps)."""
    s = systems.make_config1(frame=0)
    ff = _rigid_water(s)
    trace, flags, _ = _run_ke(s, ff, 20.0)
    dof = 3 * s.n - ff.constraints.shape[0]
    want = 0.5 * dof * KB * 300.0
    got = float(trace[50:].mean())
    print(f"rigid water: <KE>/target {got / want:.4f} (dof {dof})", flush=True)
    assert flags == [0, 0, 0, 0]
    assert abs(got / want - 1.0) < 0.03, got / want


def test_config1_thousand_steps_flexible(hip_device):
    """1,000 steps at 300 K of config 1 as specified (flexible water): no flags, finite
    energies, and the mean kinetic energy over the last 500 steps within [0.97, 1.08] of
    3N/2 kT.  The kinetic energy is that of the half-step velocities (the reference
    ignores the time shift, ReferenceSDMKernels.cpp:112-119), which a leapfrog integrator
    overestimates for the stiff O-H bonds (omega dt ~ 0.7 at 1 fs): about +4% here."""
    s = systems.make_config1(frame=0)
    trace, flags, ctx = _run_ke(s, None, 20.0)
    want = 1.5 * s.n * KB * 300.0
    got = float(trace[50:].mean())
    print(f"flexible water: <KE>/target {got / want:.4f}", flush=True)
    assert flags == [0, 0, 0, 0]
    assert all(np.isfinite(e["pot_energy"]) for e in ctx.energies())
    assert 0.97 < got / want < 1.08, got / want
