"""The reference's example system example/temoa-g1-g4.dms on the GPU.

20,446 atoms: receptor (residue 1), two guests (residues 2 and 3), 6,734 rigid waters;
torsions, 1-4 exceptions, flat-bottom positional restraints and X-H/water constraints
read by openmm_sdm_plugin_amd.dms (tests/golden/temoa_g1_g4.npz, made by
tests/golden/make_dms_fixture.py).  Set up as example/test_explicit.py:24-191: guest 1
displaced by +2.2 nm and guest 2 by -2.2 nm along each axis, centroid restraints of each
guest to the receptor and the alignment force between the guests, ILogistic bias at
lambda 0.5 with the rational soft core.  The example's PME is replaced by the reaction
field (eps_rf 78.5, rc 1.0 nm), the engine's long-range treatment; the oracle evaluates
the same model."""
import dataclasses
import os

import numpy as np
import pytest

from openmm_sdm_plugin_amd import SDMUtils, _native, systems
from openmm_sdm_plugin_amd.dms import DMSForceField, to_sdm_system
from oracle import oracle as O
from oracle import restraints as OR
from oracle import terms as OT

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KCAL = 4.184


def load_temoa():
    z = np.load(os.path.join(HERE, "golden", "temoa_g1_g4.npz"))
    names = [f.name for f in dataclasses.fields(DMSForceField)]
    kw = {k: z[k] for k in names if k in z.files}
    cutoff, opls, cm = z["scalars"]
    ff = DMSForceField(**kw, opls=bool(opls), nonbonded_method="PME", cutoff=float(cutoff),
                       remove_cm_motion=bool(cm))
    resid = z["resid"]
    lig1, lig2 = np.nonzero(resid == 2)[0], np.nonzero(resid == 3)[0]
    rcpt = np.nonzero(resid == 1)[0]
    disp = np.zeros((ff.n, 3))
    disp[lig1] = 2.2
    disp[lig2] = -2.2
    s = to_sdm_system(ff, z["positions"], z["velocities"], displacement=disp, name="temoa")
    u = SDMUtils(s)
    kf, r0 = 25.0 * KCAL * 100, 0.45  # 25 kcal/mol/A^2, 4.5 A
    u.addRestraintForce(lig_cm_particles=lig1, rcpt_cm_particles=rcpt, kfcm=kf, tolcm=r0)
    u.addRestraintForce(lig_cm_particles=lig2, rcpt_cm_particles=rcpt, kfcm=kf, tolcm=r0,
                        offset=[2.2, 2.2, 2.2])
    u.addAlignmentForce(liga_ref_particles=[lig1[0] + k for k in (8, 6, 4)],
                        ligb_ref_particles=[lig2[0] + k for k in (3, 5, 1)],
                        kfdispl=kf, ktheta=50.0 * KCAL, kpsi=50.0 * KCAL,
                        offset=[2.2, 2.2, 2.2])
    return s, ff


def temoa_params():
    p = systems.params_config1()
    p.temperature, p.friction, p.step_size = 300.0, 0.1, 0.001
    p.bias_method, p.lambda_, p.lambda1, p.lambda2 = 2, 0.5, 0.5, 0.5
    p.alpha, p.u0, p.w0coeff = 0.0, 0.0, 0.0
    p.softcore_method, p.umax, p.ubcore, p.acore = 2, 100 * KCAL, 50 * KCAL, 0.0625
    return p


def rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def test_temoa_components_match_oracle(hip_device):
    s, ff = load_temoa()
    print(f"case temoa n={s.n} constraints={ff.constraints.shape[0]}", flush=True)
    ctx = _native.NativeContext(s, 1)
    ctx.set_params(0, temoa_params())
    ctx.set_positions(0, s.positions)
    ctx.compute_components()
    e = ctx.energies()[0]
    x = s.positions
    topo = O.Topology(s)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, x)
    e1, e2, f1, f2 = OT.energy_forces(x, ff, np.asarray(s.box))
    er, fr = OR.energy_forces(x, s.restraints, np.asarray(s.mass))
    assert abs(e["U0"] - (E1 + e2)) <= 1e-5 * abs(E1 + e2), (e["U0"], E1 + e2)
    assert abs(e["u"] - (E2 - E1)) <= 1e-5 * max(1.0, abs(E2 - E1))
    assert abs(e["Ub"] - (Eb + e1 + er)) <= 1e-9 * abs(Eb + e1 + er)
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE1), F1 + f2) < 1e-4
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE2), F2 + f2) < 1e-4
    assert np.max(np.abs(ctx.forces(0, _native.FORCE_BONDED) - (Fb + f1 + fr))) < 1e-5
    assert e["flags"] == 0


def test_temoa_runs_with_constraints(hip_device):
    s, ff = load_temoa()
    ctx = _native.NativeContext(s, 2)
    for r in range(2):
        ctx.set_params(r, temoa_params())
        ctx.set_positions(r, s.positions)
        ctx.set_velocities(r, s.velocities)
    ctx.step(200)
    ke = ctx.kinetic_energy()
    for r, e in enumerate(ctx.energies()):
        assert e["flags"] == 0 and np.isfinite(e["pot_energy"]) and e["step"] == 200
        x = ctx.get_positions(r)
        c = ff.constraints
        d = np.linalg.norm(x[c[:, 0]] - x[c[:, 1]], axis=1)
        assert np.max(np.abs(d - ff.constraint_r) / ff.constraint_r) < 1e-8
        # kinetic temperature with 3N - constraints degrees of freedom stays near 300 K
        dof = 3 * s.n - c.shape[0]
        t = 2 * ke[r] / (dof * 0.0083144626)
        assert 250.0 < t < 350.0, t


def test_temoa_kinetic_energy_projects_constraints(hip_device):
    """computeKineticEnergy on the constrained system equals the oracle's: the velocities
    projected onto the 20,278 constraints at the current positions (ReferenceSDMKernels.cpp
    :108-135), for the file's velocities after 20 steps and for random velocities, whose
    bond components the projection removes (the stored velocities are unchanged)."""
    s, ff = load_temoa()
    ctx = _native.NativeContext(s, 2)
    for r in range(2):
        ctx.set_params(r, temoa_params())
        ctx.set_positions(r, s.positions)
        ctx.set_velocities(r, s.velocities)
    ctx.step(20)
    rng = np.random.default_rng(11)
    v_rand = rng.normal(scale=0.5, size=(s.n, 3))
    ctx.set_velocities(1, v_rand)
    ke = ctx.kinetic_energy()
    m = np.asarray(s.mass)
    for r in range(2):
        x, v = ctx.get_positions(r), ctx.get_velocities(r)
        want = OT.kinetic_energy(x, v, m, ff.constraints)
        raw = OT.kinetic_energy(x, v, m)
        assert abs(ke[r] - want) <= 1e-9 * want, (r, ke[r], want, raw)
    assert np.array_equal(ctx.get_velocities(1), v_rand)
    # the random velocities carry much kinetic energy along the bonds
    assert OT.kinetic_energy(x, v_rand, m) > 1.1 * want
    assert all(e["flags"] == 0 for e in ctx.energies())
