"""The reference's first example (example/test.py) on the GPU, through the Python facade.

example/oa-g6-g3-align-restr_0_displaced.dms: a 230-atom host (residue 1) with guests g6
(residue 2) and g3 (residue 3) in vacuum, read as example/test.py:63 does
(CutoffNonPeriodic, 15 nm cutoff: no periodic images, reaction field with OpenMM's
default dielectric 78.3) from tests/golden/oa_g6_g3.npz (tests/golden/make_dms_fixture.py).
Set up as example/test.py:24-191: g3 displaced by (-15.559, -3.0, 8.6) A and g6 by the
opposite vector, centroid restraints of both guests to the host (the second with the
displacement as offset), the alignment force between them, the ILogistic bias at lambda
0.025 with the rational soft core (umax 100, ubcore 50 kcal/mol, acore 1/16), 300 K,
friction 0.1/ps, 1 fs.  The steps run as example/test.py:198-212 does (step(n) in
chunks, then getPotEnergy / getBindE), checked against the FP64 oracle (group 2 at x and
x+d, bonded, DMS terms, restraints, scalar stage) at the positions the last step of each
chunk evaluated.  Tolerances: energies 1e-5 relative, forces 1e-4 RMS relative.
Both tests run the file as example/test.py reads it and as createSystem(OPLS=True) reads
it (tests/golden/oa_g6_g3_opls.npz): its nonbonded_info declares the geometric rule, and
with OPLS the reference's reader moves LJ to a geometric-sigma CustomNonbondedForce and
writes torsions as f cos(n phi - phi0) (desmonddmsfile75.py:652-656, :780-801).
"""
import dataclasses
import os

import numpy as np
import pytest

from openmm_sdm_plugin_amd import SDMUtils, _native
from openmm_sdm_plugin_amd.dms import DMSForceField, to_sdm_system
from openmm_sdm_plugin_amd.integrator import Context, LangevinIntegratorSDM
from oracle import oracle as O
from oracle import restraints as OR
from oracle import terms as OT

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KCAL = 4.184
ANG = 0.1
DISPL = np.array([-15.559, -3.000, 8.600]) * ANG  # example/test.py:36


def rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def load_oa(opls=False):
    z = np.load(os.path.join(HERE, "golden", "oa_g6_g3_opls.npz" if opls else "oa_g6_g3.npz"))
    names = [f.name for f in dataclasses.fields(DMSForceField)]
    kw = {k: z[k] for k in names if k in z.files}
    cutoff, opls, cm = z["scalars"]
    ff = DMSForceField(**kw, opls=bool(opls), nonbonded_method="CutoffNonPeriodic",
                       cutoff=float(cutoff), remove_cm_motion=bool(cm), box=None)
    resid = z["resid"]
    rcpt = np.nonzero(resid == 1)[0]
    lig1, lig2 = np.nonzero(resid == 3)[0], np.nonzero(resid == 2)[0]
    disp = np.zeros((ff.n, 3))
    disp[lig1] = DISPL
    disp[lig2] = -DISPL
    s = to_sdm_system(ff, z["positions"], z["velocities"], displacement=disp, name="oa-g6-g3")
    u = SDMUtils(s)  # example/test.py:103-161
    kf, r0 = 25.0 * KCAL / ANG ** 2, 4.5 * ANG
    u.addRestraintForce(lig_cm_particles=lig1, rcpt_cm_particles=rcpt, kfcm=kf, tolcm=r0,
                        offset=[0.0, 0.0, 0.0])
    u.addRestraintForce(lig_cm_particles=lig2, rcpt_cm_particles=rcpt, kfcm=kf, tolcm=r0,
                        offset=list(DISPL))
    u.addAlignmentForce(liga_ref_particles=[lig1[0] + k for k in (4, 3, 2)],
                        ligb_ref_particles=[lig2[0] + k for k in (5, 4, 3)],
                        kfdispl=kf, ktheta=50.0 * KCAL, kpsi=50.0 * KCAL, offset=list(DISPL))
    return s, ff, lig1, lig2


def make_integrator(s, lig1, lig2):
    """example/test.py:178-199, in kJ/mol and nm."""
    integ = LangevinIntegratorSDM(300.0, 0.1, 0.001, s.n)
    integ.setBiasMethod(LangevinIntegratorSDM.ILogisticMethod)
    lmbd = 0.025
    integ.setLambda(lmbd)
    integ.setLambda1(lmbd)
    integ.setLambda2(lmbd)
    integ.setAlpha(0.0)
    integ.setU0(0.0)
    integ.setW0coeff(0.0)
    for i in lig1:
        integ.setDisplacement(int(i), *DISPL)
    for i in lig2:
        integ.setDisplacement(int(i), *(-DISPL))
    integ.setSoftCoreMethod(LangevinIntegratorSDM.RationalMethod)
    integ.setUmax(100.0 * KCAL)
    integ.setUbcore(50.0 * KCAL)
    integ.setAcore(0.0625)
    integ.setRandomNumberSeed(2024)
    return integ


def oracle_eval(s, ff, x):
    """(E1, E2, Eb, F1, F2, Fb) of the whole force field at x: group 2 = LJ + reaction
    field + 1-4 exceptions, group 1 = bonds, angles, torsions, impropers, positional and
    SDMUtils restraints."""
    topo = O.Topology(s)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, x)
    e1, e2, f1, f2 = OT.energy_forces(x, ff, None)
    e1y, e2y, f1y, f2y = OT.energy_forces(np.asarray(x) + s.displacement, ff, None)
    er, fr = OR.energy_forces(x, s.restraints, np.asarray(s.mass))
    # exceptions join a pair of equal displacement: the same at x and at x + d
    assert abs(e2y - e2) <= 1e-9 * max(1.0, abs(e2))
    return E1 + e2, E2 + e2, Eb + e1 + er, F1 + f2, F2 + f2, Fb + f1 + fr


@pytest.mark.parametrize("opls", [False, True])
def test_vacuum_components_match_oracle(hip_device, opls):
    s, ff, lig1, lig2 = load_oa(opls)
    assert not s.periodic and ff.constraints.shape[0] == 86 and s.lj_rule == int(opls)
    print(f"case oa-g6-g3 n={s.n} non-periodic cutoff={s.cutoff} opls={opls}", flush=True)
    ctx = _native.NativeContext(s, 1)
    integ = make_integrator(s, lig1, lig2)
    ctx.set_params(0, integ._p)
    ctx.set_positions(0, s.positions)
    ctx.compute_components()
    e = ctx.energies()[0]
    E1, E2, Eb, F1, F2, Fb = oracle_eval(s, ff, s.positions)
    assert e["flags"] == 0
    assert abs(e["U0"] - E1) <= 1e-5 * max(1.0, abs(E1)), (e["U0"], E1)
    assert abs(e["u"] - (E2 - E1)) <= 1e-5 * max(1.0, abs(E2 - E1)), (e["u"], E2 - E1)
    assert abs(e["Ub"] - Eb) <= 1e-5 * max(1.0, abs(Eb)), (e["Ub"], Eb)
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE1), F1) < 1e-4
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE2), F2) < 1e-4
    assert rms_rel(ctx.forces(0, _native.FORCE_BONDED), Fb) < 1e-4
    sc = O.execute_scalars(O.make_params(integ._p), E1, E2, Eb, 0.0)
    assert abs(e["bind_e"] - sc.bind_e) <= 1e-5 * max(1.0, abs(sc.bind_e))
    assert abs(e["pot_energy"] - sc.pot_energy) <= 1e-5 * (abs(E1) + abs(sc.ebias) + abs(Eb))
    # periodic images play no part: translating the system by 30 nm changes nothing
    ctx.set_positions(0, np.asarray(s.positions) + np.array([30.0, -30.0, 30.0]))
    ctx.compute()
    e2 = ctx.energies()[0]
    assert abs(e2["U0"] - E1) <= 1e-5 * max(1.0, abs(E1)) and e2["flags"] == 0


@pytest.mark.parametrize("opls", [False, True])
def test_vacuum_facade_steps_match_oracle(hip_device, opls):
    """example/test.py's loop through the facade: step(n) in chunks, then getPotEnergy()
    and getBindE(), each against the oracle at the positions the chunk's last step
    evaluated (the record a step returns is that of the positions before its update)."""
    s, ff, lig1, lig2 = load_oa(opls)
    print(f"case oa-g6-g3 facade steps opls={opls}", flush=True)
    integ = make_integrator(s, lig1, lig2)
    ctx = Context(s, integ)
    assert integ.getConstraintTolerance() == 1e-5
    nprnt = 25
    lines = []
    for _ in range(4):
        integ.step(nprnt - 1)
        x = ctx.getPositions()
        integ.step(1)
        pot, bind = integ.getPotEnergy(), integ.getBindE()
        E1, E2, Eb, *_ = oracle_eval(s, ff, x)
        sc = O.execute_scalars(O.make_params(integ._p), E1, E2, Eb, 0.0)
        assert abs(bind - sc.bind_e) <= 1e-5 * max(1.0, abs(sc.bind_e)), (bind, sc.bind_e)
        assert abs(pot - sc.pot_energy) <= 1e-5 * (abs(E1) + abs(sc.ebias) + abs(Eb)), \
            (pot, sc.pot_energy)
        # the report line of example/test.py:212 (kcal/mol)
        lines.append("%f %f %f %f %f %f %f %f %f" % (
            integ.getTemperature(), integ.getLambda(), integ.getLambda1(),
            integ.getLambda2(), integ.getAlpha() * KCAL, integ.getU0() / KCAL,
            integ.getW0coeff() / KCAL, pot / KCAL, bind / KCAL))
    print("\n".join(lines), flush=True)
    # the constraints hold to the facade's tolerance, and the kinetic energy is finite
    xf = ctx.getPositions()
    c = ff.constraints
    d = np.linalg.norm(xf[c[:, 0]] - xf[c[:, 1]], axis=1)
    assert np.max(np.abs(d - ff.constraint_r) / ff.constraint_r) < 1e-4
    assert np.isfinite(integ.computeKineticEnergy())
    assert integ._last["flags"] == 0 and integ._last["step"] == 4 * nprnt
