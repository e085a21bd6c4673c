"""GPU parity of the SDMUtils restraints (group 1) against the oracle.

The restraint terms are FP64 on both sides; forces pass through the 2^-32 fixed-point
accumulator.  Bounds: group-1 energy 1e-9 relative; group-1 forces 1e-5 kJ/mol/nm per
component (fixed point, and the oracle's central differences at ~1e-9 relative)."""
import numpy as np
import pytest

from openmm_sdm_plugin_amd import SDMUtils, _native, systems
from oracle import oracle as O
from oracle import restraints as OR

pytestmark = pytest.mark.gpu


def _restrained(frame=1, kfcm=800.0, with_alignment=True):
    """config 1 (3,000 atoms) with a centroid restraint between the ligand and 40 nearby
    atoms, its angle and dihedral terms in windows that leave them active, and an
    alignment restraint between two triples of ligand atoms."""
    s = systems.make_config1(frame=frame)
    x = np.asarray(s.positions)
    lig = np.asarray(s.ligand)
    c = x[lig].mean(0)
    others = np.setdiff1d(np.arange(s.n), lig)
    near = others[np.argsort(np.linalg.norm(x[others] - c, axis=1))[:40]]
    rref = [int(a) for a in near[:3]]
    lref = [int(a) for a in lig[:3]]
    t = OR.angle(x[rref[0]], x[lref[0]], x[lref[1]])
    p1 = OR.dihedral(x[rref[1]], x[rref[0]], x[lref[0]], x[lref[1]])
    p2 = OR.dihedral(x[rref[0]], x[lref[0]], x[lref[1]], x[lref[2]])
    u = SDMUtils(s)
    u.addRestraintForce(lig_cm_particles=lig.tolist(), rcpt_cm_particles=near.tolist(),
                        kfcm=kfcm, tolcm=0.01, lig_ref_particles=lref,
                        rcpt_ref_particles=rref, angle_center=t + 0.5, kfangle=60.0,
                        angletol=0.2, dihedral1center=p1 - 0.7, kfdihedral1=40.0,
                        dihedral1tol=0.3, dihedral2center=p2 + 0.9, kfdihedral2=30.0,
                        dihedral2tol=0.25, offset=[0.05, -0.02, 0.01])
    if with_alignment:
        u.addAlignmentForce(liga_ref_particles=[int(a) for a in lig[3:6]],
                            ligb_ref_particles=[int(a) for a in lig[6:9]], kfdispl=200.0,
                            ktheta=30.0, kpsi=25.0, offset=[0.1, 0.0, -0.1])
    return s


def _oracle_group1(s, x):
    topo = O.Topology(s)
    eb, fb = O.bonded(topo, x)
    er, fr = OR.energy_forces(x, s.restraints, np.asarray(s.mass))
    return eb + er, fb + fr, er


def test_restraint_components(hip_device):
    s = _restrained()
    print(f"case restraints config1 n={s.n}", flush=True)
    ctx = _native.NativeContext(s, 1)
    ctx.set_params(0, systems.params_config1())
    ctx.set_positions(0, s.positions)
    ctx.compute_components()
    e = ctx.energies()[0]
    Eg1, Fg1, Er = _oracle_group1(s, s.positions)
    assert Er > 1.0  # every term active
    assert abs(e["Ub"] - Eg1) <= 1e-9 * abs(Eg1), (e["Ub"], Eg1)
    Fb = ctx.forces(0, _native.FORCE_BONDED)
    assert np.max(np.abs(Fb - Fg1)) < 1e-5
    # the hybrid force of the normal path carries the same group-1 forces
    ctx.compute()
    Fh = ctx.forces(0, _native.FORCE_HYBRID)
    topo = O.Topology(s)
    E1, E2, Eb, F1, F2, _ = O.reference_eval(topo, s.positions)
    sc = O.execute_scalars(O.make_params(systems.params_config1()), E1, E2, Eb + Er, 0.0)
    Fref = (1 - sc.sp) * F1 + sc.sp * F2 + Fg1
    rel = np.sqrt(np.mean((Fh - Fref) ** 2) / np.mean(Fref ** 2))
    assert rel < 1e-4
    assert abs(ctx.energies()[0]["pot_energy"] - sc.pot_energy) <= 1e-5 * (
        abs(E1) + abs(sc.ebias) + abs(Eb + Er))


def test_restraints_in_steps_and_removal(hip_device):
    s = _restrained(frame=2)
    ctx = _native.NativeContext(s, 2)
    for r in range(2):
        ctx.set_params(r, systems.params_config1())
        ctx.set_positions(r, s.positions)
        ctx.set_velocities(r, s.velocities)
    ctx.step(1)  # the step's record holds the energies at the starting positions
    Eg1, _, Er = _oracle_group1(s, s.positions)
    for e in ctx.energies():
        assert abs(e["Ub"] - Eg1) <= 1e-9 * abs(Eg1)
        assert e["flags"] == 0
    ctx.set_restraints([])
    x = ctx.get_positions(0)
    ctx.compute_components()
    topo = O.Topology(s)
    eb, fb = O.bonded(topo, x)
    assert abs(ctx.energies()[0]["Ub"] - eb) <= 1e-9 * abs(eb)
    assert np.max(np.abs(ctx.forces(0, _native.FORCE_BONDED) - fb)) < 1e-5


def test_invalid_restraints_rejected(hip_device):
    s = systems.make_small(frame=0)
    ctx = _native.NativeContext(s, 1)
    u = SDMUtils(s)
    u.addRestraintForce(lig_cm_particles=[0, 1], rcpt_cm_particles=[s.n + 5], kfcm=1.0)
    with pytest.raises(_native.SDMError):
        ctx.set_restraints(s.restraints)
    with pytest.raises(ValueError):
        u.addAlignmentForce(liga_ref_particles=[0, 1], ligb_ref_particles=[2, 3, 4])
    # an empty group adds nothing, as in the reference
    n = len(s.restraints)
    u.addRestraintForce(lig_cm_particles=[], rcpt_cm_particles=[1, 2], kfcm=1.0)
    assert len(s.restraints) == n
