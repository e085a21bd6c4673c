"""The per-replica error flags (sdm_energies.flags) are raised by the conditions they name
and cleared by sdm_clear_flags.  Each case evaluates finite inputs once, and every loop
involved is bounded (the fixed-point conversion clamps out-of-range and NaN values; SHAKE
stops after a fixed number of iterations):
  SDM_FLAG_OVERFLOW    two atoms 0.002 nm apart: an FP32 pair force beyond the 2^31
                       kJ/mol/nm fixed-point range (the FP64 energy stays finite)
  SDM_FLAG_NONFINITE   two atoms at one position: an infinite pair energy
  SDM_FLAG_CAPACITY    a dense cluster in a large box, whose neighbour lists exceed the
                       capacity sized from the mean density
  SDM_FLAG_CONSTRAINT  a constraint triangle of lengths 0.1, 0.1 and 0.3 nm, which no
                       positions satisfy
OpenMM reports such states as exceptions after the fact (NaN coordinates, constraint
failure); here the step completes and the flag is sticky until cleared."""
import dataclasses
import types

import numpy as np
import pytest

from openmm_sdm_plugin_amd import _native, systems

pytestmark = pytest.mark.gpu

FLAG_CONSTRAINT = 8


def _waters(s):
    lig = set(int(a) for a in s.ligand)
    return [(o, h1, h2) for h1, o, h2 in np.asarray(s.angles).reshape(-1, 3).tolist()
            if o not in lig]


def _compute(s):
    ctx = _native.NativeContext(s, 1)
    ctx.set_params(0, systems.params_config1())
    ctx.set_positions(0, s.positions)
    ctx.compute()
    return ctx


@pytest.mark.parametrize("gap, want", [(0.002, _native.FLAG_OVERFLOW),
                                       (0.0, _native.FLAG_NONFINITE)])
def test_close_pair_flags(hip_device, gap, want):
    s = systems.make_small(frame=0)
    (o1, _, _), (o2, _, _) = _waters(s)[:2]
    x = np.array(s.positions)
    x[o2] = x[o1] + np.array([gap, 0.0, 0.0])
    print(f"case close pair gap={gap}", flush=True)
    ctx = _compute(s.with_coordinates(x, s.velocities))
    e = ctx.energies()[0]
    assert e["flags"] & want
    if gap > 0:
        # only the force is out of range: the FP64 energies and u' stay finite
        assert e["flags"] == _native.FLAG_OVERFLOW
        assert np.isfinite(e["U0"]) and e["U0"] > 1e20 and np.isfinite(e["bind_e"])
    else:
        assert not np.isfinite(e["U0"])
    ctx.clear_flags()
    assert ctx.energies()[0]["flags"] == 0


def test_neighbour_list_capacity_flag(hip_device):
    """38,400 parameter-free atoms in a 3 nm cube inside a 30 nm box: each 64-atom block
    has a few hundred neighbour blocks, the lists hold 128 (sized from the mean density)."""
    rng = np.random.default_rng(5)
    n = 64 * 600
    x = 13.5 + 3.0 * rng.random((n, 3))
    disp = np.zeros((n, 3))
    disp[:2] = (3.0, 0.0, 0.0)
    z = np.zeros(n)
    e2 = np.zeros((0, 2), dtype=np.int32)
    s = systems.SDMSystem(
        box=np.full(3, 30.0), cutoff=1.0, eps_rf=78.5, charge=z, sigma=z.copy(),
        epsilon=z.copy(), mass=np.full(n, 16.0), exclusions=e2, bonds=e2, bond_r0=z[:0],
        bond_k=z[:0], angles=np.zeros((0, 3), dtype=np.int32), angle_t0=z[:0],
        angle_k=z[:0], displacement=disp, positions=x, velocities=np.zeros((n, 3)),
        ligand=np.array([0, 1], dtype=np.int32), name="cluster")
    print("case capacity cluster n=38400 box 30 nm", flush=True)
    ctx = _compute(s)
    e = ctx.energies()[0]
    assert e["flags"] & _native.FLAG_CAPACITY
    assert e["U0"] == 0.0
    # the same atoms spread over the box fit their lists
    ok = _compute(dataclasses.replace(s, positions=30.0 * rng.random((n, 3))))
    assert ok.energies()[0]["flags"] == 0


def test_constraint_flag(hip_device):
    s = systems.make_small(frame=0)
    o, h1, h2 = _waters(s)[0]
    i4 = np.zeros((0, 4), dtype=np.int32)
    i3 = np.zeros((0, 3), dtype=np.int32)
    f0 = np.zeros(0)
    ff = types.SimpleNamespace(
        torsions=i4, torsion_n=np.zeros(0, dtype=np.int32), torsion_phase=f0, torsion_k=f0,
        impropers=i4, improper_t0=f0, improper_k=f0,
        posre_fb_atoms=np.zeros(0, dtype=np.int32), posre_fb_x0=i3.astype(float),
        posre_fb_hk=f0, posre_fb_tol=f0,
        posre_h_atoms=np.zeros(0, dtype=np.int32), posre_h_x0=i3.astype(float),
        posre_h_hk=i3.astype(float),
        exception_pairs=np.zeros((0, 2), dtype=np.int32), exception_qq=f0,
        exception_sigma=f0, exception_epsilon=f0,
        constraints=np.array([[o, h1], [o, h2], [h1, h2]], dtype=np.int32),
        constraint_r=np.array([0.1, 0.1, 0.3]))
    print("case constraint triangle 0.1/0.1/0.3", flush=True)
    ctx = _native.NativeContext(s, 1)
    ctx.set_forcefield_terms(ff)
    p = systems.params_config1()
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.set_velocities(0, s.velocities)
    ctx.step(1)
    assert ctx.energies()[0]["flags"] & FLAG_CONSTRAINT
    ctx.clear_flags()
    assert ctx.energies()[0]["flags"] == 0
