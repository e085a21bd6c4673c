"""Virtual sites on the GPU against the oracle (oracle/vsites.py, pinned by
tests/test_oracle_vsites.py).

The system is the small water box made four-site: every water's charge moves from the
oxygen to a massless site M = (1 - 2a) O + a H1 + a H2 (ThreeParticleAverageSite, a =
0.128), and the displaced guest gets a two-particle site between its first two atoms and an
out-of-plane site on atoms 1-3 (displaced with the guest).  The reference places such sites
after every step (ReferenceStochasticDynamicsSDM.cpp:264); OpenMM moves their forces to the
parents before the update.  Tolerances as in the other parity tests: energies 1e-5
relative, forces 1e-4 RMS relative, a noise-free step's positions 1e-7 nm."""
import dataclasses
import types

import numpy as np
import pytest

from openmm_sdm_plugin_amd import _native, systems
from oracle import oracle as O
from oracle import vsites as OV
from vsite_cases import four_site

pytestmark = pytest.mark.gpu

def rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _context(s, p, **tuning):
    ctx = _native.NativeContext(s, 1)
    if tuning:
        ctx.set_tuning(**tuning)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.set_velocities(0, s.velocities)
    return ctx


def test_vsite_components_match_oracle(hip_device):
    s = four_site()
    vs = s.virtual_sites
    p = systems.params_config1()
    print(f"case 4-site small n={s.n} sites={len(vs.kind)}", flush=True)
    ctx = _context(s, p)
    ctx.compute_components()
    e = ctx.energies()[0]
    topo = O.Topology(s)
    x = np.asarray(s.positions)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, x)
    U0, u, *_ = O.sdm_direct(topo, x)
    assert abs(e["U0"] - E1) <= 1e-5 * max(1.0, abs(E1)), (e["U0"], E1)
    assert abs(e["U1"] - E2) <= 1e-5 * max(abs(E1), abs(E2))
    assert abs(e["u"] - u) <= 1e-5 * max(1.0, abs(u))
    sites = vs.atoms[:, 0]
    for which, ref in ((_native.FORCE_STATE1, F1), (_native.FORCE_STATE2, F2),
                       (_native.FORCE_BONDED, Fb)):
        got = ctx.forces(0, which)
        want = OV.distribute(ref, x, vs)
        assert np.all(got[sites] == 0.0)
        assert rms_rel(got, want) < 1e-4, which
    sc = O.execute_scalars(O.make_params(p), E1, E2, Eb, 0.0)
    Fh = OV.distribute((1 - sc.sp) * F1 + sc.sp * F2 + Fb, x, vs)
    assert rms_rel(ctx.forces(0, _native.FORCE_HYBRID), Fh) < 1e-4
    # the charge sites carry real force before the move: without the move the parents'
    # forces would be visibly off
    assert rms_rel(OV.distribute(F1, x, vs)[:-len(sites)], F1[:-len(sites)]) > 1e-2


@pytest.mark.parametrize("use_graph", [0, 1])
def test_vsite_step_matches_oracle(hip_device, use_graph):
    """One noise-free step: the hybrid force moved to the parents, the Langevin update of
    the massive atoms, then the sites placed from their parents."""
    s = four_site(frame=2)
    vs = s.virtual_sites
    p = systems.params_config1()
    p.temperature = 0.0
    print(f"case 4-site step use_graph={use_graph}", flush=True)
    ctx = _context(s, p, use_graph=use_graph)
    ctx.step(1)
    x_gpu, v_gpu = ctx.get_positions(0), ctx.get_velocities(0)
    topo = O.Topology(s)
    x0 = np.asarray(s.positions)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, x0)
    sc = O.execute_scalars(O.make_params(p), E1, E2, Eb, 0.0)
    F = OV.distribute((1 - sc.sp) * F1 + sc.sp * F2 + Fb, x0, vs)
    xu, vu = O.langevin_update(x0, s.velocities, F, s.mass, p.step_size, p.friction, 0.0,
                               np.zeros_like(x0), form="cl")
    xr = OV.positions(xu, vs)
    d = float(np.max(np.abs(x_gpu - xr)))
    print(f"max |dx| {d:.2e} nm", flush=True)
    assert d < 1e-7
    massive = np.asarray(s.mass) > 0
    assert rms_rel(v_gpu[massive], vu[massive]) < 1e-4
    assert ctx.energies()[0]["flags"] == 0


def test_vsites_follow_their_parents(hip_device):
    """20 steps at 300 K: every site sits where its parents put it, the sites carry no
    force, and graph replay and direct launches give the same bits."""
    s = four_site(frame=0)
    vs = s.virtual_sites
    p = systems.params_config1()
    runs = []
    for use_graph in (0, 1):
        print(f"case 4-site 20 steps use_graph={use_graph}", flush=True)
        ctx = _context(s, p, use_graph=use_graph)
        ctx.step(20)
        e = ctx.energies()[0]
        assert e["flags"] == 0 and np.isfinite(e["pot_energy"])
        x = ctx.get_positions(0)
        assert np.max(np.abs(x - OV.positions(x, vs))) < 1e-12
        assert np.all(ctx.forces(0, _native.FORCE_HYBRID)[vs.atoms[:, 0]] == 0.0)
        runs.append(x)
    assert np.array_equal(runs[0], runs[1])
    # the sites moved with their molecules
    assert np.max(np.abs(runs[0][vs.atoms[:, 0]] - np.asarray(s.positions)[vs.atoms[:, 0]])) > 1e-4


def test_vsite_definitions_refused(hip_device):
    s = four_site()
    vs = s.virtual_sites
    ctx = _native.NativeContext(dataclasses.replace(s, virtual_sites=None), 1)
    site, o = int(vs.atoms[0, 0]), int(vs.atoms[0, 1])

    def bad(kind, atoms, w=(0.5, 0.5, 0.0)):
        with pytest.raises(_native.SDMError):
            ctx.set_virtual_sites(types.SimpleNamespace(kind=[kind], atoms=[atoms],
                                                        weights=[w]))
    bad(7, (site, o, o + 1, o + 2))              # unknown kind
    bad(OV.AVERAGE2, (o, o + 1, o + 2, 0))       # the site has mass
    bad(OV.AVERAGE2, (site, site + 1, o, 0))     # a parent is a site (massless)
    bad(OV.AVERAGE2, (site, o, s.n, 0))          # index out of range
    g2 = int(vs.atoms[-2, 0])
    bad(OV.AVERAGE2, (g2, o, o + 1, 0))          # guest site, water parents
    ctx.set_virtual_sites(vs)
    ctx.set_virtual_sites(None)
