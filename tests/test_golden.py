"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle still reproduces every fixture (a regression guard on the checker).
GPU: the HIP path against the fixtures, within 1e-5 relative (energies) and 1e-4 RMS
relative (forces)."""
import glob
import os

import numpy as np
import pytest

from openmm_sdm_plugin_amd import _native, systems
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
MAKERS = {"small_frame0": lambda: systems.make_small(frame=0),
          "small_frame1": lambda: systems.make_small(frame=1),
          "small_frame2": lambda: systems.make_small(frame=2),
          "config1_frame0": lambda: systems.make_config1(frame=0)}
# the oracle fixtures of make_golden.py (temoa_g1_g4.npz is the DMS system's fixture)
FIXTURES = sorted(p for p in glob.glob(os.path.join(HERE, "golden", "*.npz"))
                  if os.path.basename(p)[:-4] in MAKERS)


def load(path):
    name = os.path.basename(path)[:-4]
    return name, MAKERS[name](), np.load(path)


def params_from_row(row):
    p = systems.params_config1()
    (p.bias_method, p.softcore_method) = int(row[0]), int(row[1])
    (p.lambda_, p.lambda1, p.lambda2, p.alpha, p.u0, p.w0coeff, p.gamma, p.wbcoeff, p.umax,
     p.acore, p.ubcore) = [float(v) for v in row[2:]]
    return p


def rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def test_fixtures_present():
    assert len(FIXTURES) == 4


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_oracle_reproduces_fixture(path):
    name, s, g = load(path)
    assert np.array_equal(s.positions, g["positions"])
    topo = O.Topology(s)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, s.positions)
    assert E1 == float(g["E1"]) and E2 == float(g["E2"]) and Eb == float(g["Eb"])
    assert np.array_equal(F1, g["F1"]) and np.array_equal(F2, g["F2"])
    for row, ref in zip(g["scalar_params"], g["scalars"]):
        sc = O.execute_scalars(O.make_params(params_from_row(row)), E1, E2, Eb, 0.0)
        assert np.allclose([sc.bind_e, sc.ebias, sc.bfp, sc.fp, sc.sp, sc.pot_energy], ref,
                           rtol=1e-14, atol=0)


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p))
def test_hip_matches_fixture(hip_device, path):
    name, s, g = load(path)
    print(f"case golden {name}", flush=True)
    E1, E2, Eb = float(g["E1"]), float(g["E2"]), float(g["Eb"])
    ctx = _native.NativeContext(s, 1)
    ctx.set_params(0, params_from_row(g["scalar_params"][0]))
    ctx.set_positions(0, g["positions"])
    ctx.compute_components()
    e = ctx.energies()[0]
    # the pair energies are FP64 sums (the forces FP32): plain 1e-5 relative, also on the
    # strongly cancelling 186-atom systems (U0 ~ 30 kJ/mol of |pair energies| ~ 4.5e4)
    assert abs(e["U0"] - E1) <= 1e-5 * max(1.0, abs(E1)), (e["U0"], E1)
    assert abs(e["u"] - (E2 - E1)) <= 1e-5 * max(1.0, abs(E2 - E1))
    assert abs(e["Ub"] - Eb) <= 1e-5 * max(1.0, abs(Eb))
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE1), g["F1"]) < 1e-4
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE2), g["F2"]) < 1e-4
    assert rms_rel(ctx.forces(0, _native.FORCE_BONDED), g["Fb"]) < 1e-4
    p0, pd = ctx.count_pairs()
    assert abs(int(p0[0]) - int(g["npairs0"])) <= 2 and int(pd[0]) == int(g["npairs_delta"])
    for row, ref in zip(g["scalar_params"], g["scalars"]):
        ctx.set_params(0, params_from_row(row))
        ctx.compute()
        e = ctx.energies()[0]
        for k, v in zip(("bind_e", "W", "Wp", "fp", "sp"), ref[:5]):
            assert abs(e[k] - v) <= 1e-5 * max(1e-3, abs(v)), (k, e[k], v)
        assert abs(e["pot_energy"] - ref[5]) <= 1e-5 * (abs(E1) + abs(ref[1]) + abs(Eb))
    # trajectory with Philox noise
    pstep = systems.params_config1()
    pstep.alpha, pstep.lambda1, pstep.lambda2, pstep.u0 = 0.1, 0.2, 0.4, 30.0
    pstep.seed = int(g["step_seed"])
    ctx.set_params(0, pstep)
    ctx.set_positions(0, g["positions"])
    ctx.set_velocities(0, g["velocities"])
    ctx.set_time(0, 0.0, 0)
    ctx.step(int(g["steps"]))
    assert np.max(np.abs(ctx.get_positions(0) - g["step_x"])) < 1e-5
    assert rms_rel(ctx.get_velocities(0), g["step_v"]) < 1e-3
