"""The C++ host API (SDMPlugin.h: LangevinIntegratorSDM, SDMForce, Context) driven from a
compiled C++ program gives the same observables as the C ABI called directly, and keeps
the reference's error behaviour."""
import dataclasses
import json
import os
import subprocess

import numpy as np
import pytest

from openmm_sdm_plugin_amd import _native, systems

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "facade_driver")


def write_system(s, path):
    with open(path, "w") as f:
        hdr = [float(s.cutoff), float(s.eps_rf)] + [float(b) for b in s.box]
        f.write(f"{s.n} " + " ".join(repr(v) for v in hdr) + "\n")
        for i in range(s.n):
            vals = [s.charge[i], s.sigma[i], s.epsilon[i], s.mass[i], *s.displacement[i],
                    *s.positions[i], *s.velocities[i]]
            f.write(" ".join(repr(float(v)) for v in vals) + "\n")
        f.write(f"{len(s.exclusions)}\n")
        for a, b in s.exclusions:
            f.write(f"{a} {b}\n")
        f.write(f"{len(s.bonds)}\n")
        for (a, b), r0, k in zip(s.bonds, s.bond_r0, s.bond_k):
            f.write(f"{a} {b} {float(r0)!r} {float(k)!r}\n")
        f.write(f"{len(s.angles)}\n")
        for (a, b, c), t0, k in zip(s.angles, s.angle_t0, s.angle_k):
            f.write(f"{a} {b} {c} {float(t0)!r} {float(k)!r}\n")
        vs = s.virtual_sites
        if s.lj_rule or vs is not None:
            nvs = 0 if vs is None else len(vs.kind)
            f.write(f"{int(s.lj_rule)} {nvs}\n")
            for k in range(nvs):
                a, w = vs.atoms[k], vs.weights[k]
                f.write(f"{int(vs.kind[k])} {a[0]} {a[1]} {a[2]} {a[3]} "
                        + " ".join(repr(float(v)) for v in w) + "\n")


def run_driver(s, tmp_path, nsteps=20):
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/facade_driver was not built (__graft_entry__.build())")
    path = tmp_path / "sys.txt"
    write_system(s, str(path))
    out = subprocess.run([DRIVER, str(path), str(nsteps)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    return json.loads(out.stdout.strip().splitlines()[-1])


def check_against_c_abi(s, res):
    """the façade's observables against the C ABI on the same system and parameters"""
    assert res["bound_error"] == "This Integrator is already bound to a context"
    assert res["softcore_error"] == "Unknown soft core method"
    p = systems.params_config4()  # the parameters facade_driver.cpp sets
    p.lambda_ = 1.0
    p.seed = 4242
    ctx = _native.NativeContext(s, 1)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.compute()
    e = ctx.energies()[0]
    assert res["force_uprime"] == e["bind_e"]
    assert res["force_U0"] == e["U0"]
    F = ctx.forces(0)
    assert np.isclose(res["force_checksum"], float((F[:, 0] + 2 * F[:, 1] + 3 * F[:, 2]).sum()),
                      rtol=1e-12, atol=1e-6)
    ctx.set_velocities(0, s.velocities)
    ctx.step(20)
    e = ctx.energies()[0]
    assert res["bind_e"] == e["bind_e"] and res["pot_energy"] == e["pot_energy"]
    assert np.isclose(res["x_checksum"], float(ctx.get_positions(0).sum()), rtol=1e-13)
    assert np.isclose(res["time"], 0.02, rtol=1e-12)
    assert np.isclose(res["ke"], float(ctx.kinetic_energy()[0]), rtol=1e-12)


def test_cpp_facade_matches_c_abi(hip_device, tmp_path):
    s = systems.make_config1(frame=5)
    print("case facade config1 frame 5", flush=True)
    check_against_c_abi(s, run_driver(s, tmp_path))


def test_cpp_facade_virtual_sites_and_geometric_rule(hip_device, tmp_path):
    """The four-site water box (TIP4P-style charge sites, two guest sites) with the OPLS
    combining rule, given to the façade as OpenMM-shaped VirtualSite objects and
    System::setCombiningRule."""
    from vsite_cases import four_site
    s = dataclasses.replace(four_site(frame=3), lj_rule=1)
    print("case facade 4-site small frame 3 geometric", flush=True)
    res = run_driver(s, tmp_path)
    check_against_c_abi(s, res)
    # the rule and the sites both reach the device: the arithmetic rule gives another
    # energy, and without the site definitions the massless sites stay where they started
    p = systems.params_config4()
    p.lambda_ = 1.0
    p.seed = 4242
    ctx = _native.NativeContext(dataclasses.replace(s, lj_rule=0), 1)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.compute()
    assert ctx.energies()[0]["U0"] != res["force_U0"]
    ctx = _native.NativeContext(dataclasses.replace(s, virtual_sites=None), 1)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.set_velocities(0, s.velocities)
    ctx.step(20)
    assert not np.isclose(res["x_checksum"], float(ctx.get_positions(0).sum()), rtol=1e-9)
