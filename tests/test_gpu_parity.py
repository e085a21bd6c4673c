"""GPU parity of the HIP SDM path against the FP64 oracle (called through the C ABI).

Tolerances (mixed FP32-compute / fixed-point mode):
  U0, U1, u, Ub, W, u'   within 1e-5 relative
  forces                 within 1e-4 RMS relative (per force component array)
"""
import numpy as np
import pytest

from openmm_sdm_plugin_amd import _native, systems
from oracle import oracle as O

pytestmark = pytest.mark.gpu

E_TOL = 1e-5
F_TOL = 1e-4

# pair kernels: block waves with / without Newton-3 (with Newton-3 also 4 and 16 waves
# per i-block), grouped waves of 16 and 8 atoms
KERNELS = {"block_n3": dict(pair_group=0, pair_newton3=1, pair_split=1),
           "block_n3_split4": dict(pair_group=0, pair_newton3=1, pair_split=4),
           "block_n3_split16": dict(pair_group=0, pair_newton3=1, pair_split=16),
           "block_full": dict(pair_group=0, pair_newton3=0),
           "grp16": dict(pair_group=16, pair_newton3=1),
           "grp8": dict(pair_group=8, pair_newton3=1)}


def rms_rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))


def close(a, b, tol=E_TOL, floor=1.0):
    return abs(a - b) <= tol * max(floor, abs(b))


def gpu_components(s, params=None, displacement=None, kernel="grp16"):
    ctx = _native.NativeContext(s, 1, displacement=displacement)
    ctx.set_tuning(**KERNELS[kernel])
    ctx.set_params(0, params or systems.params_config1())
    ctx.set_positions(0, s.positions)
    ctx.set_velocities(0, s.velocities)
    ctx.compute_components()
    e = ctx.energies()[0]
    F = {w: ctx.forces(0, w) for w in range(4)}
    return ctx, e, F


def check_components(s, params=None, kernel="grp16"):
    print(f"case {s.name} n={s.n} kernel={kernel}", flush=True)
    ctx, e, F = gpu_components(s, params, kernel=kernel)
    topo = O.Topology(s)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, s.positions)
    U0, u, F0d, dF, _, _ = O.sdm_direct(topo, s.positions)
    assert abs(e["U0"] - E1) <= E_TOL * max(1.0, abs(E1)), (e["U0"], E1)
    assert abs(e["U1"] - E2) <= E_TOL * max(abs(E1), abs(E2)), (e["U1"], E2)
    assert close(e["u"], u), (e["u"], u, E2 - E1)
    assert abs(e["Ub"] - Eb) <= E_TOL * max(1.0, abs(Eb)), (e["Ub"], Eb)
    assert rms_rel(F[_native.FORCE_STATE1], F1) < F_TOL
    assert rms_rel(F[_native.FORCE_STATE2], F2) < F_TOL
    assert rms_rel(F[_native.FORCE_BONDED], Fb) < F_TOL
    p = O.make_params(params or systems.params_config1())
    sc = O.execute_scalars(p, E1, E2, Eb, 0.0)
    Fh = np.empty_like(F1)
    O.lib().orc_force_mix(s.n, sc.sp, F1.ctypes.data_as(O._dp), F2.ctypes.data_as(O._dp),
                          Fb.ctypes.data_as(O._dp), Fh.ctypes.data_as(O._dp))
    assert rms_rel(F[_native.FORCE_HYBRID], Fh) < F_TOL
    assert close(e["bind_e"], sc.bind_e), (e["bind_e"], sc.bind_e)
    assert close(e["W"], sc.ebias), (e["W"], sc.ebias)
    scale = abs(E1) + abs(sc.ebias) + abs(Eb)  # PotEnergy = U0 + W + Ub
    assert abs(e["pot_energy"] - sc.pot_energy) <= E_TOL * scale, (e["pot_energy"], sc.pot_energy)
    # the normal (non-component) path gives the same hybrid force
    ctx.compute()
    Fn = ctx.forces(0, _native.FORCE_HYBRID)
    assert rms_rel(Fn, Fh) < F_TOL
    return ctx


@pytest.mark.parametrize("kernel", ["block_n3_split4", "block_full"])
def test_exclusions_across_the_displaced_set(hip_device, kernel):
    """Exclusions between displaced and undisplaced atoms (a guest bonded to its host):
    the first two guest atoms each get exclusions with their three nearest water atoms
    (0.24-0.37 nm away, inside the cutoff), so that the displaced-state screen has to
    leave those pairs out at both geometries and the pair kernels at x."""
    import dataclasses
    s = systems.make_small(frame=1)
    x = np.asarray(s.positions)
    lig = np.nonzero(np.any(np.asarray(s.displacement) != 0.0, axis=1))[0]
    env = np.setdiff1d(np.arange(s.n), lig)
    box = np.asarray(s.box)
    extra = []
    for a in lig[:2]:
        d = x[env] - x[a]
        d -= box * np.round(d / box)
        near = env[np.argsort((d * d).sum(1))[:3]]
        extra += [(int(a), int(b)) for b in near]
    ex = np.concatenate([np.asarray(s.exclusions).reshape(-1, 2), np.asarray(extra)], axis=0)
    s2 = dataclasses.replace(s, exclusions=ex.astype(np.int32), name="small-cross-excl")
    check_components(s2, kernel=kernel)
    # the exclusions change the result: without them u differs
    _, e_ref, _ = gpu_components(s, kernel=kernel)
    _, e_x, _ = gpu_components(s2, kernel=kernel)
    assert abs(e_x["u"] - e_ref["u"]) > 1e-6 * max(1.0, abs(e_ref["u"]))


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("frame", [0, 1, 2])
def test_small_components(hip_device, frame, kernel):
    s = systems.make_small(frame=frame)
    check_components(s, kernel=kernel)


@pytest.mark.parametrize("kernel", ["block_n3", "block_n3_split4", "block_n3_split16"])
@pytest.mark.parametrize("frame", [0, 1, 2])
def test_geometric_rule_components(hip_device, frame, kernel):
    """The OPLS rule (sigma_ij = sqrt(s_i s_j), the reference's CustomNonbondedForce for
    OPLS, desmonddmsfile75.py:781) on the block Newton-3 kernel, against the oracle."""
    s = systems.make_small(frame=frame)
    s.lj_rule = 1
    check_components(s, kernel=kernel)
    s = systems.make_config1(frame=frame)
    s.lj_rule = 1
    check_components(s, kernel=kernel)


def test_geometric_rule_other_kernels_refused(hip_device):
    s = systems.make_small()
    s.lj_rule = 1
    ctx = _native.NativeContext(s, 1)
    for kernel in ("block_full", "grp16", "grp8"):
        with pytest.raises(_native.SDMError):
            ctx.set_tuning(**KERNELS[kernel])
    assert ctx.autotune(5)["pair_newton3"] == 1 and ctx.get_tuning()["pair_group"] == 0


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("frame", [0, 1])
def test_config1_components(hip_device, frame, kernel):
    check_components(systems.make_config1(frame=frame), kernel=kernel)


@pytest.mark.parametrize("kernel", list(KERNELS))
def test_config2_components(hip_device, kernel):
    check_components(systems.make_config2(frame=0), systems.params_config2(), kernel=kernel)


@pytest.mark.parametrize("kernel", list(KERNELS))
def test_config5_components(hip_device, kernel):
    s = systems.make_config5(frame=0)
    check_components(s, systems.params_config5()[3], kernel=kernel)


@pytest.mark.parametrize("bias", [0, 1, 2])
@pytest.mark.parametrize("softcore", [0, 1, 2])
def test_scalar_stage_methods(hip_device, bias, softcore):
    s = systems.make_small(frame=1)
    topo = O.Topology(s)
    E1, E2, Eb, *_ = O.reference_eval(topo, s.positions)
    p = systems.params_config1()
    p.bias_method, p.softcore_method = bias, softcore
    p.lambda_, p.lambda1, p.lambda2, p.alpha, p.u0, p.w0coeff = 0.35, 0.2, 0.6, 0.05, 40.0, 3.0
    p.gamma, p.wbcoeff = 0.01, 0.4
    p.umax, p.acore, p.ubcore = 50.0, 0.0625, 10.0
    ctx = _native.NativeContext(s, 1)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.compute()
    e = ctx.energies()[0]
    sc = O.execute_scalars(O.make_params(p), E1, E2, Eb, 0.0)
    for k, ref in (("bind_e", sc.bind_e), ("W", sc.ebias), ("Wp", sc.bfp), ("fp", sc.fp),
                   ("sp", sc.sp)):
        assert close(e[k], ref, floor=1e-3), (k, e[k], ref)
    # PotEnergy = U0 + W + Ub: its bound is the sum of the components' 1e-5 bounds
    scale = abs(E1) + abs(sc.ebias) + abs(Eb)
    assert abs(e["pot_energy"] - sc.pot_energy) <= E_TOL * scale, (e["pot_energy"], sc.pot_energy)


def test_step_parity_with_noise(hip_device):
    s = systems.make_small(frame=2)
    p = systems.params_config1()
    p.alpha, p.lambda1, p.lambda2, p.u0 = 0.1, 0.2, 0.4, 30.0
    p.seed = 424242
    ctx = _native.NativeContext(s, 1)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.set_velocities(0, s.velocities)
    nsteps = 5
    ctx.step(nsteps)
    e = ctx.energies()[0]
    x_gpu = ctx.get_positions(0)
    v_gpu = ctx.get_velocities(0)
    topo = O.Topology(s)
    x, v, t, st, rep = O.ref_step(topo, O.make_params(p), s.positions, s.velocities, 0.0, 0,
                                  p.seed, 0, nsteps, use_cl_langevin=True)
    assert st == nsteps and e["step"] == nsteps
    assert abs(e["time"] - t) < 1e-12
    assert np.max(np.abs(x_gpu - x)) < 1e-5
    assert rms_rel(v_gpu, v) < 1e-3
    assert close(e["bind_e"], rep.s.bind_e)
    assert close(e["pot_energy"], rep.s.pot_energy)


@pytest.mark.parametrize("m_alpha", [0.0, 0.04])
def test_nonequilibrium_schedule(hip_device, m_alpha):
    """m_alpha = 0 is the reference's switching; 0.04 also schedules alpha, whose work term
    uses dW/dalpha (pinned by test_oracle_pins.test_work_alpha_slope_is_dW_dalpha)."""
    s = systems.make_small(frame=0)
    p = systems.params_config1()
    p.bias_method = 2
    p.alpha, p.m_alpha = 0.1, m_alpha
    p.nonequilibrium, p.noneq_tmax = 1, 0.004
    p.m_l1, p.m_l2, p.m_u0, p.m_w0 = 0.5, 0.8, 10.0, 1.0
    p.b_l1, p.b_l2, p.b_u0, p.b_w0 = 0.1, 0.1, 20.0, 0.0
    p.friction, p.temperature = 1.0, 0.0
    ctx = _native.NativeContext(s, 1)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.set_velocities(0, s.velocities)
    ctx.step(3)
    e = ctx.energies()[0]
    got = ctx.get_params(0, systems.SDMParams())
    topo = O.Topology(s)
    op = O.make_params(p)
    x, v, t, st, rep = O.ref_step(topo, op, s.positions, s.velocities, 0.0, 0, p.seed, 0, 3)
    for k in ("lambda_", "lambda1", "lambda2", "u0", "w0coeff"):
        assert abs(getattr(got, k) - getattr(op, k)) < 1e-12, k
    assert close(got.work, op.work, tol=1e-5, floor=1e-6), (got.work, op.work)
    assert close(e["pot_energy"], rep.s.pot_energy)


@pytest.mark.parametrize("kernel", list(KERNELS))
def test_determinism_bitwise(hip_device, kernel):
    s = systems.make_config1(frame=3)
    ctx = _native.NativeContext(s, 1)
    ctx.set_tuning(**KERNELS[kernel])
    ctx.set_params(0, systems.params_config1())
    ctx.set_positions(0, s.positions)
    ctx.compute()
    e1 = ctx.energies()[0]
    f1 = ctx.forces(0)
    for _ in range(3):
        ctx.compute()
        e2 = ctx.energies()[0]
        f2 = ctx.forces(0)
        assert e1["U0"] == e2["U0"] and e1["u"] == e2["u"] and e1["Ub"] == e2["Ub"]
        assert np.array_equal(f1, f2)


def test_batched_equals_single(hip_device):
    frames = [systems.make_config1(frame=k) for k in range(3)]
    params = []
    for k in range(3):
        p = systems.params_config1()
        p.lambda1 = p.lambda2 = 0.1 + 0.2 * k
        p.alpha = 0.05 * k
        params.append(p)
    multi = _native.NativeContext(frames[0], 3)
    for k in range(3):
        multi.set_params(k, params[k])
        multi.set_positions(k, frames[k].positions)
    multi.compute()
    em = multi.energies()
    for k in range(3):
        single = _native.NativeContext(frames[0], 1)
        single.set_params(0, params[k])
        single.set_positions(0, frames[k].positions)
        single.compute()
        es = single.energies()[0]
        for key in ("U0", "u", "Ub", "bind_e", "W", "sp", "pot_energy"):
            assert em[k][key] == es[key], (k, key)
        assert np.array_equal(multi.forces(k), single.forces(0))


@pytest.mark.parametrize("kernel", list(KERNELS))
def test_pair_counts(hip_device, kernel):
    s = systems.make_config1(frame=0)
    ctx = _native.NativeContext(s, 1)
    ctx.set_tuning(**KERNELS[kernel])
    ctx.set_params(0, systems.params_config1())
    ctx.set_positions(0, s.positions)
    p0, pd = ctx.count_pairs()
    topo = O.Topology(s)
    _, _, _, _, n0, nd = O.sdm_direct(topo, s.positions)
    # FP32 r^2 < rc^2 may admit/drop pairs within rounding of the cutoff (energy ~0 there)
    assert abs(int(p0[0]) - n0) <= max(2, 1e-5 * n0), (int(p0[0]), n0)
    assert int(pd[0]) == nd  # the displaced-state pairs are tested in FP64


def test_box_vector_displacement_gives_zero_u(hip_device):
    s = systems.make_config2(frame=1)
    d = np.zeros((s.n, 3))
    d[s.ligand] = (s.box[0], 0.0, 0.0)
    ctx, e, F = gpu_components(s, systems.params_config2(), displacement=d)
    assert abs(e["u"]) < 1e-6
    assert np.max(np.abs(F[_native.FORCE_STATE2] - F[_native.FORCE_STATE1])) < 1e-6


def test_no_displacement(hip_device):
    s = systems.make_small(frame=0)
    ctx, e, F = gpu_components(s, displacement=np.zeros((s.n, 3)))
    assert e["u"] == 0.0
    assert np.array_equal(F[_native.FORCE_STATE2], F[_native.FORCE_STATE1])


def _translated_run(s, kernel, interval, nsteps, drift=30.0):
    """nsteps of deterministic dynamics (no friction, no noise) with every atom given an
    extra velocity of `drift` nm/ps along x, so that ~0.03 nm per step of the system
    crosses the x faces of the box: blocks of the kept slot order straddle a face."""
    print(f"translated run {s.name} kernel={kernel} interval={interval} steps={nsteps}",
          flush=True)
    p = systems.params_config1()
    p.friction, p.temperature = 0.0, 0.0
    v = np.array(s.velocities, dtype=np.float64)
    v[:, 0] += drift
    ctx = _native.NativeContext(s, 1)
    ctx.set_tuning(**KERNELS[kernel], sort_interval=interval)
    ctx.set_params(0, p)
    ctx.set_positions(0, s.positions)
    ctx.set_velocities(0, v)
    ctx.step(nsteps)
    return ctx, p


@pytest.mark.parametrize("kernel", ["block_n3", "block_n3_split4", "block_full", "grp16"])
def test_components_between_sorts(hip_device, kernel):
    """Energies and forces evaluated in a kept slot order (sorted 10 steps and ~0.3 nm
    of translation earlier, atoms wrapped across the x faces since) match the oracle at
    the current positions."""
    s = systems.make_config1(frame=1)
    ctx, p = _translated_run(s, kernel, interval=50, nsteps=10)
    x = ctx.get_positions(0)
    L = np.asarray(s.box)
    crossed = np.any(np.floor(x / L) != np.floor(np.asarray(s.positions) / L), axis=1)
    assert crossed.sum() > 50, int(crossed.sum())
    ctx.compute_components()
    e = ctx.energies()[0]
    assert e["flags"] == 0
    topo = O.Topology(s)
    E1, E2, Eb, F1, F2, Fb = O.reference_eval(topo, x)
    U0, u, F0d, dF, _, _ = O.sdm_direct(topo, x)
    assert close(e["U0"], E1), (e["U0"], E1)
    assert abs(e["U1"] - E2) <= E_TOL * max(abs(E1), abs(E2)), (e["U1"], E2)
    assert close(e["u"], u), (e["u"], u)
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE1), F1) < F_TOL
    assert rms_rel(ctx.forces(0, _native.FORCE_STATE2), F2) < F_TOL
    assert rms_rel(ctx.forces(0, _native.FORCE_BONDED), Fb) < F_TOL


def test_sort_interval_trajectory(hip_device):
    """Sorting every step or every 7th step gives the same trajectory up to the FP32
    summation order of the pair forces; a kept order is bit-reproducible."""
    s = systems.make_config1(frame=2)
    a, _ = _translated_run(s, "block_n3", interval=1, nsteps=20)
    b, _ = _translated_run(s, "block_n3", interval=7, nsteps=20)
    c, _ = _translated_run(s, "block_n3", interval=7, nsteps=20)
    xa, xb, xc = a.get_positions(0), b.get_positions(0), c.get_positions(0)
    assert np.max(np.abs(xa - xb)) < 1e-6
    assert np.array_equal(xb, xc)
    ea, eb, ec = a.energies()[0], b.energies()[0], c.energies()[0]
    assert close(eb["pot_energy"], ea["pot_energy"]), (eb["pot_energy"], ea["pot_energy"])
    assert eb["pot_energy"] == ec["pot_energy"] and eb["u"] == ec["u"]
    assert ea["flags"] == eb["flags"] == 0


@pytest.mark.parametrize("use_graph", [0, 1])
def test_steps_agree_across_contexts(hip_device, use_graph):
    """The same 9 steps in many contexts of one process, created one after another with
    0-4 others alive (so the process's streams take many spreads over the hardware
    queues), give bit-identical trajectories: every launch keeps the step's data
    dependencies whatever the stream mapping."""
    s = systems.make_config1(frame=0)
    ref = None
    for k, nidle in enumerate((0, 2, 4, 1, 3, 0, 4, 2, 1, 3, 0, 2)):
        idle = [_translated_run(s, "block_n3", interval=4, nsteps=0)[0] for _ in range(nidle)]
        ctx, _ = _translated_run(s, "block_n3", interval=4, nsteps=0)
        ctx.set_tuning(use_graph=use_graph)
        ctx.step(9)
        x = ctx.get_positions(0)
        assert ctx.energies()[0]["flags"] == 0
        if ref is None:
            ref = x
        assert np.array_equal(x, ref), (k, nidle, float(np.max(np.abs(x - ref))))
        del ctx, idle


def test_energy_calls_keep_step_order(hip_device):
    """Energy evaluations between steps do not change the slot orders the steps use."""
    s = systems.make_config1(frame=0)
    a, _ = _translated_run(s, "block_n3", interval=4, nsteps=0)
    b, _ = _translated_run(s, "block_n3", interval=4, nsteps=0)
    a.step(9)
    b.compute()
    for _ in range(3):
        b.step(3)
        b.compute_components()
    assert np.array_equal(a.get_positions(0), b.get_positions(0))
    a.compute()
    b.compute()
    ea, eb = a.energies()[0], b.energies()[0]
    assert ea["U0"] == eb["U0"] and ea["u"] == eb["u"] and ea["pot_energy"] == eb["pot_energy"]
    assert np.array_equal(a.forces(0), b.forces(0))


def test_energy_on_last_step_of_call(hip_device):
    """Pair energies are evaluated on the last step of each sdm_step call (the record
    that energies() returns); the dynamics do not depend on them, so the trajectory and
    the final energies equal those of a run that evaluates them every step."""
    s = systems.make_config1(frame=1)
    runs = []
    for every in (0, 1):
        ctx = _native.NativeContext(s, 1)
        ctx.set_tuning(**KERNELS["block_n3"], energy_every_step=every)
        p = systems.params_config1()
        p.seed = 77
        ctx.set_params(0, p)
        ctx.set_positions(0, s.positions)
        ctx.set_velocities(0, s.velocities)
        ctx.step(7)
        e7 = ctx.energies()[0]
        ctx.step(1)
        e8 = ctx.energies()[0]
        runs.append((ctx.get_positions(0), e7, e8))
    (xa, a7, a8), (xb, b7, b8) = runs
    assert np.max(np.abs(xa - xb)) < 1e-9
    for ea, eb in ((a7, b7), (a8, b8)):
        for k in ("U0", "U1", "u", "pot_energy", "W", "sp"):
            assert close(ea[k], eb[k], tol=1e-6), (k, ea[k], eb[k])
        assert ea["flags"] == eb["flags"] == 0
    assert a7["U0"] != a8["U0"]  # the single-step call evaluated its own energies


@pytest.mark.parametrize("use_graph,split", [(0, 1), (0, 4), (1, 8), (1, 16), (0, 32)])
def test_forces_only_steps_match_energy_steps_on_ladder(hip_device, use_graph, split):
    """The forces-only pair kernel that steps use between energy evaluations gives the
    forces of the energy variant bit for bit on a full-size batch (22 replicas of the
    22,750-atom system, where every wave of the pair kernel shares its SIMD with
    others): 10 steps with energies on the last step only equal 10 steps with energies
    on every step."""
    R = 22
    s = systems.make_config2(frame=0)
    params = systems.params_config3(R)
    print(f"case ladder R={R} use_graph={use_graph} split={split}", flush=True)
    xs = []
    for every in (0, 1):
        ctx = _native.NativeContext(s, R)
        for r in range(R):
            ctx.set_params(r, params[r])
            ctx.set_positions(r, s.positions)
            ctx.set_velocities(r, s.velocities)
        ctx.set_tuning(**{**KERNELS["block_n3"], "pair_split": split}, use_graph=use_graph,
                       sort_interval=20, energy_every_step=every)
        ctx.step(10)
        assert all(e["flags"] == 0 for e in ctx.energies())
        xs.append(np.stack([ctx.get_positions(r) for r in range(R)]))
        del ctx
    assert np.array_equal(xs[0], xs[1]), float(np.max(np.abs(xs[0] - xs[1])))


@pytest.mark.parametrize("field,value", [("pair_newton3", 2), ("use_graph", 2),
                                         ("pair_group", 4), ("pair_waves_per_group", 3),
                                         ("sort_interval", 0), ("energy_every_step", -1),
                                         ("pair_split", 3), ("side_fork", 5),
                                         ("list_skin", 0.2), ("list_skin", -0.01)])
def test_tuning_out_of_range_refused(hip_device, field, value):
    """sdm_set_tuning refuses values outside each field's documented set and keeps the
    previous tuning."""
    s = systems.make_small()
    ctx = _native.NativeContext(s, 1)
    before = ctx.get_tuning()
    with pytest.raises(_native.SDMError):
        ctx.set_tuning(**{field: value})
    assert ctx.get_tuning() == before


@pytest.mark.parametrize("use_graph", [0, 1])
def test_side_work_placement_gives_same_bits(hip_device, use_graph):
    """The bonded terms and the displaced-state chain on side streams forked after the
    tile lists, forked before them (joined after the pair kernel or before it), or
    serially on the main stream: the same trajectory bit for bit (fixed-point force sums
    commute; every FP64 partial has one writer)."""
    frames = [systems.make_config1(frame=k) for k in range(3)]
    params = systems.params_config3(3)
    xs = []
    for fork in (0, 1, 2, 3, 4):
        print(f"case side_fork={fork} use_graph={use_graph}", flush=True)
        ctx = _native.NativeContext(frames[0], 3)
        ctx.set_tuning(side_fork=fork, use_graph=use_graph, pair_split=4)
        for r in range(3):
            ctx.set_params(r, params[r])
            ctx.set_positions(r, frames[r].positions)
            ctx.set_velocities(r, frames[r].velocities)
        ctx.step(12)
        assert all(e["flags"] == 0 for e in ctx.energies())
        xs.append((np.stack([ctx.get_positions(r) for r in range(3)]),
                   [e["pot_energy"] for e in ctx.energies()]))
    for x, pe in xs[1:]:
        assert np.array_equal(x, xs[0][0]) and pe == xs[0][1]



def _skin_run(frames, params, tuning, skin, n1=16, n2=5):
    """n1 steps, positions set again (replica r takes frame r + 1), n2 more steps."""
    R = len(frames)
    ctx = _native.NativeContext(frames[0], R)
    ctx.set_tuning(**tuning, list_skin=skin)
    for r in range(R):
        ctx.set_params(r, params[r])
        ctx.set_positions(r, frames[r].positions)
        ctx.set_velocities(r, frames[r].velocities)
    ctx.step(n1)
    x1 = np.stack([ctx.get_positions(r) for r in range(R)])
    for r in range(R):
        ctx.set_positions(r, frames[(r + 1) % R].positions)
    ctx.step(n2)
    es = ctx.energies()
    assert all(e["flags"] == 0 for e in es)
    assert ctx.get_tuning()["list_skin"] == pytest.approx(skin)
    return x1, np.stack([ctx.get_positions(r) for r in range(R)]), [e["pot_energy"] for e in es]


@pytest.mark.parametrize("use_graph", [0, 1])
@pytest.mark.parametrize("kernel", ["block_n3", "block_full", "grp16"])
def test_list_skin_gives_same_bits(hip_device, kernel, use_graph):
    """Tile lists kept over steps with a Verlet skin (built for r_c + skin, kept while no
    atom of the replica has moved half the skin): the pair kernel still applies the exact
    cutoff, the pairs added by the skin contribute exact zeros, and with one wave per
    i-block the i-force sums see the same nonzero terms in the same order, so the
    trajectory is the same bits as with lists built on every step (the FP64 pair energy
    agrees to rounding: its kernel's four waves per i-block take other shares of the
    longer tile lists).  Skins of 0.02 nm
    (lists kept about one step at 300 K and 1 fs) and 0.1 nm (several steps), full sorts
    every 7 steps, positions set again in the middle (the lists are rebuilt from the new
    coordinates).  The grouped kernel ignores the skin (its lists are built every time)."""
    frames = [systems.make_config1(frame=k) for k in range(3)]
    params = systems.params_config3(3)
    tuning = dict(KERNELS[kernel], sort_interval=7, use_graph=use_graph)
    runs = []
    for skin in (0.0, 0.02, 0.1):
        print(f"case kernel={kernel} use_graph={use_graph} list_skin={skin}", flush=True)
        runs.append(_skin_run(frames, params, tuning, skin))
    for x1, x2, pe in runs[1:]:
        assert np.array_equal(x1, runs[0][0])
        assert np.array_equal(x2, runs[0][1])
        for a, b in zip(pe, runs[0][2]):
            assert abs(a - b) <= 1e-13 * max(1.0, abs(b)), (a, b)


def test_list_skin_with_split_waves(hip_device):
    """With several waves per i-block the skin's extra tiles change which wave takes
    which tile, so the i forces are rounded to fixed point in other groups: the
    trajectory agrees with lists built every step to within that rounding."""
    frames = [systems.make_config1(frame=k) for k in range(3)]
    params = systems.params_config3(3)
    tuning = dict(KERNELS["block_n3_split4"], sort_interval=7, use_graph=0)
    print("case block_n3_split4 list_skin 0 / 0.05", flush=True)
    a = _skin_run(frames, params, tuning, 0.0, n1=8, n2=4)
    b = _skin_run(frames, params, tuning, 0.05, n1=8, n2=4)
    for xa, xb in ((a[0], b[0]), (a[1], b[1])):
        d = float(np.max(np.abs(xa - xb)))
        print(f"max |dx| {d:.2e} nm", flush=True)
        assert d < 1e-7
    for pa, pb in zip(a[2], b[2]):
        assert abs(pa - pb) <= 1e-7 * max(1.0, abs(pa))
