"""The batched lambda ladder, as it is benchmarked, against the FP64 oracle.

Config 3: one context of 22 replicas of the 22,750-atom system, every replica with its
own coordinates (make_config2(frame=k)) and its own ladder parameters, run with the
benchmark's launch configuration (block Newton-3 pair kernel, 4 waves per i-block,
direct launches, pair energies on the last step of each call only).  Config 5: 8
replicas of the 99,920-atom two-guest system.  Config 1: 20 frames in one context.
Each replica is compared with the oracle evaluated on its own coordinates and
parameters (LangevinIntegratorSDM.cpp:153-183, ReferenceSDMKernels.cpp:264-318).

Tolerances: U0, U1, u, Ub, u', W and the scalars within 1e-5 relative; forces within
1e-4 RMS relative; trajectories: the velocity change over the steps within 1e-4 RMS
relative (the force's share of it, so a 1e-4 force error would show) and positions within
1e-5 nm, with the oracle drawing the same Philox noise.  The position bound is set by
the FP32 coordinates of the pair kernel (an ulp is 4.8e-7 nm at 4-8 nm): close pairs'
forces carry ~1e-5 relative errors from them, which over 3 steps move a hydrogen by
~1e-6 nm (measured 1.2e-6 nm).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from openmm_sdm_plugin_amd import _native, systems
from oracle import oracle as O

pytestmark = pytest.mark.gpu

E_TOL = 1e-5
F_TOL = 1e-4
# the benchmark's launch configuration of the 22-replica ladder (what sdm_autotune picks
# on MI355X) and of the 8-replica large ladder
BENCH_TUNING = dict(pair_group=0, pair_newton3=1, pair_split=4, pair_waves_per_group=1,
                    cell_cells_per_dim=20, sort_interval=20, use_graph=0, energy_every_step=0)
LARGE_TUNING = dict(pair_group=0, pair_newton3=1, pair_split=2, pair_waves_per_group=1,
                    cell_cells_per_dim=33, sort_interval=20, use_graph=1, energy_every_step=0)
# ctypes releases the GIL, so the oracle's replicas run in parallel threads
POOL = 8


def rms_rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))


def close(a, b, tol=E_TOL, floor=1.0):
    return abs(a - b) <= tol * max(floor, abs(b))


_cache = {}


def config3_frames(R=22):
    if "c3" not in _cache:
        _cache["c3"] = [systems.make_config2(frame=k) for k in range(R)]
    return _cache["c3"]


def config5_frames(R=8):
    """Replica k: frame 0 translated by a seeded random vector (periodic) with every
    coordinate jittered by up to 5 pm, so every replica has its own coordinates, slot
    order and periodic images."""
    if "c5" not in _cache:
        s = systems.make_config5(frame=0)
        rng = np.random.default_rng(55)
        out = []
        for k in range(R):
            x = np.asarray(s.positions) + rng.uniform(0, 10.0, 3) * (k > 0)
            x = x + rng.uniform(-0.005, 0.005, x.shape)
            out.append(s.with_coordinates(x.astype(np.float32).astype(np.float64)))
        _cache["c5"] = out
    return _cache["c5"]


def oracle_components(frames, params):
    """Per replica: U0 = U(x), u = U(x+d) - U(x), F0, dF (oracle sdm_direct, pinned
    against the two-sweep form), Ub, Fb, and the scalar stage for its parameters."""
    topo = O.Topology(frames[0])

    def one(k):
        x = frames[k].positions
        U0, u, F0, dF, _, _ = O.sdm_direct(topo, x)
        Eb, Fb = O.bonded(topo, x)
        sc = O.execute_scalars(O.make_params(params[k]), U0, U0 + u, Eb, 0.0)
        return dict(U0=U0, u=u, F0=F0, dF=dF, Eb=Eb, Fb=Fb, sc=sc)
    with ThreadPoolExecutor(POOL) as ex:
        return list(ex.map(one, range(len(frames))))


def ladder(frames, params, tuning):
    R = len(frames)
    ctx = _native.NativeContext(frames[0], R)
    ctx.set_tuning(**tuning)
    for r in range(R):
        ctx.set_params(r, params[r])
        ctx.set_positions(r, frames[r].positions)
        ctx.set_velocities(r, frames[r].velocities)
    return ctx


def check_ladder_components(ctx, frames, params):
    ctx.compute_components()
    es = ctx.energies()
    ref = oracle_components(frames, params)
    worst = {}
    for r, (e, o) in enumerate(zip(es, ref)):
        sc = o["sc"]
        U0, u, Eb = o["U0"], o["u"], o["Eb"]
        assert e["flags"] == 0, (r, e["flags"])
        assert close(e["U0"], U0), (r, e["U0"], U0)
        assert abs(e["U1"] - (U0 + u)) <= E_TOL * max(abs(U0), abs(U0 + u)), (r, e["U1"])
        assert close(e["u"], u), (r, e["u"], u)
        assert close(e["Ub"], Eb), (r, e["Ub"], Eb)
        for k, want in (("bind_e", sc.bind_e), ("W", sc.ebias), ("Wp", sc.bfp),
                        ("sp", sc.sp)):
            assert close(e[k], want, floor=1e-3), (r, k, e[k], want)
        scale = abs(U0) + abs(sc.ebias) + abs(Eb)
        assert abs(e["pot_energy"] - sc.pot_energy) <= E_TOL * scale, (r, e["pot_energy"])
        F0, F1 = o["F0"], o["F0"] + o["dF"]
        Fh = F0 + sc.sp * o["dF"] + o["Fb"]
        got = {w: ctx.forces(r, w) for w in range(4)}
        for w, want in ((_native.FORCE_STATE1, F0), (_native.FORCE_STATE2, F1),
                        (_native.FORCE_BONDED, o["Fb"]), (_native.FORCE_HYBRID, Fh)):
            err = rms_rel(got[w], want)
            worst[w] = max(worst.get(w, 0.0), err)
            assert err < F_TOL, (r, w, err)
    print(f"ladder R={len(frames)}: worst force rms rel {worst}", flush=True)
    # the stepping path (not the component one) gives the same hybrid force
    ctx.compute()
    for r in (0, len(frames) // 2, len(frames) - 1):
        o = ref[r]
        Fh = o["F0"] + o["sc"].sp * o["dF"] + o["Fb"]
        assert rms_rel(ctx.forces(r), Fh) < F_TOL


def test_config3_ladder_components_match_oracle(hip_device):
    R = 22
    frames = config3_frames(R)
    params = systems.params_config3(R)
    print(f"case config3 ladder R={R} tuning={BENCH_TUNING}", flush=True)
    ctx = ladder(frames, params, BENCH_TUNING)
    check_ladder_components(ctx, frames, params)


@pytest.mark.parametrize("use_graph", [0, 1])
def test_config3_ladder_steps_match_oracle(hip_device, use_graph):
    """3 steps of the 22-replica ladder (the first two with the forces-only pair kernel,
    the last with the energies) against the oracle's Reference-form steps of replicas 0,
    10 and 21 on the same Philox noise."""
    R, n = 22, 3
    frames = config3_frames(R)
    params = systems.params_config3(R)
    print(f"case config3 ladder steps R={R} use_graph={use_graph}", flush=True)
    ctx = ladder(frames, params, {**BENCH_TUNING, "use_graph": use_graph})
    ctx.step(n)
    es = ctx.energies()
    topo = O.Topology(frames[0])

    def one(r):
        p = params[r]
        return O.ref_step(topo, O.make_params(p), frames[r].positions, frames[r].velocities,
                          0.0, 0, p.seed, r, n, use_cl_langevin=True)
    picks = [0, 10, 21]
    with ThreadPoolExecutor(POOL) as ex:
        refs = list(ex.map(one, picks))
    for r, (x, v, t, st, rep) in zip(picks, refs):
        e = es[r]
        assert e["flags"] == 0 and e["step"] == n and st == n
        xg, vg = ctx.get_positions(r), ctx.get_velocities(r)
        v0 = np.asarray(frames[r].velocities)
        dx = float(np.max(np.abs(xg - x)))
        dv = rms_rel(vg - v0, v - v0)
        print(f"replica {r}: max |dx| {dx:.2e} nm, velocity-change rms rel {dv:.2e}",
              flush=True)
        assert dx < 1e-5, (r, dx)
        assert dv < 1e-4, (r, dv)
        assert close(e["bind_e"], rep.s.bind_e), (r, e["bind_e"], rep.s.bind_e)
        assert close(e["W"], rep.s.ebias), (r, e["W"], rep.s.ebias)
        scale = abs(rep.E1) + abs(rep.s.ebias) + abs(rep.Eb)
        assert abs(e["pot_energy"] - rep.s.pot_energy) <= E_TOL * scale, (r, e["pot_energy"])
    # replicas that share nothing but the topology stay independent: none of them moved
    # like another
    assert not np.array_equal(ctx.get_positions(0), ctx.get_positions(21))


def test_config5_ladder_components_match_oracle(hip_device):
    R = 8
    frames = config5_frames(R)
    params = systems.params_config5(R)
    print(f"case config5 ladder R={R} tuning={LARGE_TUNING}", flush=True)
    ctx = ladder(frames, params, LARGE_TUNING)
    check_ladder_components(ctx, frames, params)


def test_config1_twenty_frames_match_oracle(hip_device):
    """Config 1 over 20 frames (seeds 1000..1019) in one batched context, each replica
    with its own lambda, alpha and u0."""
    R = 20
    frames = [systems.make_config1(frame=k) for k in range(R)]
    params = []
    for k in range(R):
        p = systems.params_config1()
        p.lambda1 = 0.05 + 0.02 * k
        p.lambda2 = 0.6 - 0.01 * k
        p.alpha, p.u0 = 0.01 * (k % 5), 50.0 + 5.0 * k
        params.append(p)
    print(f"case config1 frames R={R}", flush=True)
    ctx = ladder(frames, params, dict(pair_group=0, pair_newton3=1, pair_split=4,
                                      use_graph=1))
    check_ladder_components(ctx, frames, params)


def test_u_history_matches_step_records(hip_device):
    """The exchange record: u' taken on steps 0, every, 2 every, .. equals the u' that the
    step records of those steps reported; the newest records come back oldest first,
    also after the ring wraps and when fewer are asked for."""
    s = systems.make_small(frame=1)
    R = 2
    params = []
    for r in range(R):
        p = systems.params_config1()
        p.lambda1 = p.lambda2 = 0.2 + 0.3 * r
        p.seed = 40 + r
        params.append(p)
    ctx = ladder([s] * R, params, dict(use_graph=0))
    ctx.set_u_history(100, 8)
    rec = [[], []]
    for _ in range(300):
        ctx.step(1)
        for r, e in enumerate(ctx.energies()):
            rec[r].append(e["bind_e"])
    for r in range(R):
        h = ctx.u_history(r, 8)
        assert np.array_equal(h, [rec[r][0], rec[r][100], rec[r][200]]), (h, rec[r][::100])
        assert np.array_equal(ctx.u_history(r, 2), [rec[r][100], rec[r][200]])
    # a ring of 4 every 10 steps: after 300 + 100 steps, 10 records of the second run
    # (steps 300, 310, .. 390) were taken; the ring holds the last 4
    ctx.set_u_history(10, 4)
    more = [[], []]
    for _ in range(100):
        ctx.step(1)
        for r, e in enumerate(ctx.energies()):
            more[r].append(e["bind_e"])
    for r in range(R):
        assert np.array_equal(ctx.u_history(r, 4), [more[r][k] for k in (60, 70, 80, 90)])
        assert np.array_equal(ctx.u_history(r, 3), [more[r][k] for k in (70, 80, 90)])
