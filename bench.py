#!/usr/bin/env python3
"""Benchmark: aggregate MD steps/s of the batched 22-replica SDM lambda ladder at 22,750 atoms.

Workload (one "step" = one SDM Langevin step of every replica, inputs resident in HBM):
  22 replicas of the 22,750-atom host-guest system (7,500 water-like triatomics, 200-atom
  host shell, 50-atom guest displaced by (2.5,0,0) nm), 6.1 nm periodic box, LJ + reaction
  field (rc 1.0 nm, eps_rf 78.5), ILogistic bias (alpha 0.1, u0 110), rational soft core,
  lambda ladder 0->0.5 (11 replicas) and 0.5->0 (11 replicas), 1 fs, 300 K, 1/ps friction.
  Synthetic seeded coordinates (no datasets); every replica has its own coordinates.

Printed: one JSON line (rank 0).  `value` = aggregate steps/s of all replicas (22 x steps /
elapsed).  As in the reference, where the potential energy of a step(n) call is read
after it, the pair energies are evaluated on the last step of each call; u, W, sp, the
work and the forces on every step.  Also reported: the same ladder with the pair energies
on every step, the single-replica plain path (R=1) steps/s and ns/day, the batched/plain
ratio, the 8-replica ladder of the 99,920-atom relative-binding system (two guests
displaced by +-3 nm, rc 1.2 nm), the roofline of the pair kernel as the steps run it
(forces only; FP32 VALU, 42 FLOP per in-cutoff pair, and the 50-FLOP count that includes
the energy terms) timed live with HIP events on the library stream, with the VALU
counters of the committed rocprofv3 summary, and the FP64 CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 50            # FMA = 2; see the roofline accounting in the design notes
FLOP_PER_PAIR_FORCES = 42     # the same without the energy terms (LJ 2, RF 4, sums 2)
FP32_PEAK_TFLOPS = 157.3      # MI355X dense FP32 vector (published, packed FP32)
FP32_NONPACKED_TFLOPS = 78.6  # one FP32 FMA per lane per cycle (packed FP32 is not used)
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--replicas", type=int, default=22)
    ap.add_argument("--single-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the 8-replica 99,920-atom ladder")
    ap.add_argument("--large-steps", type=int, default=300)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--profile", type=int, default=0,
                    help="profiling mode: N eager steps of the 22-replica ladder, no other legs")
    ap.add_argument("--side-fork", type=int, default=None,
                    help="profiling modes: placement of the side work (sdm_tuning.side_fork)")
    ap.add_argument("--list-skin", type=float, default=None,
                    help="profiling modes: Verlet skin of the tile lists in nm "
                         "(sdm_tuning.list_skin)")
    ap.add_argument("--equilibrate", type=int, default=0,
                    help="profiling modes: steps run (from captured graphs) before the "
                         "profiled ones, so that the lattice-built start has relaxed")
    ap.add_argument("--profile-large", type=int, default=0,
                    help="profiling mode: N eager steps of the 8-replica 99,920-atom ladder "
                         "(its benchmark tuning, kernels launched one by one), no other legs")
    ap.add_argument("--profile-single", type=int, default=0,
                    help="profiling mode: N steps of one replica of the 22,750-atom system "
                         "replayed from captured graphs (32 waves per i-block, side work "
                         "forked after the slot order), no other legs")
    ap.add_argument("--profile-pair", type=int, default=0,
                    help="profiling mode: the benchmark's autotuning and warm-up, then the "
                         "pair kernel alone N times (the launches the roofline figure "
                         "times), no other legs")
    return ap.parse_args()


def read_profile_value(name, key):
    """A `key: value` line of a committed rocprofv3 summary under profiles/ (None if absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def read_traffic():
    """HBM bytes per pair-kernel launch from the committed rocprofv3 --pmc summary."""
    return read_profile_value("k_pair_n3_config3_pmc.md", "traffic_bytes_per_launch")


def note(msg):
    """Progress of the legs on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_info():
    """Core count (nproc) and model name (lscpu) of the machine the CPU legs run on."""
    import subprocess
    info = {"nproc": os.cpu_count(), "model": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True,
                                           timeout=10).stdout.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(n_steps, warmup=10, runs=3, omp_steps=2):
    """The FP64 oracle's Reference-platform step (E at x, E at x+d, bonded, execute,
    Langevin) on the 22,750-atom replica: a warm-up, then the median of `runs` timed runs
    of n_steps on one thread.  22 sequential replicas on one core give the same aggregate
    steps/s.  Beside it, the one-replica-per-core form: as many replicas as threads (the
    machine's cores, at most 16), each stepped omp_steps times after one warm-up step."""
    import numpy as np
    from openmm_sdm_plugin_amd import systems
    from oracle import oracle as O
    s = systems.make_config2(frame=0)
    params = systems.params_config3(22)
    p = params[11]
    topo = O.Topology(s)
    op = O.make_params(p)
    x, v = s.positions.copy(), s.velocities.copy()
    x, v, t, st, _ = O.ref_step(topo, op, x, v, 0.0, 0, p.seed, 0, warmup)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        x, v, t, st, _ = O.ref_step(topo, op, x, v, t, st, p.seed, 0, n_steps)
        times.append(time.perf_counter() - t0)
    assert np.all(np.isfinite(x))
    med = float(np.median(times))
    info = cpu_info()
    threads = max(1, min(16, int(info["nproc"] or 1)))
    reps = [params[k % 22] for k in range(threads)]
    ops = [O.make_params(q) for q in reps]
    xs, vs = [s.positions] * threads, [s.velocities] * threads
    seeds = [q.seed for q in reps]
    xs, vs = O.ref_step_replicas(topo, ops, xs, vs, seeds, 1, threads)  # warm-up
    t0 = time.perf_counter()
    xs, vs = O.ref_step_replicas(topo, ops, xs, vs, seeds, omp_steps, threads)
    omp_dt = time.perf_counter() - t0
    return {"value": n_steps / med, "unit": "steps/s", "cores": 1, "kind": "port",
            "sample": f"median of {runs} runs of {n_steps} steps (after {warmup} warm-up "
                      "steps) of one 22,750-atom ladder replica, FP64 oracle in its "
                      "Reference-platform form (E at x, E at x+d, bonded, execute, "
                      "Langevin), one thread; 22 sequential replicas give the same "
                      "aggregate steps/s",
            "run_seconds": times, "nproc": info["nproc"], "cpu_model": info["model"],
            "openmp_one_replica_per_core": {
                "value": threads * omp_steps / omp_dt, "unit": "aggregate steps/s",
                "threads": threads, "replicas": threads,
                "sample": f"{threads} replicas x {omp_steps} steps after 1 warm-up step, "
                          "one replica per OpenMP thread"}}


# the launch configurations sdm_autotune picks on MI355X for the 22-replica 22,750-atom
# ladder and the 8-replica 99,920-atom ladder (used by the profiling modes, which do not
# autotune)
LADDER_TUNING = dict(pair_group=0, pair_newton3=1, pair_split=4, pair_waves_per_group=1,
                     cell_cells_per_dim=20, sort_interval=20, use_graph=0, energy_every_step=0,
                     side_fork=1)
SINGLE_TUNING = dict(pair_group=0, pair_newton3=1, pair_split=32, pair_waves_per_group=1,
                     cell_cells_per_dim=20, sort_interval=20, use_graph=1, energy_every_step=0,
                     side_fork=4)
LARGE_TUNING = dict(pair_group=0, pair_newton3=1, pair_split=4, pair_waves_per_group=1,
                    cell_cells_per_dim=33, sort_interval=20, energy_every_step=0, side_fork=1)


def large_ladder_context(systems, _native):
    """8 replicas of the 99,920-atom two-guest system (lambda 0.1-0.45), from one frame."""
    s5 = systems.make_config5(frame=0)
    p5 = systems.params_config5(8)
    big = _native.NativeContext(s5, 8)
    for r in range(8):
        big.set_params(r, p5[r])
        big.set_positions(r, s5.positions)
        big.set_velocities(r, s5.velocities)
    return big


def main():
    a = parse()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the replica ladder would shard by replica; see the design notes).",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from openmm_sdm_plugin_amd import _native, systems

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if a.profile_large:
        big = large_ladder_context(systems, _native)
        big.set_tuning(**LARGE_TUNING, use_graph=1)
        if a.side_fork is not None:
            big.set_tuning(side_fork=a.side_fork)
        if a.list_skin is not None:
            big.set_tuning(list_skin=a.list_skin)
        big.step(a.equilibrate)
        big.set_tuning(use_graph=0)
        big.step(a.profile_large)
        e5 = big.energies()
        print(json.dumps({"profile_large_steps": a.profile_large, "step": e5[0]["step"],
                          "flags": [x["flags"] for x in e5], "tuning": big.get_tuning()}))
        return
    if a.profile_single:
        s2 = systems.make_config2(frame=0)
        one = _native.NativeContext(s2, 1)
        one.set_params(0, systems.params_config3(22)[11])
        one.set_positions(0, s2.positions)
        one.set_velocities(0, s2.velocities)
        one.set_tuning(**SINGLE_TUNING)
        one.step(max(a.equilibrate, 100))
        one.energies()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        one.step(a.profile_single)
        e1 = one.energies()
        dt = time.perf_counter() - t0
        print(json.dumps({"profile_single_steps": a.profile_single, "steps_per_s": a.profile_single / dt,
                          "flags": e1[0]["flags"], "tuning": one.get_tuning()}))
        return
    R = a.replicas
    frames = [systems.make_config2(frame=k) for k in range(R)]
    params = systems.params_config3(R)
    ctx = _native.NativeContext(frames[0], R)
    for r in range(R):
        ctx.set_params(r, params[r])
        ctx.set_positions(r, frames[r].positions)
        ctx.set_velocities(r, frames[r].velocities)
    ctx.set_u_history(100, 4096)
    if a.profile or a.profile_pair:
        if a.profile_pair:
            # the same sequence as the benchmark up to its roofline timing (with
            # --no-autotune: the configuration the autotuner picks on MI355X, so a counter
            # run does not profile the autotuner's trials)
            if not a.no_autotune:
                ctx.autotune(40)
            else:
                ctx.set_tuning(**LADDER_TUNING)
            ctx.step(a.warmup)
            ms = ctx.time_pair_kernel(a.profile_pair)
            print(json.dumps({"pair_kernel_ms_events": ms, "tuning": ctx.get_tuning()}))
        else:
            # one launch per kernel (no graph) so a profiler attributes every dispatch
            ctx.set_tuning(**dict(LADDER_TUNING, use_graph=1))
            if a.side_fork is not None:
                ctx.set_tuning(side_fork=a.side_fork)
            if a.list_skin is not None:
                ctx.set_tuning(list_skin=a.list_skin)
            ctx.step(a.equilibrate)
            ctx.set_tuning(use_graph=0)
            ctx.step(a.profile)
        e = ctx.energies()
        # calibration for the HBM-traffic counters: a fill that writes 64 MiB and a sum
        # that reads the same 64 MiB (plain streaming kernels of known size); a second
        # 64 MiB fill in between evicts the first buffer from the L2s (32 MiB in all)
        src = torch.ones(16 * 1024 * 1024, dtype=torch.float32, device="cuda")
        evict = torch.ones(16 * 1024 * 1024, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        tot = float(src.sum())
        torch.cuda.synchronize()
        del evict
        print(json.dumps({"profile_steps": a.profile or a.profile_pair, "step": e[0]["step"],
                          "flags": [x["flags"] for x in e], "tuning": ctx.get_tuning(),
                          "calibration_bytes": src.numel() * 4, "calibration_sum": tot}))
        return
    tuning = ctx.autotune(40) if not a.no_autotune else ctx.get_tuning()
    note(f"ladder R={R} tuned: {tuning}")

    # ---- batched ladder: warm-up ----
    ctx.step(a.warmup)
    ctx.energies()

    # ---- roofline of the pair kernel (live HIP events on the library stream), at the
    # state after the warm-up, before the timed steps (it does not move the positions;
    # the next step recomputes the forces) ----
    p0_a, _ = ctx.count_pairs()
    pair_ms = ctx.time_pair_kernel(20)
    p0_b, _ = ctx.count_pairs()
    pairs = 0.5 * (float(p0_a.sum()) + float(p0_b.sum()))
    achieved = FLOP_PER_PAIR_FORCES * pairs / (pair_ms * 1e-3) / 1e12
    achieved_50 = FLOP_PER_PAIR * pairs / (pair_ms * 1e-3) / 1e12
    traffic = read_traffic()
    # VALU evidence from the committed counter summary (SQ_INSTS_VALU per launch, and the
    # busy fraction derived from SQ_ACTIVE_INST_VALU and GRBM_GUI_ACTIVE)
    vname = "k_pair_n3_config3_valu.md"
    valu_busy = read_profile_value(vname, "valu_busy")
    vinst = read_profile_value(vname, "valu_insts_per_launch")
    valu_per_pair = 64.0 * vinst / pairs if vinst else None

    note(f"pair kernel {pair_ms:.4f} ms per launch")
    # ---- exactly K timed steps ----
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.step(a.steps)
    e = ctx.energies()  # synchronises the library stream
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    agg = R * a.steps / elapsed
    flags = [x["flags"] for x in e]
    ok = all(np.isfinite(x["pot_energy"]) for x in e) and not any(flags)

    note(f"timed ladder: {agg:.0f} steps/s")
    # ---- the same with the pair energies on every step ----
    ctx.set_tuning(energy_every_step=1)
    ctx.step(a.warmup)
    ctx.energies()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.step(a.steps)
    e_all = ctx.energies()
    torch.cuda.synchronize()
    agg_all = R * a.steps / (time.perf_counter() - t0)
    ok = ok and not any(x["flags"] for x in e_all)
    ctx.set_tuning(energy_every_step=0)

    note(f"energies every step: {agg_all:.0f} steps/s")
    # ---- plain single-replica path (R = 1) ----
    one = _native.NativeContext(frames[0], 1)
    one.set_params(0, params[R // 2] if R > 1 else params[0])
    one.set_positions(0, frames[0].positions)
    one.set_velocities(0, frames[0].velocities)
    if not a.no_autotune:
        one.autotune(200)
    one.step(a.warmup)
    one.energies()
    t1 = time.perf_counter()
    one.step(a.single_steps)
    one.energies()
    single = a.single_steps / (time.perf_counter() - t1)
    del one

    note(f"single replica: {single:.0f} steps/s")
    # ---- the large relative-binding ladder: 8 replicas of the 99,920-atom system (two
    # guests displaced by +-3 nm, rc 1.2 nm), all starting from one frame ----
    large = None
    if not a.no_large:
        big = large_ladder_context(systems, _native)
        tuning5 = big.autotune(10) if not a.no_autotune else big.get_tuning()
        big.step(a.warmup)
        big.energies()
        torch.cuda.synchronize()
        t5 = time.perf_counter()
        big.step(a.large_steps)
        e5 = big.energies()
        torch.cuda.synchronize()
        dt5 = time.perf_counter() - t5
        ok = ok and not any(x["flags"] for x in e5)
        large = {"workload": "8-replica lambda ladder (0.1-0.45) of the 99,920-atom system, "
                             "10 nm box, rc 1.2 nm, replicas from one frame",
                 "aggregate_steps_per_s": 8 * a.large_steps / dt5,
                 "ms_per_step": 1e3 * dt5 / a.large_steps, "steps": a.large_steps,
                 "tuning": tuning5}
        del big

    if large:
        note(f"large ladder: {large['aggregate_steps_per_s']:.0f} steps/s")
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a.cpu_steps)
        note(f"cpu baseline: {cpu['value']:.2f} {cpu['unit']}")

    out = {
        "metric": "aggregate MD steps/s across batched lambda-replicas at 23k atoms; "
                  "single-replica ns/day",
        "value": agg,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded coordinates, random-draw guest/host parameters)",
        "config": {
            "workload": "config 3: 22-replica lambda ladder of the 22,750-atom host-guest "
                        "system, 6.1 nm box, LJ+RF rc 1.0 nm, 1 fs Langevin",
            "replicas": R, "atoms": frames[0].n, "arithmetic": "pair math f32, int64 "
            "fixed-point forces, f64 energy accumulation, f64 displaced-state terms and "
            "integration", "tuning": tuning,
        },
        "single_replica_steps_per_s": single,
        "single_replica_ns_per_day": single * 0.001 * 86400.0 / 1000.0,
        "large_ladder": large,
        "steps_per_s_energy_every_step": agg_all,
        "batched_over_plain": agg / single,
        "all_finite_no_flags": ok,
        "roofline": {
            "bound": "valu", "kernel": "k_pair_n3 (forces only)", "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic, "pair_kernel_ms": pair_ms,
            "pairs_per_launch": pairs, "flop_per_pair": FLOP_PER_PAIR_FORCES,
            "frac_50flop": achieved_50 / FP32_PEAK_TFLOPS,
            "frac_of_nonpacked_peak": achieved / FP32_NONPACKED_TFLOPS,
            "valu_busy": valu_busy,
            "valu_lane_ops_per_in_cutoff_pair": valu_per_pair,
        },
        "cpu_baseline": cpu,
    }
    if cpu:
        out["gpu_over_cpu"] = agg / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
