"""Benchmark of the per-timestep particle loop on one MI355X.

Workload: the 3D dam break (tank 1.6 x 0.67 x 0.6 m, water column
0.4 x 0.67 x 0.3 m, dx = 5e-3 m, h = 1.3 dx, 3 wall layers: 1,111,320
particles), synthetic lattice input, random-free.  A "step" is one full
iteration of MSolution::Solution over all particles, device-resident.

Prints one JSON line: particle-updates/s (real particles, walls included,
advanced one step), the roofline of the dominant kernel measured with HIP
events on the library's stream, and the CPU oracle timed on a bounded sample
of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-updates/s on 1 GPU vs plain FP64 GPU path; p50/p99 scene latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# MI355X's published vector FP32 rate (157.3 TF) counts packed FP32 FMAs; the
# library is compiled with packed FP32 turned off, so its FP32 VALU ceiling is
# one FMA per lane per clock, the same 78.6 TF as FP64
FP32_PEAK_TFLOPS = 78.6


def alg_bytes_3d(group, arm, k, compact=False):
    """Algorithmic bytes per particle for one launch of a timed kernel group,
    ndim = 3: every field the launch reads or writes counted once, its outputs
    included (the neighbour list: 4 bytes per entry); the neighbours' re-reads
    of each other's records are not algorithmic.  k = mean list length;
    compact: the FP32c momentum record is the 32-byte form (one mass per
    material, linear EOS) rather than the 40-byte one."""
    rec_s, rec_m = ((32, 32 if compact else 40) if arm == "fp32c" else (64, 80))
    if group == "k_neighbours":
        # read x, info; write nnbr, list (the half-step positions and the
        # continuity record are written by k_permute, in the bin group)
        return 24 + 4 + 4 + 4 * k
    if group == "k_strain":
        # read info, own record (fp32c) or xs+v, nnbr, list, rho, m, a (critical
        # step), cell code (fp32c); write trd, x, p, p/rho^2, rho_new, momentum record
        own = 32 + 24 + 24 + 4 if arm == "fp32c" else 24 + 24
        return 4 + own + 4 + 4 * k + 8 + 8 + 24 + 8 + 24 + 8 + 8 + 8 + rec_m
    if group == "k_momentum (+k_rigid_reduce)":
        # read info, rho_new, p/rho^2, own record (fp32c) or x+v, nnbr, list; write a
        own = rec_m if arm == "fp32c" else 48
        return 4 + 8 + 8 + own + 4 + 4 * k + 24
    if group == "k_integrate":
        # read info, a, x, v, rho_new; write a, v, x, rho
        return 4 + 24 + 24 + 24 + 8 + 24 + 24 + 24 + 8
    if group.startswith("bin"):
        # count: read x, info, write cell, slot; scatter: read cell, slot,
        # orig, write the cell's keys; permute: read x, v, a, rho, m, info,
        # orig, cell, key and write x, v, a, rho, m, vol, info, orig, cell
        # code, nnbr, the half-step positions and the continuity record
        return (24 + 4 + 4 + 4) + (4 + 4 + 4 + 4) + (24 * 3 + 8 + 8 + 4 + 4 + 4 + 4) + \
            (24 * 3 + 8 + 8 + 8 + 4 + 4 + 4 + 4 + 24 + rec_s)
    return 0


# algorithmic flops per neighbour pair (3D, reference formulas): continuity 41,
# Morris force 70 (+12 for the Monaghan term); per candidate test of the
# neighbour search 8 (3 sub, 3 mul, 2 add)
FLOPS_PER_PAIR_3D = {"k_strain": 41, "k_momentum (+k_rigid_reduce)": 82}
FLOPS_PER_CANDIDATE = 8

# The three dependency-separated phases of one particle-update and their
# compulsory traffic (one touch per field per phase), in bytes per particle
# for ndim D: bin reads x and writes a cell key (8D + 4); continuity reads
# {x, v, m, rho, flags} ((2D + 2) 8 + 4) and writes rho (8); force and
# integration read the same record and write x and v (16 D).  Flops per
# neighbour pair: continuity 41, Morris force 70 (3D; the reference's
# formulas).  The neighbour list is an intermediate of this build, not
# algorithmic.  Kernel groups of the library's timing -> phase:
PHASE_OF_GROUP = {
    "bin (grid, count, scan, scatter, rank)": "bin", "periodic images": "bin", "k_neighbours": "bin",
    "k_strain": "continuity", "k_normals+k_curvature": "continuity", "k_advance": "continuity",
    "k_momentum (+k_rigid_reduce)": "force+integrate", "k_integrate": "force+integrate",
    "k_finish": "force+integrate",
}


def phase_bytes(phase, D=3):
    rec = (2 * D + 2) * 8 + 4
    return {"bin": 8 * D + 4, "continuity": rec + 8, "force+integrate": rec + 16 * D}[phase]


PHASE_FLOPS_PER_PAIR_3D = {"bin": 0, "continuity": 41, "force+integrate": 70}
# peak of the phase's arithmetic: FP64 vector 78.6 TF/s; FP32 157.3 TF/s
# (MI355X's published vector rate, which counts packed FP32 FMAs)
PEAK_TFLOPS = {"fp64": 78.6, "fp32c": 157.3}


def phase_roofline(kern, arm, N, k):
    """t_lower = sum over phases of max(B_alg / HBM peak, F_alg / VALU peak)
    against the measured kernel times of each phase."""
    out = {}
    for ph in ("bin", "continuity", "force+integrate"):
        ms = sum(v[0] for g, v in kern.items() if v[1] and PHASE_OF_GROUP.get(g) == ph)
        b = phase_bytes(ph) * N
        f = PHASE_FLOPS_PER_PAIR_3D[ph] * k * N
        peak = PEAK_TFLOPS[arm] if ph != "bin" else PEAK_TFLOPS["fp64"]
        t_lo = max(b / (HBM_PEAK_GBS * 1e9), f / (peak * 1e12)) * 1e3
        out[ph] = {"kernels": [g for g in kern if PHASE_OF_GROUP.get(g) == ph and kern[g][1]],
                   "bytes_per_particle": phase_bytes(ph), "flops_per_pair": PHASE_FLOPS_PER_PAIR_3D[ph],
                   "measured_ms": round(ms, 4), "t_lower_ms": round(t_lo, 5),
                   "frac": round(t_lo / ms, 4) if ms else None}
    return out


def candidate_tests(x, deck):
    """Candidate pairs the neighbour search tests in one step: for every
    particle, the particles of the 3^d cells around its 2h cell, itself
    excluded (free axes: the grid of the current bounding box)."""
    import numpy as np
    hmax = max(m.h for m in deck.materials)
    gs = 1.0 / (2.0 * hmax)
    lo = x.min(axis=0) - 0.01 * hmax
    c = np.floor(gs * (x - lo)).astype(np.int64)
    lim = c.max(axis=0) + 1
    key = (c[:, 0] * lim[1] + c[:, 1]) * lim[2] + c[:, 2]
    counts = np.zeros(int(lim.prod()), dtype=np.int64)
    np.add.at(counts, key, 1)
    grid = counts.reshape(tuple(lim))
    pad = np.pad(grid, 1)
    stencil = sum(pad[1 + a:1 + a + lim[0], 1 + b:1 + b + lim[1], 1 + e:1 + e + lim[2]]
                  for a in (-1, 0, 1) for b in (-1, 0, 1) for e in (-1, 0, 1))
    return float((stencil[c[:, 0], c[:, 1], c[:, 2]] - 1).sum())


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="extra steps run with per-kernel HIP-event timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--small", action="store_true", help="reduced dam break (for quick checks)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a plain 256 MiB device copy (reference for HBM counters)")
    ap.add_argument("--arms", default="fp64,fp32c", help="comma-separated arms to run")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="steps both arms have run (the timed ones included, the rest untimed) when the "
                         "FP32c error against FP64 is measured")
    ap.add_argument("--serve-requests", type=int, default=32,
                    help="requests of the serving leg (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=6.0, help="Poisson arrival rate, 1/s")
    ap.add_argument("--serve-steps-scale", type=float, default=1.0,
                    help="scales the 500-5000 step counts of the serving mix")
    ap.add_argument("--ensemble", type=int, default=512, help="scenes of the ensemble leg (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=2000,
                    help="steps per ensemble scene (the workload's 2000)")
    args = ap.parse_args()
    if args.profile_steps < 1:
        ap.error("--profile-steps must be at least 1 (the roofline needs per-kernel times)")
    return args


def serving_leg(args):
    """Continuous batching against one-scene-at-a-time on the same seeded
    Poisson stream (40% Poiseuille, 30% cavity, 20% droplet, 10% cylinder;
    4k-500k particles), FP32-compensated arm."""
    from sph07_amd import serving
    from sph07_amd.native import ARM_FP32C
    kw = dict(seed=0x5EED0004, n=args.serve_requests, rate=args.serve_rate,
              steps_scale=args.serve_steps_scale)
    out = {"requests": args.serve_requests, "rate_per_s": args.serve_rate,
           "steps_scale": args.serve_steps_scale, "arm": "fp32c", "hbm_budget_gb": 200}
    print(f"bench: serving leg, {args.serve_requests} requests", file=sys.stderr, flush=True)
    reqs = serving.serving_stream(**kw)
    out["particles"] = [r.deck.np for r in reqs]
    out["steps"] = [r.steps for r in reqs]
    b = serving.serve_batched(reqs, ARM_FP32C, hbm_budget=200e9, chunk=25).summary()
    print(f"bench:   batched {b}", file=sys.stderr, flush=True)
    reqs = serving.serving_stream(**kw)
    o = serving.serve_one_at_a_time(reqs, ARM_FP32C).summary()
    print(f"bench:   one at a time {o}", file=sys.stderr, flush=True)
    out["batched"], out["one_at_a_time"] = b, o
    out["throughput_ratio"] = b["particle_updates_per_s"] / o["particle_updates_per_s"]
    # the same requests all queued at t = 0: the device is never idle, so the
    # throughput is each mode's capacity (p50/p99 include the queueing)
    reqs = serving.serving_stream(**kw)
    for r in reqs:
        r.arrival = 0.0
    bs = serving.serve_batched(reqs, ARM_FP32C, hbm_budget=200e9, chunk=25).summary()
    print(f"bench:   saturated, batched {bs}", file=sys.stderr, flush=True)
    reqs = serving.serving_stream(**kw)
    for r in reqs:
        r.arrival = 0.0
    os_ = serving.serve_one_at_a_time(reqs, ARM_FP32C).summary()
    print(f"bench:   saturated, one at a time {os_}", file=sys.stderr, flush=True)
    out["saturated"] = {"batched": bs, "one_at_a_time": os_,
                        "throughput_ratio": bs["particle_updates_per_s"] / os_["particle_updates_per_s"]}
    del out["particles"], out["steps"]
    out["particles_total"] = int(sum(r.deck.np for r in reqs))
    return out


def ensemble_leg(args):
    """Independent 2D scenes (alternately Poiseuille and lid-driven cavity,
    2k-64k particles, nu log-uniform 1e-6..1e-4) advanced together by the
    batched step, FP32-compensated arm."""
    from sph07_amd import serving
    from sph07_amd.native import ARM_FP32C, Batch
    reqs = serving.ensemble(n=args.ensemble, steps=args.ensemble_steps)
    print(f"bench: ensemble leg, {len(reqs)} scenes x {args.ensemble_steps} steps",
          file=sys.stderr, flush=True)
    b = Batch(2, ARM_FP32C)
    t0 = time.perf_counter()
    for r in reqs:
        b.add(r.deck, r.steps)
    t_add = time.perf_counter() - t0
    t0 = time.perf_counter()
    live = b.live
    while live:
        live = b.step(100)
    el = time.perf_counter() - t0
    errs = sum(1 for s in b.finished() if b.status(s)["error"])
    b.close()
    upd = sum(r.deck.np * r.steps for r in reqs)
    return {"scenes": len(reqs), "steps": args.ensemble_steps, "particles_total": int(sum(r.deck.np for r in reqs)),
            "seconds": round(el, 3), "init_seconds": round(t_add, 3),
            "particle_updates_per_s": upd / el, "failed": errs}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from sph07_amd import scenes
    from sph07_amd.native import Scene, ARM_FP64, ARM_FP32C, F_V, F_X, F_RHO, F_NNBR

    assert torch.cuda.is_available(), "no GPU visible"
    if args.calib_copy:
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")   # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    kw = dict(nx_fluid=20, ny=34, nz_fluid=15, nx_tank=80, nz_tank=30) if args.small else {}
    deck = scenes.dam_break(**kw)
    N = deck.np

    def run_arm(arm):
        t0 = time.perf_counter()
        sc = Scene(deck, arm)
        t_init = time.perf_counter() - t0
        sc.Solution(args.warmup)
        sc.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = sc.Solution(args.steps)
        sc.synchronize()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        assert done == args.steps, (done, args.steps)
        sc.set_timing(True)
        sc.Solution(args.profile_steps)
        tm = sc.timing()
        sc.set_timing(False)
        kern = {k: (ms / cnt if cnt else 0.0, cnt) for k, (ms, cnt) in tm.items()}
        sc.launch_shape = sc.autotune(-1)
        return sc, elapsed, t_init, kern

    arms = {}
    wanted = args.arms.split(",")
    for arm_name, arm in (("fp64", ARM_FP64), ("fp32c", ARM_FP32C)):
        if arm_name not in wanted:
            continue
        print(f"bench: dam_break {N} particles ({deck.n_fluid} fluid), {arm_name} arm",
              file=sys.stderr, flush=True)
        sc, elapsed, t_init, kern = run_arm(arm)
        rate = N * args.steps / elapsed
        print(f"bench: {arm_name}: {args.steps} steps in {elapsed:.3f} s -> {rate:.4e} particle-updates/s",
              file=sys.stderr, flush=True)
        for k, (avg, cnt) in kern.items():
            if cnt:
                print(f"bench:   {k:40s} {avg * 1e3:9.1f} us x {cnt}", file=sys.stderr, flush=True)
        arms[arm_name] = dict(scene=sc, elapsed=elapsed, t_init=t_init, kern=kern, rate=rate,
                              cand=candidate_tests(sc.field(F_X), deck), pairs=float(sc.field(F_NNBR).sum()))
        # untimed: on to the step count at which the arms are compared
        ran = args.warmup + args.steps + args.profile_steps
        if args.error_steps > ran:
            assert sc.Solution(args.error_steps - ran) == args.error_steps - ran
        arms[arm_name]["steps_run"] = max(ran, args.error_steps)
    # both arms advanced the same number of steps from the same deck
    moving = deck.dispbc == 0
    vel_err = rho_err = None
    if "fp64" in arms and "fp32c" in arms:
        v64 = arms["fp64"]["scene"].field(F_V)[moving]
        v32 = arms["fp32c"]["scene"].field(F_V)[moving]
        vel_err = float(np.linalg.norm(v32 - v64) / max(np.linalg.norm(v64), 1e-300))
        r64 = arms["fp64"]["scene"].field(F_RHO)
        r32 = arms["fp32c"]["scene"].field(F_RHO)
        rho0 = deck.materials[0].rho
        # against the density deviation rho - rho0 (rho itself is ~rho0 everywhere)
        rho_err = float(np.linalg.norm(r32 - r64) / max(np.linalg.norm(r64 - rho0), 1e-300))
    main_arm = "fp32c" if "fp32c" in arms else "fp64"
    pairs = arms[main_arm]["pairs"]
    kmean = pairs / N
    elapsed, kern, t_init = arms[main_arm]["elapsed"], arms[main_arm]["kern"], arms[main_arm]["t_init"]
    value = arms[main_arm]["rate"]
    for a in arms.values():
        a["scene"].close()

    # the library's rule for the compact FP32c momentum record
    compact = deck.massopt != 1 and all(m.eos == 29 for m in deck.materials)
    # dominant kernel of the main arm: the longest timed group; its algorithmic
    # bytes are its phase's compulsory bytes per particle x N
    dom = max((g for g in kern if kern[g][1] and g in PHASE_OF_GROUP), key=lambda g: kern[g][0])
    avg_ms = kern[dom][0]
    per_particle = phase_bytes(PHASE_OF_GROUP[dom])
    bytes_alg = per_particle * N
    achieved = bytes_alg / (avg_ms * 1e-3) / 1e9 if avg_ms else 0.0
    traffic = occupancy = None
    tf_path = os.path.join(ROOT, "profiles", "traffic_dam_break.json")
    if os.path.exists(tf_path):
        traffic = json.load(open(tf_path)).get(main_arm, {}).get(dom)
    oc_path = os.path.join(ROOT, "profiles", "occupancy_dam_break.json")
    occ_all = json.load(open(oc_path)) if os.path.exists(oc_path) else {}
    occupancy = occ_all.get(main_arm, {}).get(dom)
    list_bytes = alg_bytes_3d(dom, main_arm, kmean, compact) * N
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "avg_launch_ms": round(avg_ms, 4), "phase": PHASE_OF_GROUP[dom],
            "algorithmic_bytes_per_particle": per_particle, "particles_per_launch": N,
            "algorithmic_bytes_per_launch": int(bytes_alg),
            "list_inclusive_bytes_per_launch": int(list_bytes),
            "list_inclusive_frac": round(list_bytes / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "occupancy_waves_per_simd": occupancy}
    phases = {a: phase_roofline(arms[a]["kern"], a, N, kmean) for a in arms}
    step_roof = {}
    for a in arms:
        t_lo = sum(p["t_lower_ms"] for p in phases[a].values())
        t_ms = arms[a]["elapsed"] / args.steps * 1e3
        step_roof[a] = {"t_lower_ms": round(t_lo, 5), "t_measured_ms": round(t_ms, 4),
                        "frac": round(t_lo / t_ms, 4)}
    # the same kernels against the vector-ALU roof (FP64 for the neighbour
    # test and the FP64 arm, FP32 for the compensated sweeps)
    valu = {}
    for a in arms:
        for g, ms in ((g, v[0]) for g, v in arms[a]["kern"].items() if v[1]):
            if g == "k_neighbours":
                flops, peak = arms[a]["cand"] * FLOPS_PER_CANDIDATE, FP64_PEAK_TFLOPS
            elif g in FLOPS_PER_PAIR_3D:
                flops = pairs * FLOPS_PER_PAIR_3D[g]
                peak = FP32_PEAK_TFLOPS if a == "fp32c" else FP64_PEAK_TFLOPS
            else:
                continue
            tf = flops / (ms * 1e-3) / 1e12
            valu.setdefault(a, {})[g] = {"achieved_tflops": round(tf, 3), "peak": peak,
                                         "frac": round(tf / peak, 4)}
    hbm = {a: {g: round(alg_bytes_3d(g, a, kmean, compact) * N / (v[0] * 1e-3) / 1e9, 1)
               for g, v in arms[a]["kern"].items() if v[1] and alg_bytes_3d(g, a, kmean, compact)}
           for a in arms}
    kernels = {a: {k: round(v[0], 4) for k, v in arms[a]["kern"].items() if v[1]} for a in arms}

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as orc
        print(f"bench: cpu baseline: oracle, {args.cpu_steps} steps of the same scene",
              file=sys.stderr, flush=True)
        o = orc.Oracle(deck)
        t0 = time.perf_counter()
        o.run(args.cpu_steps)
        tc = time.perf_counter() - t0
        nthr = orc.threads()
        cpu = {"value": N * args.cpu_steps / tc, "unit": "particle-updates/s", "cores": nthr,
               "kind": "port",
               "sample": f"{args.cpu_steps} full steps of the same {N}-particle dam break "
                         f"(after its untimed initialisation), FP64 restatement, {nthr} OpenMP "
                         f"threads over particles, {tc:.1f} s"}
        del o

    serve = serving_leg(args) if args.serve_requests > 0 else None
    ens = ensemble_leg(args) if args.ensemble > 0 else None

    out = {
        "metric": METRIC, "value": value, "unit": "particle-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "dam_break_3d" + ("_small" if args.small else ""), "particles": N,
                   "fluid_particles": deck.n_fluid, "dx": deck.meta["dx"], "h_over_dx": 1.3,
                   "arm": main_arm, "monaghan_alpha": 0.1, "parallelism": "single-gpu"},
        "fp64_value": arms["fp64"]["rate"] if "fp64" in arms else None,
        "speedup_vs_fp64": value / arms["fp64"]["rate"] if "fp64" in arms else None,
        "fp32c_rel_l2_velocity_error_vs_fp64": vel_err,
        "fp32c_rel_l2_density_deviation_error_vs_fp64": rho_err,
        "error_at_steps": arms[main_arm]["steps_run"],
        "neighbour_interactions_per_s": pairs * args.steps / elapsed,
        "init_s": round(t_init, 3),
        "kernels_avg_ms": kernels,
        "launch_shapes": {a: arms[a]["scene"].launch_shape for a in arms},
        "roofline": roof,
        "step_roofline": step_roof,
        "phase_roofline": phases,
        "occupancy_waves_per_simd": occ_all or None,
        "list_inclusive_gbs_per_kernel": hbm,
        "valu_roofline_per_kernel": valu,
        "cpu_baseline": cpu,
        "serving": serve,
        "p50_time_to_solution_s": serve["batched"]["p50_time_to_solution_s"] if serve else None,
        "p99_time_to_solution_s": serve["batched"]["p99_time_to_solution_s"] if serve else None,
        "ensemble": ens,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
