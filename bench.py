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

# Algorithmic bytes per particle per launch (each field touched once; the
# neighbours' re-reads and the list entries are not counted) and algorithmic
# FP64 flops per neighbour pair, for ndim = 3.
BYTES_3D = {
    "k_neighbours": 8 * 3 + 8 + 8 + 4 + 8 * 3 + 4 + 8 * 3 + 8,   # x,m,rho,info,v | nnbr,xs,vol
    "k_strain": 8 * 3 + 8 * 3 + 8 + 8 + 4 + 4 + 8 * 3 + 8 + 8 * 3 + 8 + 8 + 8,
    "k_momentum (+k_rigid_reduce)": 8 * 3 + 8 * 3 + 8 + 8 + 8 + 4 + 4 + 8 * 3,
    "k_integrate": 4 + 8 * 3 + 8 * 3 + 8 * 3 + 8 * 3 + 8 * 3 + 8 + 8,
}
FLOPS_PER_PAIR_3D = {"k_strain": 41, "k_momentum (+k_rigid_reduce)": 70 + 12}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="extra steps run with per-kernel HIP-event timing")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--small", action="store_true", help="reduced dam break (for quick checks)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run a plain 256 MiB device copy (reference for HBM counters)")
    ap.add_argument("--arms", default="fp64,fp32c", help="comma-separated arms to run")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from sph07_amd import scenes
    from sph07_amd.native import Scene, ARM_FP64, ARM_FP32C, F_V, F_NNBR

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
        arms[arm_name] = dict(scene=sc, elapsed=elapsed, t_init=t_init, kern=kern, rate=rate)
    # both arms advanced the same number of steps from the same deck
    moving = deck.dispbc == 0
    vel_err = None
    if "fp64" in arms and "fp32c" in arms:
        v64 = arms["fp64"]["scene"].field(F_V)[moving]
        v32 = arms["fp32c"]["scene"].field(F_V)[moving]
        vel_err = float(np.linalg.norm(v32 - v64) / max(np.linalg.norm(v64), 1e-300))
    main_arm = "fp32c" if "fp32c" in arms else "fp64"
    pairs = float(arms[main_arm]["scene"].field(F_NNBR).sum())
    elapsed, kern, t_init = arms[main_arm]["elapsed"], arms[main_arm]["kern"], arms[main_arm]["t_init"]
    value = arms[main_arm]["rate"]
    for a in arms.values():
        a["scene"].close()

    dom = max(kern, key=lambda k: kern[k][0])
    avg_ms = kern[dom][0]
    bytes_alg = BYTES_3D.get(dom, 0) * N
    achieved = bytes_alg / (avg_ms * 1e-3) / 1e9 if avg_ms else 0.0
    traffic = None
    tf_path = os.path.join(ROOT, "profiles", "traffic_dam_break.json")
    if os.path.exists(tf_path):
        traffic = json.load(open(tf_path)).get(dom)
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": bytes_alg}
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
        cpu = {"value": N * args.cpu_steps / tc, "unit": "particle-updates/s", "cores": 1,
               "kind": "port",
               "sample": f"{args.cpu_steps} full steps of the same {N}-particle dam break "
                         f"(after its untimed initialisation), single thread, {tc:.1f} s"}
        del o

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
        "neighbour_interactions_per_s": pairs * args.steps / elapsed,
        "init_s": round(t_init, 3),
        "kernels_avg_ms": kernels,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
