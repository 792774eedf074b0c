"""Accuracy of the FP32-compensated arm against the FP64 arm on the three
single-scene workloads at their full sizes and step counts, and of both arms
against the analytic channel-flow solution.

    python -m sph07_amd.accuracy [--only poiseuille,dam_break,cylinder]

prints one JSON object:

* poiseuille -- 40 x 40 fluid particles, 2000 steps: the largest deviation of
  each arm's velocity profile from the start-up series of Morris et al. (1997)
  eq. 21 (channel width L + dx: the walls put the no-slip plane half a spacing
  out), relative to the series' peak, and the FP32c arm's relative L2 velocity
  error against the FP64 arm;
* dam_break -- 1,111,320 particles, 1000 steps: relative L2 errors of the
  fluid's velocity and of the density deviation rho - rho0;
* cylinder -- 509,036 particles, 5000 steps: the rigid body's trajectory (its
  centre every 250 steps) in the FP32c arm against the FP64 arm: the largest
  distance between the two, also relative to the distance the body has moved
  by the end, and the body's velocity in both; beside it the FP64 arm run
  from positions one ulp apart, which shows how far rounding-level
  differences grow in this flow whatever the arithmetic.

The FP64 arm reproduces the oracle bit for bit (the parity tests), so it is the
reference here.  The inputs are the synthetic lattices of sph07_amd.scenes.
"""
import argparse
import json
import sys
import time

import numpy as np


def poiseuille():
    from . import scenes
    from .native import Scene, ARM_FP64, ARM_FP32C, F_X, F_V
    deck = scenes.poiseuille()
    m = deck.meta
    nf = deck.n_fluid
    out, v = {"particles": deck.np, "steps": 2000}, {}
    for name, arm in (("fp64", ARM_FP64), ("fp32c", ARM_FP32C)):
        s = Scene(deck, arm)
        assert s.Solution(2000) == 2000
        g = s.globals()
        x, vel = s.field(F_X)[:nf], s.field(F_V)[:nf]
        ua = scenes.poiseuille_series(x[:, 1] + m["dx"] / 2, g["ptime"], m["L"] + m["dx"], m["F"], m["nu"])
        out[f"{name}_max_profile_error_vs_series"] = float(np.abs(vel[:, 0] - ua).max() / ua.max())
        out["time_s"] = g["ptime"]
        v[name] = vel
        s.close()
    out["fp32c_rel_l2_velocity_error_vs_fp64"] = float(np.linalg.norm(v["fp32c"] - v["fp64"]) /
                                                       np.linalg.norm(v["fp64"]))
    return out


def dam_break(steps=1000):
    from . import scenes
    from .native import Scene, ARM_FP64, ARM_FP32C, F_V, F_RHO
    deck = scenes.dam_break()
    fluid = deck.dispbc == 0
    st = {}
    for name, arm in (("fp64", ARM_FP64), ("fp32c", ARM_FP32C)):
        s = Scene(deck, arm)
        assert s.Solution(steps) == steps
        st[name] = (s.field(F_V)[fluid], s.field(F_RHO)[fluid], s.globals()["ptime"])
        s.close()
    rho0 = deck.materials[0].rho
    (v64, r64, t64), (v32, r32, t32) = st["fp64"], st["fp32c"]
    return {"particles": deck.np, "steps": steps, "time_s": t64,
            "fp32c_rel_l2_velocity_error_vs_fp64": float(np.linalg.norm(v32 - v64) / np.linalg.norm(v64)),
            "fp32c_rel_l2_density_deviation_error_vs_fp64": float(np.linalg.norm(r32 - r64) /
                                                                  np.linalg.norm(r64 - rho0)),
            "fp32c_time_relative_difference": abs(t32 - t64) / t64}


def cylinder(steps=5000, every=250):
    from . import scenes
    from .native import Scene, ARM_FP64, ARM_FP32C
    import copy
    deck = scenes.cylinder()
    # the FP64 arm again from positions one ulp apart (x of every fluid
    # particle moved to the next double): how far rounding-level differences
    # carry in this flow, whatever the arithmetic
    nudged = copy.deepcopy(deck)
    fl = np.arange(deck.np) < deck.n_fluid
    nudged.x[fl, 0] = np.nextafter(deck.x[fl, 0], np.inf)
    sc = {"fp64": Scene(deck, ARM_FP64), "fp32c": Scene(deck, ARM_FP32C), "fp64_nudged": Scene(nudged, ARM_FP64)}

    def centre(g):
        return np.array([g["rbR0"], g["rbR1"]])

    # the body's starting centre: the unweighted mean of its particles
    # (MRigidBody.cpp:37-49), which the first step computes
    r0 = deck.x[deck.mat_id == deck.rb_id, :2].mean(axis=0)
    track = []
    for _ in range(steps // every):
        for s in sc.values():
            assert s.Solution(every) == every
        ga, gb, gn = sc["fp64"].globals(), sc["fp32c"].globals(), sc["fp64_nudged"].globals()
        va, vb = np.array([ga["rbV0"], ga["rbV1"]]), np.array([gb["rbV0"], gb["rbV1"]])
        vn = np.array([gn["rbV0"], gn["rbV1"]])
        track.append({"step": ga["timestep"] - 1, "t": ga["ptime"],
                      "displacement": float(np.linalg.norm(centre(ga) - r0)),
                      "dR": float(np.linalg.norm(centre(gb) - centre(ga))),
                      "dR_nudged": float(np.linalg.norm(centre(gn) - centre(ga))),
                      "V": float(np.linalg.norm(va)), "dV": float(np.linalg.norm(vb - va)),
                      "dV_nudged": float(np.linalg.norm(vn - va))})
    moved = track[-1]["displacement"]
    out = {"particles": deck.np, "steps": steps, "time_s": track[-1]["t"],
           "displacement_m": moved,
           "fp32c_max_trajectory_error_m": max(r["dR"] for r in track),
           "fp32c_trajectory_error_rel_final_displacement": max(r["dR"] for r in track) / max(moved, 1e-300),
           "fp32c_final_velocity_rel_error": track[-1]["dV"] / max(track[-1]["V"], 1e-300),
           "fp64_nudged_max_trajectory_difference_m": max(r["dR_nudged"] for r in track),
           "fp64_nudged_final_velocity_rel_difference": track[-1]["dV_nudged"] / max(track[-1]["V"], 1e-300),
           "track": track}
    for s in sc.values():
        s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", default="poiseuille,dam_break,cylinder")
    args = ap.parse_args()
    runs = {"poiseuille": poiseuille, "dam_break": dam_break, "cylinder": cylinder}
    out = {}
    for name in args.only.split(","):
        t0 = time.perf_counter()
        print(f"accuracy: {name}", file=sys.stderr, flush=True)
        out[name] = runs[name]()
        out[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        print(f"accuracy: {name}: {out[name]}", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
