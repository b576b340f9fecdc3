"""Collision-pass throughput on one MI355X.

Workload: one broadphase + narrowphase pass over 1,000,000 convex bodies
(40 % boxes, 30 % spheres, 20 % capsules, 10 % hulls of 16-64 points),
random orientations, packed uniformly at ~4 overlapping AABB pairs per body,
AABB margin 0.04 m.  A step = world AABBs -> sorted pair set -> contact
manifolds for every pair, bodies resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--bodies N]
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import platform
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STAGES = ["k_aabb", "tree_build", "k_traverse", "pair_list", "k_narrow", "k_epa", "contact_compaction"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(sc, pairs, offs, stage):
    """Algorithmic HBM bytes one launch of `stage` must move.

    k_aabb   per body: pose 28 B + shape record 16 B read, AABB 24 B written,
             plus 12 B per hull vertex read.
    k_narrow per pair: pair 8 B + both bodies' pose+shape 2 x 44 B read,
             count 4 B + 28 B per contact written, plus 12 B per hull
             vertex of each hull side read once.
    """
    t = sc["shape_type"]
    hc = sc["hull_count"].astype(np.int64)
    if stage == "k_aabb":
        return int(t.shape[0] * (28 + 16 + 24) + 12 * int(hc[t == 3].sum()))
    if stage == "k_narrow":
        P = pairs.shape[0]
        Cn = int(offs[-1])
        hv = int(hc[pairs[:, 0]].sum() + hc[pairs[:, 1]].sum())
        return int(P * (8 + 88 + 4) + 28 * Cn + 12 * hv)
    return None


def cpu_baseline(scenes, oracle_mod, sc, target_bodies, margin):
    """Single-thread oracle pass over the bodies inside a corner sub-cube."""
    side = float(sc["side"])
    n = sc["shape_type"].shape[0]
    f = (target_bodies / n) ** (1.0 / 3.0)
    m = np.all(sc["position"] < np.float32(side * f), axis=1)
    idx = np.nonzero(m)[0]
    sub = {k: sc[k][idx] for k in ("position", "orientation", "shape_type", "shape_params", "hull_first",
                                   "hull_count")}
    sub["hull_vertices"] = sc["hull_vertices"]
    P, Cn, sec = oracle_mod.collision_pass(sub, margin)
    return dict(value=P / sec, unit="pairs/s", cores=1, kind="port",
                sample=(f"single-thread C++ oracle pass (AABBs + exact sort-and-sweep pairs + collide_seeded per "
                        f"pair) over the {idx.shape[0]} bodies of the same scene inside a corner cube of side "
                        f"{side * f:.1f} m: {P} pairs, {Cn} contacts in {sec:.2f} s"),
                contacts_per_s=Cn / sec, seconds=sec, pairs=P)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bodies", type=int, default=1_000_000)
    ap.add_argument("--margin", type=float, default=0.04)
    ap.add_argument("--cpu-bodies", type=int, default=250_000)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    bpa = importlib.import_module("bullet-physics_amd")
    t0 = time.time()
    sc = bpa.scenes.mixed_convex(args.bodies, 2)
    log(f"scene: {args.bodies} bodies, cube side {sc['side']:.1f} m, generated in {time.time() - t0:.1f} s")
    w = bpa.CollisionWorld(sc)
    for k in range(args.warmup):
        P, Cn = w.collide(args.margin)
        log(f"warmup {k}: {P} pairs, {Cn} contacts")
    stage_ms = np.zeros(len(STAGES))
    w.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P, Cn = w.collide(args.margin)
        stage_ms += w.timings()
    w.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    stage_ms /= args.steps
    log("stages (ms): " + ", ".join(f"{s} {v:.3f}" for s, v in zip(STAGES, stage_ms)))
    pairs, offs, cont = w.results()
    value = P / (ms * 1e-3)

    dom = STAGES[int(np.argmax(stage_ms))]
    roof_stage = dom if dom in ("k_aabb", "k_narrow") else "k_narrow"
    ab = algorithmic_bytes(sc, pairs, offs, roof_stage)
    dur = stage_ms[STAGES.index(roof_stage)] * 1e-3
    achieved = ab / dur / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"{roof_stage}_mixed_1m_traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("bytes_per_launch")
    roof = {"bound": "hbm", "kernel": roof_stage, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "algorithmic_bytes": ab, "kernel_ms": round(dur * 1e3, 4), "dominant_stage": dom}

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cpu = cpu_baseline(bpa.scenes, O, sc, args.cpu_bodies, args.margin)
        try:
            cpu["cpu_model"] = subprocess.run(["lscpu"], capture_output=True, text=True).stdout.split(
                "Model name:")[1].splitlines()[0].strip()
        except Exception:
            cpu["cpu_model"] = platform.processor()
        log(f"cpu baseline: {cpu['value']:.0f} pairs/s ({cpu['sample']})")

    out = {
        "metric": "contact pairs/sec (broad+narrow, f32) on 1 MI355X; batched vs plain speedup",
        "value": round(value, 1), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded reference LCG)",
        "config": {"workload": f"mixed convex scene, {args.bodies} bodies (40% box, 30% sphere, 20% capsule, "
                               f"10% hull 16-64 pts), ~4 AABB pairs/body, margin {args.margin} m, one pass",
                   "bodies": args.bodies, "pairs": int(P), "contacts": int(Cn)},
        "contacts_per_s": round(Cn / (ms * 1e-3), 1),
        "stage_ms": {s: round(float(v), 4) for s, v in zip(STAGES, stage_ms)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
