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
STAGES = ["aabb", "tree_build", "tree_traverse", "pair_list", "narrow_fork", "narrow_gjk", "epa",
          "contact_compaction"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(sc, pairs, offs, routes, kernel):
    """Algorithmic HBM bytes one launch of `kernel` must move, each datum once:
    a body (and its hull) is charged once per launch however many of the
    launch's pairs touch it.

    Body record = pose 28 B + shape 16 B = 44 B; hull vertex = 12 B; contact
    = position 12 + normal 12 + depth 4 = 28 B.
    k_traverse  per body: its AABB 24 B read; per pair: 8 B written.
    k_np_hull   per pair with a hull: pair index 4 + pair 8 + count 4 + its
                band contacts 28 each, + 64 B queue entry for a pair sent on
                to EPA; per distinct body of those pairs: 44 + its vertices.
    k_epa_quad  per queued pair: queue entry 64 + pair 8 + count 4 + route 8
                (read, write) + 28 per contact; per distinct body of those
                pairs: 44 + its vertices (the ~2 % of entries it hands on to
                the two-pair and one-pair kernels are counted as if finished
                there: a 64-B queue entry written instead of the outputs).
    pass        the whole pass: N bodies x 44 + 12 x every hull vertex +
                P x (8 pair + 4 count) + 28 per contact.
    """
    P = pairs.shape[0]
    hc = sc["hull_count"].astype(np.int64)
    ncon = np.diff(offs.astype(np.int64))
    cls = routes & 7
    epa = (routes & 8) != 0

    def bodies(m):
        ids = np.unique(pairs[m].reshape(-1))
        return int(ids.shape[0] * 44 + 12 * hc[ids].sum())

    if kernel == "k_traverse":
        return int(sc["shape_type"].shape[0] * 24 + 8 * P)
    if kernel == "k_np_hull":
        m = cls == 6
        return int(m.sum() * (4 + 8 + 4) + 28 * ncon[m & ~epa].sum() + 64 * (m & epa).sum()) + bodies(m)
    if kernel == "k_epa_quad":
        return int(epa.sum() * (64 + 8 + 4 + 8) + 28 * ncon[epa].sum()) + bodies(epa)
    if kernel == "pass":
        return int(sc["shape_type"].shape[0] * 44 + 12 * hc.sum() + P * 12 + 28 * ncon.sum())
    return None


def cpu_baseline(scenes, oracle_mod, sc, target_bodies, margin, repeats=3):
    """The oracle's pass (AABBs, DbvtBroadPhase pairs, collide_seeded per pair)
    over the whole scene, or over the bodies inside a corner sub-cube when
    `target_bodies` is below the scene's count: one warm-up, then the median of
    `repeats` runs, single-threaded (the reference is single-threaded), and
    the same with the AABBs and the narrowphase on all the host's cores."""
    side = float(sc["side"])
    n = sc["shape_type"].shape[0]
    whole = target_bodies >= n
    f = 1.0 if whole else (target_bodies / n) ** (1.0 / 3.0)
    idx = np.arange(n) if whole else np.nonzero(np.all(sc["position"] < np.float32(side * f), axis=1))[0]
    region = (f"all {n} bodies of the scene" if whole else
              f"the {idx.shape[0]} bodies of the same scene inside a corner cube of side {side * f:.1f} m")
    sub = {k: sc[k][idx] for k in ("position", "orientation", "shape_type", "shape_params", "hull_first",
                                   "hull_count")}
    sub["hull_vertices"] = sc["hull_vertices"]
    runs = [oracle_mod.collision_pass(sub, margin) for _ in range(repeats + 1)][1:]
    P, Cn = runs[0][0], runs[0][1]
    sec = float(np.median([r[2] for r in runs]))
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    mt = [oracle_mod.collision_pass_mt(sub, margin, threads) for _ in range(repeats + 1)][1:]
    assert all(r[0] == P and r[1] == Cn for r in runs + mt)
    sec_mt = float(np.median([r[2] for r in mt]))
    return dict(value=P / sec, unit="pairs/s", cores=1, kind="port",
                sample=(f"single-thread C++ oracle pass (AABBs + DbvtBroadPhase pairs + collide_seeded per pair) "
                        f"over {region}: {P} pairs, {Cn} contacts; median of {repeats} runs after a warm-up, "
                        f"{sec:.2f} s each"),
                contacts_per_s=Cn / sec, seconds=sec, pairs=P,
                all_cores={"value": P / sec_mt, "cores": threads, "seconds": sec_mt,
                           "note": "AABBs and narrowphase on OpenMP threads, DBVT pairs serial"})


def captured(w, margin, steps, ref):
    """The same pass captured in a HIP graph (bch_world_capture): per step
    one graph launch plus the count read-back.  Checks that the replay's
    results equal the eager pass's `ref`."""
    caps = w.capture(margin, headroom=1.25)
    w.replay()
    w.replay_counts()
    got = w.results()
    same = all(np.array_equal(a, b) for a, b in zip(got, ref))
    w.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(steps, 1)):
        w.replay()
        P, Cn = w.replay_counts()
    secs = (time.perf_counter() - t0) / max(steps, 1)
    w.release_graph()
    return {"ms_per_step": round(secs * 1e3, 4), "pairs_per_s": round(P / secs, 1), "capacities": list(caps),
            "results_equal_eager": bool(same)}


def manifold_frames(w, sc, margin, frames):
    """The pair pass with persistent manifolds (bch_world_manifold_pass) on the
    same scene over `frames` frames: frame 0 detects every pair, later frames
    move a third of the bodies by ~4 mm and turn a tenth by ~0.01 rad, so the
    cache serves the pairs that stayed within the redetect bounds."""
    rng = np.random.default_rng(11)
    n = sc["shape_type"].shape[0]
    pos = np.ascontiguousarray(sc["position"], np.float32).copy()
    quat = np.ascontiguousarray(sc["orientation"], np.float32).copy()
    w.manifold_reset()
    w.set_poses(pos, quat)
    w.manifold_pass(margin)  # untimed: first use of the buffers
    w.manifold_reset()
    out = []
    for f in range(frames):
        if f:
            drift = rng.random(n) < 0.33
            pos[drift] += rng.normal(0.0, 0.004, (int(drift.sum()), 3)).astype(np.float32)
            turn = rng.random(n) < 0.1
            half = rng.normal(0.0, 0.005, int(turn.sum())).astype(np.float32)
            q = quat[turn]
            # rotate about y by 2 * half: (0, s, 0, c) * q
            c_, s_ = np.cos(half), np.sin(half)
            quat[turn] = np.stack([c_ * q[:, 0] + s_ * q[:, 2], c_ * q[:, 1] + s_ * q[:, 3],
                                   c_ * q[:, 2] - s_ * q[:, 0], c_ * q[:, 3] - s_ * q[:, 1]], 1)
        w.set_poses(pos, quat)
        w.synchronize()
        t0 = time.perf_counter()
        P, Np = w.manifold_pass(margin)
        ms = (time.perf_counter() - t0) * 1e3
        _, route, _, _ = w.manifold_results()
        rc = np.bincount(route, minlength=6)
        out.append({"frame": f, "ms": round(ms, 3), "pairs": int(P), "points": int(Np),
                    "served": int(rc[3]), "detected": int(rc[4]), "removed": int(rc[5])})
        log(f"manifold frame {f}: {out[-1]}")
    warm = [o["ms"] for o in out[1:]] or [out[0]["ms"]]
    return {"frames": out, "cold_ms": out[0]["ms"], "warm_ms_median": float(np.median(warm)),
            "warm_pairs_per_s": round(out[-1]["pairs"] / (float(np.median(warm)) * 1e-3), 1)}


def mesh_batch(bpa, steps, n_bodies=100_000, grid=362, cpu_sample=2000, cpu=True):
    """Convex bodies against a static terrain mesh (bch_collide_mesh_batch):
    `n_bodies` mixed convex bodies (boxes, spheres, capsules, hulls) near a
    grid x grid-vertex wavy terrain (2 (grid-1)^2 triangles, 0.25 m cells),
    one (body, mesh) pair each.  The call takes host buffers (bodies, pairs,
    poses in; contacts out), so the time includes those transfers.  The
    oracle's convex_mesh on `cpu_sample` of the pairs gives the CPU rate."""
    rng = np.random.default_rng(13)
    h = (grid - 1) * 0.25 / 2.0
    xs = np.linspace(-h, h, grid).astype(np.float32)
    X, Z = np.meshgrid(xs, xs)
    Y = (0.8 * np.sin(0.3 * X) * np.cos(0.25 * Z)).astype(np.float32)
    v = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    r = np.arange(grid - 1)
    a = (r[:, None] * grid + r[None, :]).ravel()
    tris = np.concatenate([np.stack([a, a + grid, a + 1], 1), np.stack([a + 1, a + grid, a + grid + 1], 1)])
    tris = tris.astype(np.uint32)
    ctx = bpa._Context(0)
    t0 = time.perf_counter()
    mesh = bpa.TriangleMesh(v, tris, ctx=ctx)
    build_s = time.perf_counter() - t0
    sc = bpa.scenes.mixed_convex(n_bodies, 17)
    keep = np.nonzero(sc["shape_type"] != 4)[0]
    x, z = rng.uniform(-h + 1, h - 1, (2, keep.shape[0]))
    pos = sc["position"].copy()
    pos[keep] = np.stack([x, 0.8 * np.sin(0.3 * x) * np.cos(0.25 * z) + rng.uniform(-0.3, 0.6, keep.shape[0]), z], 1)
    sc = dict(sc)
    sc["position"] = pos.astype(np.float32)
    pairs = np.stack([keep, np.full(keep.shape[0], mesh.id)], 1).astype(np.uint32)
    poses = np.tile(np.array([0, 0, 0, 0, 0, 0, 1], np.float32), (pairs.shape[0], 1))
    bpa.collide_mesh_batch(sc, pairs, poses, ctx=ctx)
    ts = []
    for _ in range(max(steps, 1)):
        t1 = time.perf_counter()
        cnt, pts = bpa.collide_mesh_batch(sc, pairs, poses, ctx=ctx)
        ts.append(time.perf_counter() - t1)
    ms = float(np.median(ts)) * 1e3
    out = {"workload": f"{pairs.shape[0]} convex bodies on a {tris.shape[0]}-triangle terrain",
           "bvh_build_s": round(build_s, 3), "ms_per_batch": round(ms, 3),
           "pairs_per_s": round(pairs.shape[0] / (ms * 1e-3), 1), "touching": int((cnt > 0).sum()),
           "points": int(cnt.sum()), "includes": "bodies, pairs and poses uploaded, contacts copied back"}
    if not cpu:
        ctx.close()
        return out
    try:  # CPU baseline: the oracle's convex_mesh on a sample
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        om = O.Mesh(v, tris)
        k = min(cpu_sample, pairs.shape[0])
        t1 = time.perf_counter()
        O.mesh_collide_scene(sc, [om], pairs[:k] * np.array([1, 0], np.uint32), poses[:k])
        cs = time.perf_counter() - t1
        out["cpu_pairs_per_s"] = round(k / cs, 1)
        out["cpu_sample"] = f"oracle convex_mesh, first {k} pairs, 1 thread"
    except Exception as e:  # the oracle is optional on a box without it built
        out["cpu_pairs_per_s"] = None
        out["cpu_error"] = str(e)[:120]
    ctx.close()
    return out


def sweeps(bpa, steps, n_bodies=100_000, n_sphere=262_144, n_cast=65_536, cpu_sample=1000, cpu=True):
    """CCD sweeps (bch_sweep_batch): `n_sphere` sphere sweeps (radius
    0.05-0.5 m, the CCD pass of step.rs:1141-1148) and `n_cast` convex casts
    of scene bodies (query.rs:210-218), each from 1-3 m away toward a random
    target of a mixed convex scene of `n_bodies` bodies.  Host buffers in and
    out, so the time includes the transfers.  The oracle on `cpu_sample`
    queries of each kind gives the CPU rate."""
    rng = np.random.default_rng(21)
    sc = bpa.scenes.mixed_convex(n_bodies, 23)
    ctx = bpa._Context(0)

    def queries(n):
        t = rng.integers(0, n_bodies, n).astype(np.uint32)
        c = sc["position"][t].astype(np.float64)
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        frm = (c + d * rng.uniform(1.0, 3.0, (n, 1))).astype(np.float32)
        to = (c + rng.normal(0, 0.3, (n, 3))).astype(np.float32)
        return t, frm, to

    st, sf, se = queries(n_sphere)
    rad = rng.uniform(0.05, 0.5, n_sphere).astype(np.float32)
    ct, cf, ce = queries(n_cast)
    caster = rng.integers(0, n_bodies, n_cast).astype(np.uint32)
    out = {"workload": f"{n_sphere} sphere sweeps + {n_cast} convex casts against a {n_bodies}-body mixed scene",
           "includes": "bodies and queries uploaded, hits copied back"}
    for tag, kw, (t, f, e), n in (("sphere", {"radius": rad}, (st, sf, se), n_sphere),
                                  ("cast", {"caster": caster}, (ct, cf, ce), n_cast)):
        hit, _ = bpa.sweep_batch(sc, t, f, e, ctx=ctx, **kw)
        ts = []
        for _ in range(max(steps, 1)):
            t1 = time.perf_counter()
            bpa.sweep_batch(sc, t, f, e, ctx=ctx, **kw)
            ts.append(time.perf_counter() - t1)
        ms = float(np.median(ts)) * 1e3
        out[tag] = {"ms_per_batch": round(ms, 3), "queries_per_s": round(n / (ms * 1e-3), 1), "hits": int(hit.sum())}
        if cpu:
            try:
                sys.path.insert(0, os.path.join(ROOT, "oracle"))
                import oracle as O
                k = min(cpu_sample, n)
                kk = {key: v[:k] for key, v in kw.items()}
                t1 = time.perf_counter()
                O.sweep_scene(sc, t[:k], f[:k], e[:k], **kk)
                out[tag]["cpu_queries_per_s"] = round(k / (time.perf_counter() - t1), 1)
                out[tag]["cpu_sample"] = f"oracle, first {k} queries, 1 thread"
            except Exception as ex:  # the oracle is optional on a box without it built
                out[tag]["cpu_queries_per_s"] = None
                out[tag]["cpu_error"] = str(ex)[:120]
    ctx.close()
    return out


def batched_vs_plain(bpa, n_worlds, margin, steps):
    """4096 worlds (16-2048 boxes/spheres + plane each, seed 3): one batched
    pass against the plain pass run world by world on the same GPU."""
    t0 = time.time()
    sc = bpa.scenes.batched_worlds(n_worlds, 3)
    log(f"batched scene: {n_worlds} worlds, {sc['shape_type'].shape[0]} bodies in {time.time() - t0:.1f} s")
    wb = bpa.CollisionWorld(sc)
    tuned = wb.autotune(margin, trials=3)
    Pb, Cb = wb.collide(margin)
    wb.synchronize()
    tb = []
    for _ in range(max(steps, 1)):
        t1 = time.perf_counter()
        Pb, Cb = wb.collide(margin)
        wb.synchronize()
        tb.append(time.perf_counter() - t1)
    tb = float(np.median(tb))
    graph = captured(wb, margin, steps, wb.results())
    log(f"batched captured pass: {graph}")
    # plain path: one world at a time through the same context (upload untimed)
    ctx = wb.ctx
    tp = 0.0
    Pp = Cp = 0
    for w in range(n_worlds):
        wp = bpa.CollisionWorld(bpa.scenes.world_slice(sc, w), ctx, batched=False)
        wp.collide(margin)
        t1 = time.perf_counter()
        P, Cn = wp.collide(margin)
        wp.synchronize()
        tp += time.perf_counter() - t1
        Pp += P
        Cp += Cn
    return {"workload": f"{n_worlds} worlds, 16-2048 boxes/spheres + plane each, {sc['shape_type'].shape[0]} bodies",
            "pairs": int(Pb), "contacts": int(Cb), "batched_ms": round(tb * 1e3, 3),
            "plain_per_world_ms": round(tp * 1e3, 3), "batched_pairs_per_s": round(Pb / tb, 1),
            "plain_pairs_per_s": round(Pp / tp, 1), "speedup": round(tp / tb, 2),
            "results_equal_counts": bool(Pp == Pb and Cp == Cb), "tuning": tuned, "captured": graph,
            "captured_speedup": round(tp / (graph["ms_per_step"] * 1e-3), 2)}


def differentiation(bpa, steps, n_pairs=262_144):
    """262,144 hull pairs (32 points each, half separated by U[0, 0.2] m, half
    penetrating by U[0, 0.2] m): signed distance forward and adjoint, device
    buffers.  The penetrating half is placed at its target depth by bisection
    along each pair's direction with this library's own signed distance
    (scenes.exact_penetration); the separated half is exact by construction."""
    import torch
    t0 = time.time()
    sc = bpa.scenes.hull_pairs(n_pairs, 4)
    w = bpa.CollisionWorld(sc, batched=False)
    w.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    pairs = torch.from_numpy(sc["pairs"].astype(np.int32)).cuda()
    P = n_pairs
    out = torch.empty((P, 10), device="cuda")
    st = torch.empty(P, dtype=torch.uint8, device="cuda")

    def sd_at(position):
        w.set_poses(position, sc["orientation"])
        w.sd_forward(pairs.data_ptr(), P, out.data_ptr(), st.data_ptr())
        return out[:, 0].cpu().numpy()

    sc = bpa.scenes.exact_penetration(sc, sd_at)
    w.set_poses(sc["position"], sc["orientation"])
    w.sd_forward(pairs.data_ptr(), P, out.data_ptr(), st.data_ptr())
    err = np.abs(out[:, 0].cpu().numpy().astype(np.float64) - sc["target"])
    placed = {"max_abs_sd_minus_target_m": float(err.max()), "unreached": sc["unreached"],
              "retargeted": sc["retargeted"],
              "note": "overlapping pairs whose line never reaches the drawn depth keep the deepest point of a "
                      "coarse scan (unreached), and pairs whose signed distance jumps across the crossing keep "
                      "where the bisection ends (retargeted), each with the signed distance there as target; the "
                      "max is over the separated pairs too, which include the few walks that cycle to the "
                      "iteration cap"}
    log(f"hull pairs: {n_pairs} in {time.time() - t0:.1f} s (placement {placed})")
    gen = torch.Generator(device="cuda").manual_seed(0)
    g10 = torch.randn((P, 10), device="cuda", generator=gen)
    g14 = torch.empty((P, 14), device="cuda")
    w.sd_forward(pairs.data_ptr(), P, out.data_ptr(), st.data_ptr())
    w.sd_backward(pairs.data_ptr(), P, g10.data_ptr(), g14.data_ptr(), None)
    torch.cuda.synchronize()
    k = max(steps, 1)
    t1 = time.perf_counter()
    for _ in range(k):
        w.sd_forward(pairs.data_ptr(), P, out.data_ptr(), st.data_ptr())
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(k):
        w.sd_backward(pairs.data_ptr(), P, g10.data_ptr(), g14.data_ptr(), None)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    fwd = (t2 - t1) / k
    bwd = (t3 - t2) / k
    status = np.bincount(st.cpu().numpy() & 7, minlength=6)
    w.ctx.set_stream(None)
    return {"workload": f"{P} convex-hull pairs, 32 points each, half separated by U[0,0.2] m, half penetrating "
                        f"by U[0,0.2] m", "placement": placed,
            "forward_ms": round(fwd * 1e3, 3), "backward_ms": round(bwd * 1e3, 3),
            "pairs_per_s_forward": round(P / fwd, 1), "pairs_per_s_forward_backward": round(P / (fwd + bwd), 1),
            "routes": {"walk": int(status[0]), "epa": int(status[1]), "capped": int(status[2]),
                       "separated_after_rounding": int(status[5])}}


def depenetration(bpa, steps=100):
    """1,024 worlds of 64-1,024 boxes and hulls: penetration penalty
    sum(max(0, -sd)^2) over the pairs of each world pass, gradient steps on
    all poses (broadphase, narrowphase signed distance and adjoint per step)."""
    import torch
    diff = importlib.import_module("bullet-physics_amd.diff")
    t0 = time.time()
    sc = bpa.scenes.batched_boxes_hulls(1024, 5)
    log(f"depenetration scene: {sc['shape_type'].shape[0]} bodies in {time.time() - t0:.1f} s")
    w = bpa.CollisionWorld(sc)
    pos = torch.from_numpy(sc["position"]).cuda()
    quat = torch.from_numpy(sc["orientation"]).cuda()
    diff.depenetrate(w, pos, quat, steps=1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _, _, hist = diff.depenetrate(w, pos, quat, steps=steps)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t1
    pairs = sum(h[2] for h in hist)
    # the same descent, each step one captured graph (history read every 25 steps)
    diff.depenetrate_graph(w, pos, quat, steps=1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _, _, hg = diff.depenetrate_graph(w, pos, quat, steps=steps, chunk=25)
    torch.cuda.synchronize()
    gsecs = time.perf_counter() - t1
    w.ctx.set_stream(None)
    gpairs = sum(h[2] for h in hg)
    return {"workload": f"1024 worlds, 64-1024 boxes and hulls each, {sc['shape_type'].shape[0]} bodies",
            "steps": steps, "ms_per_step": round(secs * 1e3 / steps, 3), "pairs_per_s": round(pairs / secs, 1),
            "loss": [round(hist[0][0], 6), hist[-1][0]], "max_penetration_m": [round(hist[0][1], 6), hist[-1][1]],
            "pairs_first_last": [hist[0][2], hist[-1][2]],
            "captured": {"ms_per_step": round(gsecs * 1e3 / steps, 3), "pairs_per_s": round(gpairs / gsecs, 1),
                         "loss": [round(hg[0][0], 6), hg[-1][0]],
                         "max_penetration_m": [round(hg[0][1], 6), hg[-1][1]],
                         "pairs_first_last": [hg[0][2], hg[-1][2]],
                         "includes": "capture (one uncaptured sizing step) + graph launches + a history read "
                                     "every 25 steps"}}


def valu_calibration(blocks=8192, rounds=2000, reps=5):
    """The VALU issue rate the hardware sustains: k_valu_probe (libbch_probe.so)
    runs 8 independent v_fma_f32 chains per lane, 128 FMAs per round, at 8
    blocks of 256 threads per CU.  Reported per SIMD-cycle at 2.4 GHz over
    1024 SIMDs, the same convention as roofline.valu."""
    import ctypes as C
    import torch
    sink = torch.empty(blocks * 256, device="cuda")
    L = C.CDLL(os.path.join(ROOT, "bullet-physics_amd", "libbch_probe.so"))
    L.bch_probe_valu.argtypes = [C.c_int, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_float)]
    ms = C.c_float()
    rc = L.bch_probe_valu(blocks, rounds, reps, C.c_void_p(sink.data_ptr()), C.byref(ms))
    assert rc == 0, "probe launch failed"
    fma = blocks * 4 * rounds * 128  # wave64 v_fma_f32 instructions per launch (4 waves per block)
    rate = fma / (1024 * ms.value * 1e-3 * 2.4e9)
    return {"probe": "k_valu_probe", "blocks": blocks, "rounds": rounds, "ms_per_launch": round(ms.value, 4),
            "fma_wave_instructions": fma, "per_simd_cycle_at_2.4GHz": round(rate, 4)}


def scene_boxes(sc, margin):
    """World AABBs of a scene's convex bodies (numpy, float64 rounded to f32):
    box corners |R| h, capsule segment ends +- r, hull vertex extremes,
    sphere +- r; padded by the margin.  Proxy boxes for the drop-in legs."""
    q = sc["orientation"].astype(np.float64)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], 1),
                  np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1),
                  np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)], 1)
    t = sc["shape_type"]
    prm = sc["shape_params"].astype(np.float64)
    half = np.zeros((t.shape[0], 3))
    m = t == 1
    half[m] = np.einsum("nij,nj->ni", np.abs(R[m]), prm[m, :3])
    m = t == 0
    half[m] = prm[m, :1]
    m = t == 2
    half[m] = np.abs(R[m][:, :, 1]) * prm[m, 1:2] + prm[m, :1]
    pos = sc["position"].astype(np.float64)
    lo, hi = pos - half, pos + half
    hull = np.nonzero(t == 3)[0]
    if hull.size:
        first = sc["hull_first"][hull].astype(np.int64)
        cnt = sc["hull_count"][hull].astype(np.int64)
        owner = np.repeat(np.arange(hull.size), cnt)
        vidx = np.concatenate([np.arange(f, f + c) for f, c in zip(first, cnt)])
        v = np.einsum("nij,nj->ni", R[hull][owner], sc["hull_vertices"][vidx].astype(np.float64))
        starts = np.concatenate(([0], np.cumsum(cnt)[:-1]))
        lo[hull] = pos[hull] + np.minimum.reduceat(v, starts, axis=0)
        hi[hull] = pos[hull] + np.maximum.reduceat(v, starts, axis=0)
    return (lo - margin).astype(np.float32), (hi + margin).astype(np.float32)


def dropin_broadphase(bpa, sc, margin, steps, moved=0.3):
    """The drop-in BroadPhase (bch_bp_*) over the scene's bodies as proxies:
    per step set_aabbs on a random 30 % of them (moved by ~5 mm), then
    compute_pairs into a host list, as the Rust world's update_aabbs +
    compute_pairs_into would call it.  Host wall time per step."""
    lo, hi = scene_boxes(sc, margin)
    n = lo.shape[0]
    bp = bpa.BroadPhase()
    t0 = time.perf_counter()
    for i in range(n):
        bp.create_proxy(lo[i], hi[i])
    create_s = time.perf_counter() - t0
    pairs = bp.compute_pairs()
    rng = np.random.default_rng(3)
    k = int(moved * n)
    sets, tc, ts = [], [], []
    for _ in range(max(steps, 1) + 1):
        idx = rng.choice(n, k, replace=False).astype(np.uint32)
        d = rng.normal(0.0, 0.005, (k, 3)).astype(np.float32)
        sets.append((idx, lo[idx] + d, hi[idx] + d))
    for s_ in range(len(sets)):
        idx, a, b = sets[s_]
        t1 = time.perf_counter()
        bp.set_aabbs(idx, a, b)
        t2 = time.perf_counter()
        pairs = bp.compute_pairs()
        t3 = time.perf_counter()
        if s_:  # the first round is a warm-up
            ts.append(t2 - t1)
            tc.append(t3 - t2)
    step = float(np.median(np.add(ts, tc)))
    bp.ctx.close()
    return {"workload": f"{n} proxies (the scene's AABBs), {k} moved by set_aabbs per step",
            "pairs": int(pairs.shape[0]), "ms_per_step": round(step * 1e3, 3),
            "set_aabbs_ms": round(float(np.median(ts)) * 1e3, 3),
            "compute_pairs_ms": round(float(np.median(tc)) * 1e3, 3),
            "pairs_per_s": round(pairs.shape[0] / step, 1), "create_proxies_s": round(create_s, 2),
            "includes": "host mirror updates, upload and device scatter of the moved slots, broadphase, "
                        "pairs copied into a host array"}


def dropin_collide_batch(bpa, sc, pairs, steps):
    """bch_collide_batch over the pass's pairs with host buffers both ways
    (bodies + pairs in, 4 x 7-float points per pair out), and the pose-only
    bch_collide_batch_poses that a world step would call after the first."""
    import ctypes as C
    ctx = bpa._Context(0)
    L, h = ctx.L, ctx.h
    b = bpa.Bodies(sc)
    pr = np.ascontiguousarray(pairs, np.uint32)
    P = pr.shape[0]
    # output arrays reused across steps, as a caller's buffers would be
    cnt, pts = np.zeros(P, np.uint8), np.zeros((P, 4, 7), np.float32)
    cnt2, pts2 = np.zeros(P, np.uint8), np.zeros((P, 4, 7), np.float32)
    full = lambda: L.bch_collide_batch(h, C.byref(b.c), pr.ctypes.data, P, None, cnt.ctypes.data, pts.ctypes.data)
    poses = lambda: L.bch_collide_batch_poses(h, b.pos.ctypes.data, b.quat.ctypes.data, pr.ctypes.data, P, None,
                                              cnt2.ctypes.data, pts2.ctypes.data)
    assert full() == 0 and poses() == 0
    tf, tp = [], []
    for _ in range(max(min(steps, 5), 1)):
        t1 = time.perf_counter()
        assert full() == 0
        tf.append(time.perf_counter() - t1)
        t1 = time.perf_counter()
        assert poses() == 0
        tp.append(time.perf_counter() - t1)
    ctx.close()
    f, p = float(np.median(tf)), float(np.median(tp))
    return {"workload": f"{P} caller pairs of the scene", "contacts": int(cnt.sum()),
            "results_equal": bool(np.array_equal(cnt, cnt2) and np.array_equal(pts, pts2)),
            "full_ms": round(f * 1e3, 3), "full_pairs_per_s": round(P / f, 1),
            "poses_ms": round(p * 1e3, 3), "poses_pairs_per_s": round(P / p, 1),
            "includes": "host arrays in (bodies or poses, pairs), 113 B per pair out over PCIe"}


def gradient_check(bpa, oracle_mod, n_worlds=16, directions=16, h=1e-4):
    """The de-penetration gradient (worlds of the de-penetration scene's distributions) against
    f64 central differences of the same penalty evaluated by the oracle (its own
    broadphase and f64 signed distance), along directions that move three bodies."""
    import torch
    diff = importlib.import_module("bullet-physics_amd.diff")
    sc = bpa.scenes.batched_boxes_hulls(n_worlds, 5)
    w = bpa.CollisionWorld(sc)
    pos = torch.from_numpy(sc["position"]).cuda().requires_grad_(True)
    quat = torch.from_numpy(sc["orientation"]).cuda().requires_grad_(True)
    loss, pen, pairs = diff.penetration_loss(w, pos, quat)
    loss.backward()
    gp = pos.grad.cpu().numpy().astype(np.float64)
    gq = quat.grad.cpu().numpy().astype(np.float64)
    with torch.no_grad():
        _, n32, _, _, _ = diff.signed_distance(w, pos, quat, pairs)
    n32 = n32.cpu().numpy().astype(np.float64)
    pr32 = pairs.cpu().numpy().astype(np.int64)
    pen32 = pen.detach().cpu().numpy()
    w.ctx.set_stream(None)

    def loss64(p, q):
        s2 = dict(sc)
        s2["position"] = p.astype(np.float32)
        s2["orientation"] = q.astype(np.float32)
        pr = oracle_mod.world_pairs(s2, 0.04)
        i = pr.astype(np.int64)
        pose = np.concatenate([p[i[:, 0]], q[i[:, 0]], p[i[:, 1]], q[i[:, 1]]], axis=1)
        s2["pairs"] = pr
        out, _ = oracle_mod.sd_pairs(s2, pr, pose, f64=True)
        e = np.maximum(-out[:, 0], 0.0)
        return float(np.sum(e * e))

    p0, q0 = sc["position"].astype(np.float64), sc["orientation"].astype(np.float64)
    l0 = loss64(p0, q0)
    # the f64 normals of the same pairs: where EPA's f32 and f64 runs stop on
    # different faces (two directions of near-equal depth, both within EPA's
    # 1e-4 exit), each gradient is that of its own face
    s0 = dict(sc)
    s0["pairs"] = pr32.astype(np.uint32)
    o64, _ = oracle_mod.sd_pairs(s0, s0["pairs"], np.concatenate(
        [p0[pr32[:, 0]], q0[pr32[:, 0]], p0[pr32[:, 1]], q0[pr32[:, 1]]], axis=1), f64=True)
    ang = np.arctan2(np.linalg.norm(np.cross(n32, o64[:, 1:4]), axis=1), np.sum(n32 * o64[:, 1:4], axis=1))
    rng = np.random.default_rng(8)
    rel, nonsmooth, face_tie = [], 0, []
    for _ in range(directions):
        vp, vq = np.zeros_like(p0), np.zeros_like(q0)
        bs = rng.choice(p0.shape[0], 3, replace=False)
        for b in bs:
            vp[b] = rng.normal(size=3)
            vq[b] = rng.normal(size=4)
        touched = (np.isin(pr32[:, 0], bs) | np.isin(pr32[:, 1], bs)) & (pen32 > 0)
        if touched.any() and ang[touched].max() > 1e-2:
            face_tie.append(round(float(ang[touched].max()), 4))
            continue
        g = float(np.sum(gp * vp) + np.sum(gq * vq))
        fd, jump = [], []
        for hh in (h, h / 2):
            lp, lm = loss64(p0 + hh * vp, q0 + hh * vq), loss64(p0 - hh * vp, q0 - hh * vq)
            fd.append((lp - lm) / (2 * hh))
            jump.append(abs(lp - 2 * l0 + lm) / hh)
        scale = max(abs(fd[0]), 1e-3)
        ratio = jump[1] / max(jump[0], 1e-300)
        if jump[0] <= 1e-3 * scale or 0.3 <= ratio <= 0.7:
            # the h difference, or its Richardson extrapolation with h / 2 where
            # the difference's own truncation error dominates
            rel.append(min(abs(g - fd[0]), abs(g - (4 * fd[1] - fd[0]) / 3)) / scale)
        else:
            nonsmooth += 1
    return {"workload": f"{n_worlds} worlds of boxes and hulls (the de-penetration distributions), "
                        f"{int(pairs.shape[0])} pairs, {int((pen > 0).sum())} penetrating",
            "loss_f32": float(loss.detach()), "loss_f64": l0, "directions": directions, "smooth": len(rel),
            "non_smooth_within_h": nonsmooth,
            "epa_face_ties": {"directions": len(face_tie), "normal_angle_rad": face_tie,
                              "pairs_over_1e-2_rad": int(((pen32 > 0) & (ang > 1e-2)).sum()),
                              "penetrating_pairs": int((pen32 > 0).sum()),
                              "note": "directions moving a body of a pair whose f32 (GPU) and f64 EPA normals differ "
                                      "by more than 1e-2 rad: the two runs stopped on different faces of near-equal "
                                      "depth, so their gradients are those of different faces; not held to the bar"},
            "relative_error_median": float(np.median(rel)) if rel else None,
            "relative_error_max": float(np.max(rel)) if rel else None,
            "within_1e-3": int(sum(r <= 1e-3 for r in rel)),
            "h": h, "note": "f64 central differences of the oracle's penalty along directions moving 3 bodies (the "
                            "better of the h difference and its Richardson extrapolation with h/2); a direction "
                            "is non-smooth when its second difference does not halve with h. The oracle's f64 "
                            "adjoint meets these differences to ~1e-7 in every direction, so what remains is the "
                            "f32 forward's rounding"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bodies", type=int, default=1_000_000)
    ap.add_argument("--margin", type=float, default=0.04)
    ap.add_argument("--cpu-bodies", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--only-pass", action="store_true",
                    help="skip the host-buffer, captured and manifold legs (for kernel traces of the pass alone)")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--tuning", default="",
                    help="fixed launch knobs, e.g. hull_lanes=8,epa_grid=65536,np_grid=4096, instead of autotuning")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device-to-device copy three times (HBM counter calibration)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only run the VALU issue-rate probe (independent FMA chains, full occupancy)")
    args = ap.parse_args()
    if args.valu_calibration:
        print(json.dumps(valu_calibration()), flush=True)
        return
    if args.copy_calibration:
        import torch
        n = 64 * 1024 * 1024
        x = torch.ones(n, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"copied {4 * n} bytes three times")
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    bpa = importlib.import_module("bullet-physics_amd")
    t0 = time.time()
    sc = bpa.scenes.mixed_convex(args.bodies, 2)
    log(f"scene: {args.bodies} bodies, cube side {sc['side']:.1f} m, generated in {time.time() - t0:.1f} s")
    w = bpa.CollisionWorld(sc)
    if args.tuning:
        # fixed launch knobs instead of the autotuner (profiling runs that
        # want every pass at the same launch shapes)
        knobs = {k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))}
        w.ctx.set_tuning(**knobs)
        w.collide(args.margin)  # sizes the pair buffers
        tuned = dict(w.ctx.tuning(), fixed=True)
    else:
        tuned = w.autotune(args.margin, trials=3)
    log(f"tuning: {tuned}")
    for k in range(args.warmup):
        P, Cn = w.collide(args.margin)
        log(f"warmup {k}: {P} pairs, {Cn} contacts")
    stage_ms = np.zeros(len(STAGES))
    epa_quad_ms = 0.0
    np_hull_ms = 0.0
    w.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        P, Cn = w.collide(args.margin)
        t_all = w.timings_all()
        stage_ms += t_all[:8]
        epa_quad_ms += float(t_all[8])
        np_hull_ms += float(t_all[9])
    w.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    stage_ms /= args.steps
    epa_quad_ms /= args.steps
    np_hull_ms /= args.steps
    log("stages (ms): " + ", ".join(f"{s} {v:.3f}" for s, v in zip(STAGES, stage_ms)))
    pairs, offs, cont = w.results()
    value = P / (ms * 1e-3)

    # the same pass with host buffers on both sides: poses uploaded, results
    # (pairs, offsets, contacts) copied back -- reported, never `value`
    pos = np.ascontiguousarray(sc["position"], np.float32)
    quat = np.ascontiguousarray(sc["orientation"], np.float32)
    th = []
    for _ in range(0 if args.only_pass else 3):
        t1 = time.perf_counter()
        w.set_poses(pos, quat)
        w.collide(args.margin)
        w.results()
        th.append(time.perf_counter() - t1)
    host_ms = float(np.median(th)) * 1e3 if th else float("nan")
    pcie = None if not th else {"ms_per_step": round(host_ms, 3), "pairs_per_s": round(P / (host_ms * 1e-3), 1),
            "includes": "pose upload (28 B/body) + pass + pairs, offsets and contacts copied to host"}
    log(f"host-buffer pass: {host_ms:.2f} ms")
    w.set_poses(pos, quat)
    graph = None if args.only_pass else captured(w, args.margin, args.steps, (pairs, offs, cont))
    log(f"captured pass: {graph}")
    routes = w.routes()
    dom = STAGES[int(np.argmax(stage_ms))]
    # single-kernel times from the pass's own HIP events: the traversal stage
    # is k_traverse alone, event 9 ends k_epa_quad inside the EPA stage and
    # event 10 ends k_np_hull (on the pass's stream, beside the side stream's
    # k_np launches) inside the narrowphase stage
    kernel_ms = {"k_traverse": float(stage_ms[STAGES.index("tree_traverse")]), "k_epa_quad": epa_quad_ms,
                 "k_np_hull": np_hull_ms}
    kern = max(kernel_ms, key=kernel_ms.get)
    ab = algorithmic_bytes(sc, pairs, offs, routes, kern)
    dur = kernel_ms[kern] * 1e-3
    achieved = ab / dur / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"{kern}_mixed_1m_traffic.json")
    if os.path.exists(tf):
        traffic = json.load(open(tf)).get("bytes_per_launch")
    roof = {"bound": "hbm", "kernel": kern, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
            "algorithmic_bytes": ab, "kernel_ms": round(dur * 1e3, 4), "dominant_stage": dom,
            "routes": {"epa_pairs": int(((routes & 8) != 0).sum()), "hull_pairs": int(((routes & 7) == 6).sum())}}
    pb = algorithmic_bytes(sc, pairs, offs, routes, "pass")
    roof["pass"] = {"algorithmic_bytes": pb, "ms": round(ms, 4), "achieved": round(pb / (ms * 1e-3) / 1e9, 2),
                    "frac": round(pb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                    "note": "the whole pass's bytes (each body, hull vertex, pair, count and contact once) over "
                            "the pass time"}
    # The kernel's real bound is instruction issue: its VALU wave-instructions
    # per launch (SQ counters, profiles/<kernel>_mixed_1m_valu.json) over
    # 1024 SIMDs x its own event time x 2.4 GHz, against the rate the FMA
    # probe sustains (profiles/valu_probe_calibration.json).
    vf = os.path.join(ROOT, "profiles", f"{kern}_mixed_1m_valu.json")
    if os.path.exists(vf):
        vi = json.load(open(vf))["valu_instructions_per_dispatch"]
        rate = vi / (1024 * dur * 2.4e9)
        cal = os.path.join(ROOT, "profiles", "valu_probe_calibration.json")
        ceil = json.load(open(cal))["per_simd_cycle_at_2.4GHz"] if os.path.exists(cal) else None
        roof["valu"] = {"instructions": vi, "achieved_per_simd_cycle": round(rate, 4),
                        "peak_per_simd_cycle": ceil, "frac": round(rate / ceil, 3) if ceil else None,
                        "note": "instructions from the SQ counters of the same kernel and workload over this "
                                "run's own event time of that kernel; the peak is the rate the FMA probe sustains "
                                "(bench.py --valu-calibration), both per SIMD-cycle at 2.4 GHz"}

    dropin = None
    if not args.only_pass:
        dropin = {"broadphase": dropin_broadphase(bpa, sc, args.margin, min(args.steps, 5)),
                  "collide_batch": dropin_collide_batch(bpa, sc, pairs, args.steps)}
        log(f"drop-in entry points: {dropin}")

    # last: it moves the bodies
    manifold = None if args.only_pass else manifold_frames(w, sc, args.margin, 5)

    batched = None
    if not args.no_batched:
        batched = batched_vs_plain(bpa, args.worlds, args.margin, args.steps)
        log(f"batched: {batched}")

    mesh = None if args.only_pass else mesh_batch(bpa, args.steps, cpu=not args.no_cpu)
    if mesh:
        log(f"mesh batch: {mesh}")
    sweep = None if args.only_pass else sweeps(bpa, args.steps, cpu=not args.no_cpu)
    if sweep:
        log(f"sweeps: {sweep}")

    diff_out = None
    if not args.no_diff:
        diff_out = {"signed_distance": differentiation(bpa, args.steps), "depenetration": depenetration(bpa)}
        log(f"differentiation: {diff_out}")

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
        log(f"cpu baseline: {cpu['value']:.0f} pairs/s, all cores {cpu['all_cores']['value']:.0f} ({cpu['sample']})")
        if diff_out is not None:
            # the oracle as the checker of the de-penetration gradient
            diff_out["gradient_check"] = gradient_check(bpa, O)
            log(f"gradient check: {diff_out['gradient_check']}")

    out = {
        "metric": "contact pairs/sec (broad+narrow, f32) on 1 MI355X; batched vs plain speedup",
        "value": round(value, 1), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded reference LCG)",
        "config": {"workload": f"mixed convex scene, {args.bodies} bodies (40% box, 30% sphere, 20% capsule, "
                               f"10% hull 16-64 pts), ~4 AABB pairs/body, margin {args.margin} m, one pass",
                   "bodies": args.bodies, "pairs": int(P), "contacts": int(Cn), "tuning": tuned},
        "contacts_per_s": round(Cn / (ms * 1e-3), 1),
        "stage_ms": {s: round(float(v), 4) for s, v in zip(STAGES, stage_ms)},
        "kernel_ms": {k: round(v, 4) for k, v in kernel_ms.items()},
        "roofline": roof,
        "host_buffers": pcie,
        "dropin": dropin,
        "captured": graph,
        "manifold": manifold,
        "mesh": mesh,
        "sweeps": sweep,
        "batched": batched,
        "differentiation": diff_out,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
