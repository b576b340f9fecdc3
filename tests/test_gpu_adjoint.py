"""Signed distance and its adjoint on the GPU against the oracle.

Forward: the HIP signed distance (walk with an unbounded band, EPA for
overlapping cores) against the oracle's f32 restatement of the same walks,
within the f32 tolerance of the collision path (1e-5 m, 1e-4 rad).

Backward: the hand-written adjoint against float64 central finite
differences (h = 1e-4) of the oracle's f64 signed distance, as
vector-Jacobian products with random output weights.  The bar is 1e-3
relative error per pair.  Pairs where the function has a kink within the
step (an active support feature changes: the one-sided differences disagree)
are classified as such and reported, not held to the bar.
"""
import importlib

import numpy as np
import pytest
from sd_helpers import fd as _fd, witness_margin as _witness_margin

pytestmark = pytest.mark.gpu

TOL_M = 1e-5
TOL_RAD = 1e-4
REL = 1e-3
H = 1e-4


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert t.cuda.is_available()
    return t


@pytest.fixture(scope="module")
def diff():
    return importlib.import_module("bullet-physics_amd.diff")


@pytest.fixture(autouse=True)
def _on_torch_stream(torch, ctx):
    """The library works on torch's stream here, so the tensors torch fills
    (pair lists, upstream gradients) are ordered before the kernels read
    them."""
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    yield
    ctx.set_stream(None)


def _gpu_sd(torch, bpa, sc, ctx):
    w = bpa.CollisionWorld(sc, ctx)
    pairs = torch.from_numpy(sc["pairs"].astype(np.int32)).cuda().contiguous()
    P = pairs.shape[0]
    out = torch.empty((P, 10), dtype=torch.float32, device="cuda")
    st = torch.empty(P, dtype=torch.uint8, device="cuda")
    w.sd_forward(pairs.data_ptr(), P, out.data_ptr(), st.data_ptr())
    w.synchronize()
    return w, pairs, out.cpu().numpy(), st.cpu().numpy()


def _angle(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return np.arctan2(np.linalg.norm(np.cross(a, b), axis=1), np.sum(a * b, axis=1))


@pytest.mark.parametrize("mix", [False, True])
def test_sd_forward_matches_oracle(torch, bpa, scenes, oracle, ctx, mix):
    _flush(f"case: sd forward, hull pairs, shape_mix={mix}")
    sc = scenes.hull_pairs(2048 if not mix else 1024, 4 if not mix else 5, shape_mix=mix)
    _, _, out, st = _gpu_sd(torch, bpa, sc, ctx)
    ref, rst = oracle.sd_pairs(sc, sc["pairs"])
    assert np.array_equal(st & 7, rst), f"status differs at {np.nonzero((st & 7) != rst)[0][:10]}"
    ok = rst != oracle.SD_TOUCH
    dsd = np.abs(out[ok, 0].astype(np.float64) - ref[ok, 0])
    dpa = np.linalg.norm(out[ok, 4:7].astype(np.float64) - ref[ok, 4:7], axis=1)
    dpb = np.linalg.norm(out[ok, 7:10].astype(np.float64) - ref[ok, 7:10], axis=1)
    ang = _angle(out[ok, 1:4], ref[ok, 1:4])
    exact = int(np.sum(np.all(out == ref, axis=1)))
    _flush(f"sd forward: {out.shape[0]} pairs ({np.bincount(rst, minlength=6)} walk/epa/capped/touch/unconv/separated), "
           f"{exact} bit-identical, max |dsd| {dsd.max():.3g}, |dpa| {dpa.max():.3g}, |dpb| {dpb.max():.3g}, "
           f"angle {ang.max():.3g}")
    assert dsd.max() <= TOL_M and dpa.max() <= TOL_M and dpb.max() <= TOL_M and ang.max() <= TOL_RAD
    # separated pairs of the construction sit exactly at their target distance
    sep = np.arange(out.shape[0]) % 2 == 0
    assert np.abs(out[sep, 0] - sc["target"][sep]).max() < 1e-4


def _pad_hulls(sc, count_a, count_b):
    """The scene with A's hulls padded to count_a vertices and B's to count_b
    by appending copies of their own vertices (in order, cycling)."""
    nb = sc["shape_type"].shape[0]
    verts, first, count = [], np.zeros(nb, np.uint32), np.zeros(nb, np.uint32)
    at = 0
    for b in range(nb):
        v = sc["hull_vertices"][sc["hull_first"][b]: sc["hull_first"][b] + sc["hull_count"][b]]
        n = max(count_a if b % 2 == 0 else count_b, v.shape[0])
        v = v[np.arange(n) % v.shape[0]]
        first[b], count[b] = at, n
        verts.append(v)
        at += n
    out = dict(sc)
    out["hull_vertices"] = np.concatenate(verts).astype(np.float32)
    out["hull_first"], out["hull_count"] = first, count
    return out


def test_sd_hull_staging_leaves_results_unchanged(torch, bpa, scenes, oracle, ctx):
    """The walk stages hull vertices in one 64-vertex LDS slot per pair: both
    hulls when they fit together, else A alone or B alone (<= 64 vertices),
    else neither (both read through the cache).  Padding a hull with copies
    of its own vertices changes which of these happens but no support point
    (of equal projections the first vertex wins, and the copies come after
    the originals), so status, all ten forward outputs and the fourteen pose
    gradients of the backward must be bit-identical across the four
    stagings.  The unpadded scene meets the oracle."""
    _flush("case: sd hull staging, 24-point hulls padded to 24/48/70")
    base = scenes.hull_pairs(256, 9, npts=24, spacing=0.0)
    P = base["pairs"].shape[0]
    g10 = torch.from_numpy(np.random.default_rng(11).normal(size=(P, 10)).astype(np.float32)).cuda().contiguous()
    results = {}
    for ca, cb in ((24, 24), (24, 48), (70, 24), (70, 70)):
        sc = _pad_hulls(base, ca, cb)
        w, pairs, out, st = _gpu_sd(torch, bpa, sc, ctx)
        g14 = torch.empty((P, 14), dtype=torch.float32, device="cuda")
        stb = torch.empty(P, dtype=torch.uint8, device="cuda")
        w.sd_backward(pairs.data_ptr(), P, g10.data_ptr(), g14.data_ptr(), stb.data_ptr())
        w.synchronize()
        results[(ca, cb)] = (out, st, g14.cpu().numpy(), stb.cpu().numpy())
    out0, st0, g0, stb0 = results[(24, 24)]
    ref, rst = oracle.sd_pairs(base, base["pairs"])
    assert np.array_equal(st0 & 7, rst)
    ok = rst != oracle.SD_TOUCH
    assert np.abs(out0[ok, 0].astype(np.float64) - ref[ok, 0]).max() <= TOL_M
    assert np.isfinite(g0).all()
    for key, (out, st, g14, stb) in results.items():
        _flush(f"  A {key[0]} / B {key[1]} vertices: {np.bincount(st & 7, minlength=6)} walk/epa/capped/touch/"
               f"unconv/separated, forward equal {np.array_equal(out, out0)}, gradients equal "
               f"{np.array_equal(g14, g0)}")
        assert np.array_equal(st, st0) and np.array_equal(stb, stb0)
        assert np.array_equal(out, out0) and np.array_equal(g14, g0)


@pytest.mark.parametrize("mix", [False, True])
def test_sd_adjoint_vs_f64_finite_differences(torch, bpa, scenes, oracle, ctx, mix):
    """Every pair is held to 1e-3 relative against f64 central differences
    at h = 1e-4 for the signed distance and the normal; the witness points
    too where A's support along the normal is unique by a margin (the
    reference's witness supportA(n) jumps between tied vertices of a face or
    edge contact).  An exceedance must be non-smooth within +-h in some pose
    coordinate (its second difference does not halve with h: a support
    feature changes inside the step), or the h = 1e-4 difference itself is
    off by its truncation error (the Richardson extrapolation of the h and
    h/2 differences is within the bar)."""
    _flush(f"case: sd adjoint vs f64 FD, shape_mix={mix}")
    sc = scenes.hull_pairs(192 if not mix else 128, 6 if not mix else 7, shape_mix=mix, spacing=0.0)
    w, pairs, out, st = _gpu_sd(torch, bpa, sc, ctx)
    P = pairs.shape[0]
    rng = np.random.default_rng(3)
    g10 = rng.normal(size=(P, 10))
    margin = _witness_margin(sc, out[:, 1:4].astype(np.float64))
    for label, gw, sel in (("sd+normal", np.concatenate([g10[:, :4], np.zeros((P, 6))], 1), np.ones(P, bool)),
                           ("all outputs", g10, margin > 1e-3)):
        gin = torch.from_numpy(gw.astype(np.float32)).cuda().contiguous()
        g14 = torch.empty((P, 14), dtype=torch.float32, device="cuda")
        stb = torch.empty(P, dtype=torch.uint8, device="cuda")
        w.sd_backward(pairs.data_ptr(), P, gin.data_ptr(), g14.data_ptr(), stb.data_ptr())
        w.synchronize()
        g14 = g14.cpu().numpy().astype(np.float64)
        stb = stb.cpu().numpy()
        vjp, j1, f0, st0 = _fd(oracle, sc, gw, H)
        vjp2, j2, _, _ = _fd(oracle, sc, gw, H / 2)
        scale = np.maximum(np.linalg.norm(vjp, axis=1), 1e-2)
        err = np.linalg.norm(g14 - vjp, axis=1) / scale
        # Richardson extrapolation of the two central differences: where the
        # h = 1e-4 difference is off by its own truncation error (strongly
        # curved pieces), the extrapolated one is not
        err_rich = np.linalg.norm(g14 - (4 * vjp2 - vjp) / 3, axis=1) / scale
        ratio = j2 / np.maximum(j1, 1e-300)
        smooth = np.all((j1 <= 1e-3 * scale[:, None]) | ((ratio >= 0.3) & (ratio <= 0.7)), axis=1)
        flagged = (stb & (8 | 16)) != 0
        exceed = sel & (err > REL)
        truncation = exceed & smooth & (err_rich <= REL)
        unexplained = exceed & smooth & ~flagged & ~truncation
        within = sel & ~exceed
        # the f64 reference gradient (the same adjoint restated in the oracle)
        ref, _ = oracle.sd_vjp_pairs(sc, sc["pairs"], gw)
        err_ref = np.linalg.norm(g14 - ref, axis=1) / scale
        ref_bad = sel & smooth & ~flagged & (err_ref > REL)
        _flush(f"adjoint [{label}]: {sel.sum()} pairs held, {within.sum()} within 1e-3 of f64 FD (max "
               f"{err[within].max():.3g}, median {np.median(err[sel]):.3g}); {exceed.sum()} beyond: "
               f"{(exceed & ~smooth).sum()} non-smooth within +-h, {(exceed & flagged).sum()} flagged, "
               f"{truncation.sum()} difference truncation (Richardson within 1e-3), "
               f"{unexplained.sum()} unexplained; vs f64 reference gradient: median {np.median(err_ref[sel]):.3g}, "
               f"max on smooth pairs {err_ref[sel & smooth].max():.3g}")
        assert not unexplained.any(), f"unexplained: {np.nonzero(unexplained)[0][:10]} errs {err[unexplained][:10]}"
        assert not ref_bad.any(), f"beyond 1e-3 of the f64 reference: {np.nonzero(ref_bad)[0][:10]}"
        assert within.sum() >= 0.9 * sel.sum()
        assert sel.sum() >= 0.4 * P


def test_sd_gradient_wrt_positions_is_the_normal(torch, bpa, scenes, oracle, ctx, diff):
    """d sd / d p_B = n and d sd / d p_A = -n (translation along the normal).
    For walk pairs n is the exact (f64) normal: the f32 forward normal of
    near-touching cores carries the rounding the adjoint avoids.  For EPA
    pairs it is the normal of the face the pass reported (EPA stops within
    its 1e-4 tolerance, so that face can differ from the f64 run's)."""
    _flush("case: sd position gradient = normal")
    sc = scenes.hull_pairs(512, 8)
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda().requires_grad_(True)
    quat = torch.from_numpy(sc["orientation"]).cuda().requires_grad_(True)
    pairs = torch.from_numpy(sc["pairs"].astype(np.int32)).cuda().contiguous()
    sd, n, pa, pb, st = diff.signed_distance(w, pos, quat, pairs)
    sd.sum().backward()
    gp = pos.grad.cpu().numpy()
    stn = st.cpu().numpy()
    nn = oracle.sd_pairs(sc, sc["pairs"], None, f64=True)[0][:, 1:4]
    epa = stn == 1
    nn[epa] = n.detach().cpu().numpy()[epa]
    p = sc["pairs"]
    walk = stn == 0
    # walk pairs against f64; EPA pairs within the f32 rounding of the
    # reported face normal (TOL_RAD, as for the forward)
    assert np.abs(gp[p[walk, 1]] - nn[walk]).max() < 1e-5
    assert np.abs(gp[p[walk, 0]] + nn[walk]).max() < 1e-5
    assert np.abs(gp[p[epa, 1]] - nn[epa]).max() < TOL_RAD
    assert np.abs(gp[p[epa, 0]] + nn[epa]).max() < TOL_RAD


def test_sd_rounding_exits_match_oracle_and_differentiate(torch, bpa, oracle, ctx, diff):
    """Pairs recorded from the penalty descent on which the f32 walk takes a
    rounding exit (status 5: an enclosure after the walk proved a gap; 6:
    EPA's depth beyond the walk's own depth bound), tests/golden.  Forward:
    status and all ten outputs bit-identical to the oracle.  Backward: the
    Band arm's adjoint at the closest iterate, so d sd / d p_B is a unit
    vector, d sd / d p_A its negative, and it points along the reported
    normal (within 0.05 rad: the adjoint rebuilds the iterate's simplex in
    double, and some of these pairs are within 1e-5 m of touching)."""
    from sd_helpers import rounding_exit_scene
    _flush("case: recorded rounding-exit pairs")
    sc, cases = rounding_exit_scene()
    _, _, out, st = _gpu_sd(torch, bpa, sc, ctx)
    ref, rst = oracle.sd_pairs(sc, sc["pairs"])
    assert np.array_equal(st & 7, rst) and set(rst.tolist()) == {oracle.SD_SEPARATED, oracle.SD_EPA_CONTRADICTED}
    assert np.array_equal(out, ref)
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda().requires_grad_(True)
    quat = torch.from_numpy(sc["orientation"]).cuda().requires_grad_(True)
    pairs = torch.from_numpy(sc["pairs"].astype(np.int32)).cuda().contiguous()
    sd, n, _, _, _ = diff.signed_distance(w, pos, quat, pairs)
    sd.sum().backward()
    gp = pos.grad.cpu().numpy().astype(np.float64)
    p = sc["pairs"].astype(np.int64)
    gb, ga = gp[p[:, 1]], gp[p[:, 0]]
    ang = _angle(gb.astype(np.float32), n.detach().cpu().numpy())
    _flush(f"  {len(cases)} pairs, forward bit-identical; |grad p_B| in [{np.linalg.norm(gb, axis=1).min():.6f}, "
           f"{np.linalg.norm(gb, axis=1).max():.6f}], max angle to n {ang.max():.3g} rad")
    assert np.isfinite(gp).all() and np.isfinite(quat.grad.cpu().numpy()).all()
    assert np.abs(np.linalg.norm(gb, axis=1) - 1.0).max() < 1e-5 and np.abs(ga + gb).max() < 1e-6
    assert ang.max() < 0.05


def test_sd_touching_cores_match_oracle_and_differentiate(torch, bpa, oracle, ctx, diff):
    """Box-hull pairs bisected to core touch, on which the f32 walk ends on
    the touch gate with no enclosing tetrahedron (status SD_TOUCH: sd =
    -(margins), the normal from the touching simplex).  Forward: status and
    all ten outputs bit-identical to the oracle.  Backward (autograd of the
    summed sd): d sd / d p_B is a unit vector within 1e-3 rad of the reported
    normal (the adjoint rebuilds the normal in double), d sd / d p_A its
    negative, and the pose gradient meets f64 central differences of the
    oracle's signed distance (h = 1e-4) within 1e-3 relative on the pairs
    smooth within +-h.  Without the touch arm these pairs had no gradient,
    and a de-penetration descent kept one of them 0.04 m deep."""
    from sd_helpers import fd, touch_scene
    _flush("case: touching box-hull cores")
    sc = touch_scene(oracle)
    _, _, out, st = _gpu_sd(torch, bpa, sc, ctx)
    ref, rst = oracle.sd_pairs(sc, sc["pairs"])
    assert (rst == oracle.SD_TOUCH).all() and np.array_equal(st & 7, rst)
    assert np.array_equal(out, ref)
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda().requires_grad_(True)
    quat = torch.from_numpy(sc["orientation"]).cuda().requires_grad_(True)
    pairs = torch.from_numpy(sc["pairs"].astype(np.int32)).cuda().contiguous()
    sd, n, _, _, _ = diff.signed_distance(w, pos, quat, pairs)
    sd.sum().backward()
    gp = pos.grad.cpu().numpy().astype(np.float64)
    gq = quat.grad.cpu().numpy().astype(np.float64)
    p = sc["pairs"].astype(np.int64)
    gb, ga = gp[p[:, 1]], gp[p[:, 0]]
    ang = _angle(gb.astype(np.float32), n.detach().cpu().numpy())
    g14 = np.concatenate([ga, gq[p[:, 0]], gb, gq[p[:, 1]]], axis=1)
    P = p.shape[0]
    g10 = np.zeros((P, 10))
    g10[:, 0] = 1.0
    vjp, j1, _, _ = fd(oracle, sc, g10, H)
    _, j2, _, _ = fd(oracle, sc, g10, H / 2)
    scale = np.maximum(np.linalg.norm(vjp, axis=1), 1e-2)
    err = np.linalg.norm(g14 - vjp, axis=1) / scale
    ratio = j2 / np.maximum(j1, 1e-300)
    smooth = np.all((j1 <= 1e-3 * scale[:, None]) | ((ratio >= 0.3) & (ratio <= 0.7)), axis=1)
    _flush(f"  {P} pairs, forward bit-identical; |grad p_B| in [{np.linalg.norm(gb, axis=1).min():.6f}, "
           f"{np.linalg.norm(gb, axis=1).max():.6f}], max angle to n {ang.max():.3g} rad; "
           f"{int(smooth.sum())} smooth within +-h: max rel. error {err[smooth].max():.3g}, "
           f"median {np.median(err[smooth]):.3g}")
    assert np.isfinite(g14).all()
    assert np.abs(np.linalg.norm(gb, axis=1) - 1.0).max() < 1e-5 and np.abs(ga + gb).max() < 1e-6
    assert ang.max() < 1e-3
    assert smooth.sum() >= 0.85 * P and err[smooth].max() < REL


def test_autograd_function_matches_raw_adjoint(torch, bpa, scenes, ctx, diff):
    """The torch.autograd.Function sums the per-pair adjoint over bodies."""
    _flush("case: autograd function")
    sc = scenes.hull_pairs(256, 9)
    # bodies shared between pairs: add the cross pairs (2k, 2k+3)
    k = np.arange(0, 255)
    extra = np.stack([2 * k, 2 * k + 3], axis=1).astype(np.uint32)
    sc["pairs"] = np.concatenate([sc["pairs"], extra])
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda().requires_grad_(True)
    quat = torch.from_numpy(sc["orientation"]).cuda().requires_grad_(True)
    pairs = torch.from_numpy(sc["pairs"].astype(np.int32)).cuda().contiguous()
    rng = np.random.default_rng(4)
    P = pairs.shape[0]
    wts = torch.from_numpy(rng.normal(size=(P, 10)).astype(np.float32)).cuda()
    sd, n, pa, pb, _ = diff.signed_distance(w, pos, quat, pairs)
    loss = (wts[:, 0] * sd).sum() + (wts[:, 1:4] * n).sum() + (wts[:, 4:7] * pa).sum() + (wts[:, 7:10] * pb).sum()
    loss.backward()
    g14 = torch.empty((P, 14), dtype=torch.float32, device="cuda")
    w.sd_backward(pairs.data_ptr(), P, wts.contiguous().data_ptr(), g14.data_ptr(), None)
    torch.cuda.synchronize()
    g14 = g14.cpu().numpy().astype(np.float64)
    N = sc["position"].shape[0]
    gpos = np.zeros((N, 3))
    gq = np.zeros((N, 4))
    pr = sc["pairs"].astype(np.int64)
    np.add.at(gpos, pr[:, 0], g14[:, 0:3])
    np.add.at(gpos, pr[:, 1], g14[:, 7:10])
    np.add.at(gq, pr[:, 0], g14[:, 3:7])
    np.add.at(gq, pr[:, 1], g14[:, 10:14])
    # float32 atomic accumulation over pairs sharing a body: rounding relative
    # to the largest contribution
    for got, want in ((pos.grad.cpu().numpy(), gpos), (quat.grad.cpu().numpy(), gq)):
        scale = np.abs(g14).max()
        d = np.abs(got - want).max()
        _flush(f"autograd sum vs per-pair adjoint: max diff {d:.3g}, largest contribution {scale:.3g}")
        assert d <= 1e-6 * scale + 1e-6


def _loss64(oracle, sc, pos, quat, margin=0.04):
    """sum max(0, -sd)^2 over the oracle's own pairs at these poses (f64 sd),
    world by world."""
    s2 = dict(sc)
    s2["position"] = pos.astype(np.float32)
    s2["orientation"] = quat.astype(np.float32)
    pairs = oracle.world_pairs(s2, margin)
    p = pairs.astype(np.int64)
    pose = np.concatenate([pos[p[:, 0]], quat[p[:, 0]], pos[p[:, 1]], quat[p[:, 1]]], axis=1)
    s2["pairs"] = pairs
    out, _ = oracle.sd_pairs(s2, pairs, pose, f64=True)
    pen = np.maximum(-out[:, 0], 0.0)
    return float(np.sum(pen * pen))


def test_penalty_gradient_vs_f64_finite_differences(torch, bpa, scenes, oracle, ctx, diff):
    """Batched worlds of boxes and hulls: the gradient of the penetration
    penalty (world pass for the pairs, signed distance, adjoint) against
    (a) the f64 reference gradient (the oracle's adjoint over the same pairs,
    summed per body) and (b) f64 central differences (h = 1e-4) of the same
    loss evaluated by the oracle (its own broadphase and f64 signed
    distance), along directions that move three bodies each (a direction
    moving every body crosses some contact's support-feature switch)."""
    _flush("case: penalty gradient, 12 worlds of boxes and hulls")
    sc = scenes.batched_boxes_hulls(12, 15, 64, 160)
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda().requires_grad_(True)
    quat = torch.from_numpy(sc["orientation"]).cuda().requires_grad_(True)
    loss, pen, pairs = diff.penetration_loss(w, pos, quat)
    loss.backward()
    gp = pos.grad.cpu().numpy().astype(np.float64)
    gq = quat.grad.cpu().numpy().astype(np.float64)
    p0 = sc["position"].astype(np.float64)
    q0 = sc["orientation"].astype(np.float64)
    l64 = _loss64(oracle, sc, p0, q0)
    # (a) reference gradient over the pass's pairs
    pr = pairs.cpu().numpy().astype(np.uint32)
    s2 = dict(sc)
    s2["pairs"] = pr
    out64, _ = oracle.sd_pairs(s2, pr, None, f64=True)
    g10 = np.zeros((pr.shape[0], 10))
    g10[:, 0] = -2.0 * np.maximum(-out64[:, 0], 0.0)
    g14, _ = oracle.sd_vjp_pairs(s2, pr, g10)
    N = p0.shape[0]
    rp = np.zeros((N, 3))
    rq = np.zeros((N, 4))
    ii = pr.astype(np.int64)
    np.add.at(rp, ii[:, 0], g14[:, 0:3])
    np.add.at(rp, ii[:, 1], g14[:, 7:10])
    np.add.at(rq, ii[:, 0], g14[:, 3:7])
    np.add.at(rq, ii[:, 1], g14[:, 10:14])
    gall = np.concatenate([gp.ravel(), gq.ravel()])
    rall = np.concatenate([rp.ravel(), rq.ravel()])
    rel_ref = np.linalg.norm(gall - rall) / np.linalg.norm(rall)
    _flush(f"penalty: {pr.shape[0]} pairs, {int((pen > 0).sum())} penetrating, loss f32 {float(loss):.6g} "
           f"f64 {l64:.6g}; gradient vs f64 reference: relative {rel_ref:.3g}")
    assert abs(float(loss) - l64) <= 1e-4 * max(l64, 1.0)
    assert rel_ref < 1e-2
    # (b) sparse directions against f64 differences
    rng = np.random.default_rng(8)
    held = within = 0
    for k in range(12):
        vp = np.zeros_like(p0)
        vq = np.zeros_like(q0)
        for b_ in rng.choice(N, 3, replace=False):
            vp[b_] = rng.normal(size=3)
            vq[b_] = rng.normal(size=4)
        g_dir = float(np.sum(gp * vp) + np.sum(gq * vq))
        fd, jump = [], []
        for h in (H, H / 2):
            lp = _loss64(oracle, sc, p0 + h * vp, q0 + h * vq)
            lm = _loss64(oracle, sc, p0 - h * vp, q0 - h * vq)
            fd.append((lp - lm) / (2 * h))
            jump.append(abs(lp - 2 * l64 + lm) / h)
        scale = max(abs(fd[0]), 1e-3)
        ratio = jump[1] / max(jump[0], 1e-300)
        smooth = jump[0] <= 1e-3 * scale or 0.3 <= ratio <= 0.7
        rel = abs(g_dir - fd[0]) / scale
        rel_rich = abs(g_dir - (4 * fd[1] - fd[0]) / 3) / scale
        ok = rel <= REL or rel_rich <= REL
        held += smooth
        within += smooth and ok
        _flush(f"  direction {k}: adjoint {g_dir:.8g} fd {fd[0]:.8g} rel {rel:.3g} "
               f"({'smooth' if smooth else 'non-smooth within +-h'})")
        assert ok or not smooth, f"direction {k}: adjoint {g_dir} vs f64 difference {fd[0]} (rel {rel:.3g})"
    assert held >= 8


def test_depenetration_reduces_penetration(torch, bpa, scenes, ctx, diff):
    """Gradient steps on the penalty push interpenetrating bodies apart."""
    _flush("case: depenetration, 16 worlds")
    sc = scenes.batched_boxes_hulls(16, 21, 64, 256)
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda()
    quat = torch.from_numpy(sc["orientation"]).cuda()
    _, _, hist = diff.depenetrate(w, pos, quat, steps=40)
    _flush(f"depenetration: loss {hist[0][0]:.4g} -> {hist[-1][0]:.4g}, max penetration "
           f"{hist[0][1]:.4g} -> {hist[-1][1]:.4g}")
    assert hist[-1][0] < 0.05 * hist[0][0]
    assert hist[-1][1] < 0.5 * hist[0][1]


def test_captured_depenetration_matches_autograd_loop(torch, bpa, scenes, ctx, diff):
    """The captured step (bch_depen_capture: broadphase, signed distance,
    adjoint of the penetrating pairs, gradient sum and pose update in one
    graph) against the autograd loop from the same poses:
      - the first step's record (loss, largest penetration, pairs) agrees to
        f32 summation order;
      - the poses after one step agree to the rounding that the order of
        the per-body gradient sums (float atomics on both sides) leaves;
      - later steps start from poses that differ in the last bits, and EPA's
        1e-4 convergence tolerance lets that grow, so the next step's loss
        is held to 1e-3 only."""
    _flush("case: captured depenetration vs autograd loop, 16 worlds")
    sc = scenes.batched_boxes_hulls(16, 21, 64, 256)
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda()
    quat = torch.from_numpy(sc["orientation"]).cuda()
    p1_ref, q1_ref, h1_ref = diff.depenetrate(w, pos, quat, steps=1)
    p1_g, q1_g, h1_g = diff.depenetrate_graph(w, pos, quat, steps=1)
    (l_a, m_a, n_a), (l_b, m_b, n_b) = h1_ref[0], h1_g[0]
    _flush(f"step 0: loss autograd {l_a:.8g} graph {l_b:.8g}; max pen {m_a:.7g} / {m_b:.7g}; pairs {n_a} / {n_b}")
    assert n_a == n_b and m_a == m_b
    assert abs(l_a - l_b) <= 1e-6 * l_a
    dp = (p1_g - p1_ref).abs().max().item()
    dq = (q1_g - q1_ref).abs().max().item()
    moved = (p1_ref - pos).abs().max().item()
    _flush(f"after 1 step: max |dpos| {dp:.3g} (largest move {moved:.3g}), max |dquat| {dq:.3g}")
    assert dp < 1e-5 and dq < 1e-5
    # two steps, history read between chunks of one
    _, _, h_ref = diff.depenetrate(w, pos, quat, steps=2)
    _, _, h_g = diff.depenetrate_graph(w, pos, quat, steps=2, chunk=1)
    assert len(h_g) == 2
    _flush(f"step 1: loss autograd {h_ref[1][0]:.7g} graph {h_g[1][0]:.7g}; pairs {h_ref[1][2]} / {h_g[1][2]}")
    assert abs(h_ref[1][0] - h_g[1][0]) <= 1e-3 * h_ref[1][0]
    assert abs(h_ref[1][2] - h_g[1][2]) <= 2
    # the input poses are untouched
    assert torch.equal(pos, torch.from_numpy(sc["position"]).cuda())


def test_captured_depenetration_reduces_penetration(torch, bpa, scenes, ctx, diff):
    """40 captured steps push the bodies apart like the autograd loop does."""
    _flush("case: captured depenetration, 16 worlds, 40 steps")
    sc = scenes.batched_boxes_hulls(16, 21, 64, 256)
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda()
    quat = torch.from_numpy(sc["orientation"]).cuda()
    p, q, hist = diff.depenetrate_graph(w, pos, quat, steps=40, chunk=16)
    _flush(f"captured depenetration: loss {hist[0][0]:.4g} -> {hist[-1][0]:.4g}, max penetration "
           f"{hist[0][1]:.4g} -> {hist[-1][1]:.4g}")
    assert hist[-1][0] < 0.05 * hist[0][0]
    assert hist[-1][1] < 0.5 * hist[0][1]
    n = q.norm(dim=1)
    assert torch.all((n - 1).abs() < 1e-5)


def test_captured_depenetration_recovers_from_a_mid_chunk_overflow(torch, bpa, scenes, ctx, diff):
    """A capacity of exactly the first pass's pairs, and the descent run
    backwards (lr < 0 drives penetrating bodies together, so later steps find
    more pairs): a step inside a chunk overflows, the chunk is rolled back to
    its starting poses, captured again with a larger capacity and rerun, and
    the run still records every step."""
    _flush("case: captured depenetration, mid-chunk pair overflow, 16 worlds")
    sc = scenes.batched_boxes_hulls(16, 21, 64, 256)
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda()
    quat = torch.from_numpy(sc["orientation"]).cuda()
    P0, _ = w.collide(0.04)
    info = {}
    p, q, hist = diff.depenetrate_graph(w, pos, quat, steps=24, lr=-0.05, chunk=8, pair_capacity=P0, info=info)
    pairs = [h[2] for h in hist]
    _flush(f"pairs {P0} -> {max(pairs)}, recaptures {info['recaptures']}")
    assert len(hist) == 24
    assert max(pairs) > P0 and info["recaptures"] >= 1
    assert hist[-1][1] > hist[0][1]  # run backwards, the penetration grows
    n = q.norm(dim=1)
    assert torch.all((n - 1).abs() < 1e-5)


def test_both_descent_loops_end_near_zero_penetration(torch, bpa, scenes, ctx, diff):
    """100 steps of the autograd loop and of the captured loop from the same
    start: the two trajectories part after a few steps (float-atomic gradient
    sums, then EPA's 1e-4 tolerance), but both end with the loss below 1e-6
    of its start.  Pushing one pair apart can drive a body into a neighbour:
    a fresh overlap of up to a few cm then appears and decays by about 5x per
    step, so the end state depends on when the last one came.  Both loops
    must therefore also reach, at some step of the last 10, a loss below 1e-9
    of the start with no pair deeper than 0.1 mm (the f32 floor is ~0.5 um)."""
    _flush("case: autograd and captured descent, 100 steps, 16 worlds")
    sc = scenes.batched_boxes_hulls(16, 21, 64, 256)
    w = bpa.CollisionWorld(sc, ctx)
    pos = torch.from_numpy(sc["position"]).cuda()
    quat = torch.from_numpy(sc["orientation"]).cuda()
    _, _, ha = diff.depenetrate(w, pos, quat, steps=100)
    _, _, hg = diff.depenetrate_graph(w, pos, quat, steps=100, chunk=25)
    _flush(f"autograd: loss {ha[0][0]:.4g} -> {ha[-1][0]:.3g}, max pen {ha[-1][1]:.3g}; "
           f"captured: loss {hg[0][0]:.4g} -> {hg[-1][0]:.3g}, max pen {hg[-1][1]:.3g}")
    for h in (ha, hg):
        assert h[-1][0] < 1e-6 * h[0][0]
        best = min(h[-10:], key=lambda e: e[0])
        assert best[0] < 1e-9 * h[0][0] and best[1] < 1e-4, best
