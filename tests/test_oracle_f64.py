"""Error of the f32 contact path against the reference algorithm in f64.

The GPU pass reproduces the oracle's f32 build bit for bit (tests/
test_gpu_parity.py), so the f32 build's error against the same restatement
compiled in f64 is the GPU's error against the reference's f64 build.  On a
20k-body mixed scene (all of its 75k pairs), per contact path:

* depth agrees to 1e-4 m and the normal to 1e-4 rad, except band normals of
  pairs whose cores nearly touch (the direction of a tiny closest-point
  vector turns by its rounding over its length) and EPA normals, which the
  1e-4 m support-plane exit fixes only to ~1e-3 rad;
* positions agree to 1e-4 m except where the witness is not unique: shape A
  presents a flat feature along the normal (EPA: witness = A's support along
  n, convex_convex.rs:117-154) or both shapes do (band: the closest points
  span a face), and the two builds pick different points of it;
* every pair whose contact count differs is one of two f32 exits of the
  reference walk (gjk.rs:171-288): the 64-iteration cap (:211, :285-287), which reports the
  pair separated (the no-progress test misses its 1 - 1e-6 exit by rounding
  on flat cores and the walk cycles), or a core overlap found on a nearly
  degenerate simplex whose EPA then finds no depth (no contact).
"""
import numpy as np

I = (0.0, 0.0, 0.0, 1.0)


def _shape(oracle, sc, i):
    t = int(sc["shape_type"][i])
    if t == 3:
        f, n = int(sc["hull_first"][i]), int(sc["hull_count"][i])
        return oracle.Shape(3, hull=sc["hull_vertices"][f:f + n])
    return oracle.Shape(t, list(sc["shape_params"][i]))


def _qrot(q, v):
    b, w = q[:3], q[3]
    return v * (w * w - b @ b) + b * (2 * (v @ b)) + np.cross(b, v) * 2 * w


def _non_unique_support(sc, i, n, rel=1e-5):
    """Shape i's support along n is shared by several points (a flat feature):
    candidates within rel of the largest projection, in float64."""
    t = int(sc["shape_type"][i])
    pr = sc["shape_params"][i].astype(np.float64)
    q = sc["orientation"][i].astype(np.float64)
    ld = _qrot(q * np.array([-1, -1, -1, 1.0]), np.asarray(n, np.float64))
    if t == 0:
        return False
    if t == 2:  # capsule: the two tips tie when the direction is square to the axis
        return abs(ld[1]) < rel
    if t == 1:
        c = np.array([[sx * pr[0], sy * pr[1], sz * pr[2]] for sx in (1, -1) for sy in (1, -1) for sz in (1, -1)])
    else:
        f, m = int(sc["hull_first"][i]), int(sc["hull_count"][i])
        c = sc["hull_vertices"][f:f + m].astype(np.float64)
    pj = c @ ld
    return int((pj >= pj.max() - rel * np.abs(c).max()).sum()) > 1


def test_f32_contacts_against_the_f64_build(oracle, scenes):
    sc = scenes.mixed_convex(20000, 2)
    mn, mx = oracle.scene_aabbs(sc, 0.04)
    pairs = oracle.sweep_pairs(mn, mx)
    c32, p32, path = oracle.collide_pairs(sc, pairs)
    c64, p64, _ = oracle.collide_pairs(sc, pairs, f64=True)
    thr = oracle.scene_thresholds(sc)
    margin = np.where(sc["shape_type"] == 0, sc["shape_params"][:, 0],
                      np.where(sc["shape_type"] == 1, 0.04,
                               np.where(sc["shape_type"] == 2, sc["shape_params"][:, 0], 0.0))).astype(np.float64)
    same = (c32 == c64) & (c32 > 0)
    names = {1: "band", 2: "epa", 3: "closed form"}
    report = []
    for pth, name in names.items():
        m = same & (path == pth)
        for pi in np.nonzero(m)[0]:
            i, j = pairs[pi]
            for k in range(int(c32[pi])):
                a, b = p32[pi, k].astype(np.float64), p64[pi, k]
                dp = np.linalg.norm(a[:3] - b[:3])
                dd = abs(a[6] - b[6])
                ang = np.arctan2(np.linalg.norm(np.cross(a[3:6], b[3:6])), np.dot(a[3:6], b[3:6]))
                report.append((pth, dp, dd, ang))
                assert dd <= 1e-4, (name, pi, dd)
                if ang > 1e-4:
                    if pth == 1:
                        # band: the normal is the closest-point vector of the cores over its length;
                        # its f32 rounding (~1e-6 m at these coordinates) turns it by that over the
                        # core distance (= separation + margins)
                        core = -b[6] + margin[i] + margin[j]
                        assert ang * core <= 1e-5, (name, pi, ang, core)
                    else:
                        # EPA stops once the support plane gains <= 1e-4 m (epa.rs:35, :181), which
                        # fixes the face normal only to about that over the feature size
                        assert pth == 2 and ang <= 1e-3, (name, pi, ang)
                if dp > 1e-4:
                    flat_a = _non_unique_support(sc, i, b[3:6])
                    flat_b = _non_unique_support(sc, j, -b[3:6])
                    assert flat_a if pth == 2 else (flat_a or flat_b), (name, pi, dp)
    r = np.array(report)
    for pth, name in names.items():
        x = r[r[:, 0] == pth]
        print(f"{name}: {x.shape[0]} points; |dp| median {np.median(x[:, 1]):.1e} p99 {np.quantile(x[:, 1], .99):.1e}; "
              f"|dd| median {np.median(x[:, 2]):.1e} max {x[:, 2].max():.1e}; angle median {np.median(x[:, 3]):.1e} "
              f"max {x[:, 3].max():.1e}; |dp| > 1e-4: {int((x[:, 1] > 1e-4).sum())}")
    # contact counts that differ: classified by how the f32 walk ended
    diff = np.nonzero(c32 != c64)[0]
    capped = degenerate = 0
    for pi in diff:
        i, j = pairs[pi]
        kind, iters, cap = oracle.gjk_exit(_shape(oracle, sc, i), sc["position"][i], sc["orientation"][i],
                                           _shape(oracle, sc, j), sc["position"][j], sc["orientation"][j],
                                           float(min(thr[i], thr[j])))
        if cap:
            assert kind == 0 and iters == 64 and c32[pi] == 0
            capped += 1
        else:
            assert kind == 2 and path[pi] == 2 and c32[pi] == 0, (pi, kind, iters)
            degenerate += 1
    print(f"{pairs.shape[0]} pairs; counts differ on {diff.size}: {capped} f32 walks at the 64-iteration cap, "
          f"{degenerate} f32 core overlaps whose EPA finds no depth")
    assert diff.size < 1e-3 * pairs.shape[0]
