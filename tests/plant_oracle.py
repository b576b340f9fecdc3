"""Planted-mistake check of how well the CPU tests pin the oracle (CPU only).

Not collected by pytest.  Each entry below plants one plausible mistake into a
copy of oracle/oracle.hpp (a single text substitution), builds that copy into
its own liboracle.so and runs the CPU oracle tests against it through
BCH_ORACLE_LIB.  A mutation that every test still passes marks a part of the
oracle that the tests do not pin.

    python tests/plant_oracle.py            # every mutation
    python tests/plant_oracle.py epa_axis_seeds find_best_last_min
    python tests/plant_oracle.py --list

Exit status 0 when every mutation fails at least one test.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

# (name, what the reference does (file:line), text in oracle.hpp, planted text)
MUTATIONS = [
    ("epa_axis_seeds",
     "epa.rs:150-172 expands the seed tetrahedron by the six axis supports",
     "  for (auto& d : dirs) {\n    int idx = (int)verts.size();",
     "  for (auto& d : dirs) {\n    if (dirs) break;\n    int idx = (int)verts.size();"),
    ("find_best_last_min",
     "epa.rs:212-224 min_by keeps the first of equal |distance|",
     "if (std::fabs(faces[best].distance) > std::fabs(faces[i].distance)) best = (int)i;",
     "if (std::fabs(faces[best].distance) >= std::fabs(faces[i].distance)) best = (int)i;"),
    ("face_distance_no_edge_term",
     "epa.rs:262-303 face distance falls back to the edge / vertex distance",
     "  T d;\n  if (edge_distance(pa, pb, normal, d)) return d;",
     "  T d;\n  if ((false)) (void)d;\n  else return dot(pa, normal);\n  if (edge_distance(pa, pb, normal, d)) return d;"),
    ("boundary_ordered_erase",
     "epa.rs:343 cancels a horizon edge with Vec::swap_remove",
     "        boundary[found] = boundary.back();  // swap_remove\n        boundary.pop_back();",
     "        boundary.erase(boundary.begin() + found);"),
    ("gjk_rel_error_1e-4",
     "gjk.rs:292 REL_ERROR2 = 1e-6",
     "const T REL_ERROR2 = ORL(1e-6);",
     "const T REL_ERROR2 = ORL(1e-4);"),
    ("cull_first_max",
     "box_box.rs:139-145 `>=` keeps the last deepest point",
     "    if (deps[i] >= deepest_depth) {",
     "    if (deps[i] > deepest_depth) {"),
    ("clip_dep_strict",
     "box_box.rs:501 keeps clipped points with dep >= 0",
     "    if (dep >= T(0)) {\n      pen_pts[np] = point;",
     "    if (dep > T(0)) {\n      pen_pts[np] = point;"),
    ("cull_sort_unstable",
     "box_box.rs:160-168 stable descending insertion sort (strict <)",
     "    while (k > 0 && rd[k - 1] < kd) {",
     "    while (k > 0 && rd[k - 1] <= kd) {"),
    ("epa_visibility_eps",
     "epa.rs:395-398 face visible beyond PLANE_EPS = 1e-5",
     "  return dot(f.normal, p) - f.plane_distance > ORL(1e-5);",
     "  return dot(f.normal, p) - f.plane_distance > ORL(1e-4);"),
    ("epa_exit_tolerance",
     "epa.rs:35 EPSILON = 1e-4",
     "const T EPSILON = std::is_same<T, float>::value ? (T)1e-4f : (T)1e-12;",
     "const T EPSILON = std::is_same<T, float>::value ? (T)1e-3f : (T)1e-12;"),
    ("gjk_duplicate_gate",
     "gjk.rs:253-256 duplicate support gate abs_diff_eq 1e-6",
     "if (abs_diff_eq(w.w, buf[i].w, ORL(1e-6))) duplicate = true;",
     "if (abs_diff_eq(w.w, buf[i].w, ORL(1e-4))) duplicate = true;"),
    ("edge_distance_vertex_swap",
     "epa.rs:290-293 a.ba > 0 takes |a|, b.ba < 0 takes |b|",
     "    if (a_dot_ba > T(0)) out = len(pa);\n    else if (b_dot_ba < T(0)) out = len(pb);",
     "    if (a_dot_ba > T(0)) out = len(pb);\n    else if (b_dot_ba < T(0)) out = len(pa);"),
    ("box_box_fudge",
     "box_box.rs:20 FUDGE = 1.05 on the edge axes",
     "FUDGE = ORL(1.05);",
     "FUDGE = ORL(1.0);"),
    ("epa_vertex_limit",
     "epa.rs:46 MAX_VERTICES = 128",
     "if ((int)verts.size() >= 128) {",
     "if ((int)verts.size() >= 64) {"),
    ("epa_iteration_limit",
     "epa.rs:30 MAX_ITERATIONS = 255",
     "for (; it < 255; ++it) {",
     "for (; it < 32; ++it) {"),
    ("gjk_touch_scale",
     "gjk.rs:196-198 touch tolerance 1e-12 max(|w0|^2, 1)",
     "T touch2 = ORL(1e-12) * std::fmax",
     "T touch2 = ORL(1e-10) * std::fmax"),
    ("hull_support_last_max",
     "convex_hull.rs:55-66 strict > keeps the first maximum",
     "          if (d > best_dot) {",
     "          if (d >= best_dot) {"),
    ("box_support_sign",
     "box_shape.rs:30-37 sign >= 0 picks +h",
     "return mk(dir.x >= T(0) ? half.x : -half.x, dir.y >= T(0) ? half.y : -half.y,",
     "return mk(dir.x > T(0) ? half.x : -half.x, dir.y > T(0) ? half.y : -half.y,"),
    ("sphere_box_face_tie",
     "sphere.rs:60-70 inside the box, ties go x -> y -> z",
     "if (dx <= dy && dx <= dz) {",
     "if (dx < dy && dx <= dz) {"),
    ("reduce_tetra_coplanar",
     "gjk.rs:498 coplanar tetrahedron gate 1e-12",
     "std::fabs(dot(cross(ab, ac), ad)) < ORL(1e-12)",
     "std::fabs(dot(cross(ab, ac), ad)) < ORL(1e-8)"),
    ("capsule_core_tip",
     "capsule.rs:59-69 core tip picks +hh for dir.y >= 0",
     "        return mk(T(0), dir.y >= T(0) ? half_height : -half_height, T(0));",
     "        return mk(T(0), dir.y > T(0) ? half_height : -half_height, T(0));"),
    ("sd_gap_proof_off",
     "signed-distance extension: a walk that proved a core gap never reports an enclosure",
     "    if (delta > T(0) && delta * delta > v2 * gap2) proven_gap = true;",
     "    if ((false)) proven_gap = true;"),
    ("sd_epa_bound_off",
     "signed-distance extension: an EPA depth beyond the walk's depth bound is not reported",
     "    contradicted = e.depth > g.depth_bound + T(SD_EPA_SLACK);\n  }\n  if (g.kind == 1 || (g.kind == 0 && g.capped) || contradicted) {\n    if (g.kind == 1) {",
     "    contradicted = false && e.depth > g.depth_bound;\n  }\n  if (g.kind == 1 || (g.kind == 0 && g.capped) || contradicted) {\n    if (g.kind == 1) {"),
    ("sd_touch_no_normal",
     "signed-distance extension: touching cores report a normal and witnesses",
     "  if (g.touch) {\n    // the cores touch",
     "  if ((false)) {\n    // the cores touch"),
    ("sd_touch_raw_direction",
     "signed-distance extension: the touch normal drops the simplex's in-plane components",
     "  V3<T> s = g.touch_dir, n = s;\n  if (g.n_active == 3) {",
     "  V3<T> s = g.touch_dir, n = s;\n  if ((false)) {"),
    ("capped_last_iterate",
     "signed-distance extension: a capped walk reports its closest iterate, not the last",
     "    if (!(v2 >= best_v2)) {",
     "    if ((true)) {"),
]

TESTS = ["tests/test_oracle_kat.py", "tests/test_oracle_math.py", "tests/test_oracle_pins.py",
         "tests/test_oracle_manifold.py", "tests/test_oracle_mesh.py"]


def run_one(m, workdir):
    name, cite, old, new = m
    src = open(os.path.join(ORACLE, "oracle.hpp")).read()
    if src.count(old) != 1:
        return name, cite, "NOT APPLIED (text not found exactly once)", ""
    d = os.path.join(workdir, name)
    os.makedirs(d)
    for f in ("oracle_capi.cpp", "Makefile"):
        shutil.copy(os.path.join(ORACLE, f), d)
    with open(os.path.join(d, "oracle.hpp"), "w") as fh:
        fh.write(src.replace(old, new))
    b = subprocess.run(["make", "-s", "-C", d], capture_output=True, text=True)
    if b.returncode != 0:
        return name, cite, "BUILD FAILED", b.stderr[-2000:]
    env = dict(os.environ, BCH_ORACLE_LIB=os.path.join(d, "liboracle.so"))
    tests = [t for t in TESTS if os.path.exists(os.path.join(ROOT, t))]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        "-o", "addopts=", *tests], cwd=ROOT, env=env, capture_output=True, text=True)
    failed = [ln.split(" ")[1] for ln in r.stdout.splitlines() if ln.startswith("FAILED ")]
    if r.returncode == 0:
        return name, cite, "SURVIVED (no test failed)", ""
    if not failed:
        return name, cite, f"pytest exit {r.returncode}", r.stdout[-1500:]
    return name, cite, f"caught by {len(failed)} test(s)", "\n".join("    " + f for f in failed)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="*")
    ap.add_argument("--list", action="store_true")
    ap.add_argument("--jobs", type=int, default=4)
    a = ap.parse_args()
    if a.list:
        for m in MUTATIONS:
            print(f"{m[0]:28s} {m[1]}")
        return 0
    todo = [m for m in MUTATIONS if not a.names or m[0] in a.names]
    work = tempfile.mkdtemp(prefix="oracle_plant_")
    survived = 0
    try:
        with cf.ThreadPoolExecutor(a.jobs) as ex:
            for name, cite, verdict, detail in ex.map(lambda m: run_one(m, work), todo):
                print(f"{name:28s} {verdict}   [{cite}]", flush=True)
                if detail:
                    print(detail, flush=True)
                if not verdict.startswith("caught"):
                    survived += 1
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return 1 if survived else 0


if __name__ == "__main__":
    sys.exit(main())
