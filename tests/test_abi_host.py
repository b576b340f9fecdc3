"""CPU checks of the boundary and host logic (no GPU compute calls).

* the C-ABI library loads and exports every function include/*.h declares;
* the scene generators replay the reference's test LCG exactly
  (axis_sweep3.rs:823-852) and are deterministic;
* the Python mirror refuses to run without its HIP library (no fallback).
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\**\s*(bch_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("bch_create", "bch_bp_create_proxy", "bch_bp_compute_pairs", "bch_collide_batch",
                     "bch_world_collide"):
        assert required in names


def test_library_exports_every_declared_symbol(bpa):
    lib_path = bpa._lib.LIB_PATH
    assert os.path.exists(lib_path), "build() must produce the in-tree library"
    L = ctypes.CDLL(lib_path)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
    assert sorted(bpa._lib.EXPORTS) == declared_functions()


def test_lcg_matches_reference_formula(scenes):
    A, Cc, M = 6364136223846793005, 1442695040888963407, (1 << 64) - 1
    s = 12345
    ref_u64, ref_unit = [], []
    for _ in range(1000):
        s = (s * A + Cc) & M
        ref_u64.append(s)
        ref_unit.append(np.float32(float(s >> 11)) / np.float32(2.0 ** 53))
    g = scenes.Lcg(12345)
    assert [int(x) for x in g.next_u64(1000)] == ref_u64
    g = scenes.Lcg(12345)
    assert np.array_equal(g.unit(1000), np.array(ref_unit, np.float32))
    g = scenes.Lcg(12345)   # block boundaries do not change the stream
    a = np.concatenate([g.unit(7), g.unit(500), g.unit(493)])
    assert np.array_equal(a, np.array(ref_unit, np.float32))


def test_scenes_are_deterministic_and_well_formed(scenes):
    a = scenes.mixed_convex(5000, 2)
    b = scenes.mixed_convex(5000, 2)
    for k in ("position", "orientation", "shape_type", "shape_params", "hull_first", "hull_count", "hull_vertices"):
        assert np.array_equal(a[k], b[k])
    t = a["shape_type"]
    frac = np.bincount(t, minlength=4) / t.shape[0]
    assert np.allclose(frac, [0.3, 0.4, 0.2, 0.1], atol=0.03)  # sphere, box, capsule, hull
    q = a["orientation"].astype(np.float64)
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-6)
    h = np.nonzero(t == 3)[0]
    assert np.all((a["hull_count"][h] >= 16) & (a["hull_count"][h] <= 64))
    assert int((a["hull_first"][h] + a["hull_count"][h]).max()) <= a["hull_vertices"].shape[0]
    p = scenes.boxes_spheres_on_plane(2048, 1)
    assert p["shape_type"][0] == 4 and np.all(p["shape_type"][1:] < 2)
    assert np.all(p["position"][1:, 1] >= 0) and np.all(p["position"][1:, 1] <= 40)


def test_batched_scene_offsets(scenes):
    sc = scenes.batched_worlds(32, 3)
    wo = sc["world_offsets"]
    assert wo[0] == 0 and wo[-1] == sc["shape_type"].shape[0]
    sizes = np.diff(wo) - 1   # minus each world's plane
    assert np.all((sizes >= 16) & (sizes <= 2048))
    w5 = scenes.world_slice(sc, 5)
    assert w5["shape_type"][0] == 4 and w5["shape_type"].shape[0] == wo[6] - wo[5]


def test_no_cpu_fallback_without_library(bpa, monkeypatch, tmp_path):
    monkeypatch.setattr(bpa._lib, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(bpa._lib, "_lib", None)
    with pytest.raises(ImportError):
        bpa._lib.load()
