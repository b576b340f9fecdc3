"""ctypes binding of libbullet_collision_hip.so (include/bullet_collision_hip.h).

The library is built in-tree (build.py).  Loading fails loudly when it is
missing: there is no CPU fallback on this path.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbullet_collision_hip.so")

BCH_OK, BCH_OVERFLOW, BCH_INVALID_ARG, BCH_HIP_ERROR, BCH_NO_DEVICE = 0, 1, 2, 3, 4

# Every symbol include/bullet_collision_hip.h declares.
EXPORTS = [
    "bch_create", "bch_destroy", "bch_last_error", "bch_synchronize", "bch_last_timings", "bch_last_epa_routes",
    "bch_bp_create_proxy", "bch_bp_set_aabb", "bch_bp_destroy_proxy", "bch_bp_proxy_aabb", "bch_bp_set_aabbs",
    "bch_bp_compute_pairs", "bch_bp_aabb_candidates", "bch_bp_raycast_candidates", "bch_collide_batch",
    "bch_world_set_bodies", "bch_world_set_poses", "bch_world_set_poses_device", "bch_world_collide",
    "bch_world_results", "bch_world_results_device", "bch_world_routes", "bch_set_tuning", "bch_get_tuning",
    "bch_sd_forward", "bch_sd_backward", "bch_set_stream", "bch_reset_stream", "bch_autotune",
    "bch_world_capture", "bch_world_replay", "bch_world_replay_counts", "bch_world_release_graph",
    "bch_depen_capture", "bch_depen_replay", "bch_depen_history", "bch_depen_release",
    "bch_world_set_body_flags", "bch_world_manifold_pass", "bch_world_manifold_results",
    "bch_world_manifold_impulses", "bch_world_manifold_cache", "bch_world_manifold_reset",
    "bch_world_set_poses_predicted", "bch_world_set_poses_predicted_device",
    "bch_mesh_add", "bch_mesh_clear", "bch_collide_mesh_batch", "bch_sweep_batch", "bch_sweep_mesh_batch",
    "bch_collide_batch_poses", "bch_collide_batch_static",
]

# launch-shape knobs (bch_set_tuning)
TUNE_HULL_LANES, TUNE_TRAVERSE_BLOCK, TUNE_EPA_GRID, TUNE_NP_GRID, TUNE_EPA_WAVE_FACES, TUNE_SORT_TILE = 1, 2, 3, 4, 5, 6


class BchBodies(C.Structure):
    _fields_ = [("n", C.c_uint32), ("position", C.c_void_p), ("orientation", C.c_void_p),
                ("shape_type", C.c_void_p), ("shape_params", C.c_void_p), ("hull_first", C.c_void_p),
                ("hull_count", C.c_void_p), ("hull_vertices", C.c_void_p), ("n_hull_vertices", C.c_uint32)]


class BchError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"bullet_collision_hip error {code}: {msg}")
        self.code = code


_lib = None


def _share_torch_hip_runtime():
    """One HIP runtime per process.  PyTorch wheels carry their own
    libamdhip64 (same soname, libamdhip64.so.7, loaded from torch/lib by
    their own name libamdhip64.so): if this library loaded the system copy
    first, a later `import torch` would bring up a second runtime that sees
    no device.  Loading torch's copy first (without importing torch) makes
    this library bind to it by soname, and torch finds it already loaded.
    BCH_SYSTEM_HIP=1 keeps the system runtime."""
    if os.environ.get("BCH_SYSTEM_HIP") == "1":
        return
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(path):
        C.CDLL(path, mode=C.RTLD_GLOBAL)


def load():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run python bullet-physics_amd/build.py (needs hipcc)")
    _share_torch_hip_runtime()
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, f32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float
    L.bch_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.bch_destroy.argtypes = [vp]
    L.bch_destroy.restype = None
    L.bch_last_error.argtypes = [vp]
    L.bch_last_error.restype = C.c_char_p
    L.bch_synchronize.argtypes = [vp]
    L.bch_last_timings.argtypes = [vp, vp]
    L.bch_last_epa_routes.argtypes = [vp, vp]
    L.bch_bp_create_proxy.argtypes = [vp, vp, vp, C.POINTER(u32)]
    L.bch_bp_set_aabb.argtypes = [vp, u32, vp, vp]
    L.bch_bp_destroy_proxy.argtypes = [vp, u32]
    L.bch_bp_proxy_aabb.argtypes = [vp, u32, vp, vp, C.POINTER(C.c_int)]
    L.bch_bp_set_aabbs.argtypes = [vp, u32, vp, vp, vp]
    L.bch_bp_compute_pairs.argtypes = [vp, vp, u64, C.POINTER(u64)]
    L.bch_bp_aabb_candidates.argtypes = [vp, vp, vp, vp, u64, C.POINTER(u64)]
    L.bch_bp_raycast_candidates.argtypes = [vp, vp, vp, vp, u64, C.POINTER(u64)]
    L.bch_collide_batch.argtypes = [vp, C.POINTER(BchBodies), vp, u64, vp, vp, vp]
    L.bch_collide_batch_poses.argtypes = [vp, vp, vp, vp, u64, vp, vp, vp]
    L.bch_collide_batch_static.argtypes = [vp, C.POINTER(BchBodies), vp, u64, vp, vp, vp]
    L.bch_world_set_bodies.argtypes = [vp, C.POINTER(BchBodies), u32, vp]
    L.bch_world_set_poses.argtypes = [vp, vp, vp]
    L.bch_world_set_poses_device.argtypes = [vp, vp, vp]
    L.bch_world_collide.argtypes = [vp, f32, C.POINTER(u64), C.POINTER(u64)]
    L.bch_world_results.argtypes = [vp, vp, vp, vp]
    L.bch_world_results_device.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
    L.bch_world_routes.argtypes = [vp, vp]
    L.bch_sd_forward.argtypes = [vp, vp, u64, vp, vp]
    L.bch_sd_backward.argtypes = [vp, vp, u64, vp, vp, vp]
    L.bch_set_stream.argtypes = [vp, vp]
    L.bch_reset_stream.argtypes = [vp]
    L.bch_autotune.argtypes = [vp, f32, C.c_int, C.POINTER(f32)]
    L.bch_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.bch_get_tuning.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.bch_world_capture.argtypes = [vp, f32, u64, u64]
    L.bch_world_replay.argtypes = [vp]
    L.bch_world_replay_counts.argtypes = [vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]
    L.bch_world_release_graph.argtypes = [vp]
    L.bch_depen_capture.argtypes = [vp, f32, f32, vp, vp, u64, u32, C.POINTER(u64)]
    L.bch_depen_replay.argtypes = [vp, u32]
    L.bch_world_set_body_flags.argtypes = [vp, vp, vp]
    L.bch_world_manifold_pass.argtypes = [vp, f32, f32, f32, C.POINTER(u64), C.POINTER(u64)]
    L.bch_world_manifold_results.argtypes = [vp, vp, vp, vp, vp]
    L.bch_world_manifold_impulses.argtypes = [vp, vp]
    L.bch_world_manifold_cache.argtypes = [vp, vp, vp]
    L.bch_world_manifold_reset.argtypes = [vp]
    L.bch_world_set_poses_predicted.argtypes = [vp, vp, vp, vp, f32]
    L.bch_mesh_add.argtypes = [vp, vp, u32, vp, u32, C.POINTER(u32)]
    L.bch_mesh_clear.argtypes = [vp]
    L.bch_collide_mesh_batch.argtypes = [vp, C.POINTER(BchBodies), vp, vp, vp, u64, vp, vp, vp]
    L.bch_sweep_batch.argtypes = [vp, C.POINTER(BchBodies), vp, vp, vp, vp, vp, u64, vp, vp]
    L.bch_sweep_mesh_batch.argtypes = [vp, C.POINTER(BchBodies), vp, vp, vp, vp, vp, vp, u64, vp, vp]
    L.bch_world_set_poses_predicted_device.argtypes = [vp, vp, vp, vp, f32]
    L.bch_depen_release.argtypes = [vp]
    L.bch_depen_history.argtypes = [vp, vp, vp, vp, vp, u32, C.POINTER(u32), C.c_int]
    _lib = L
    return L
