"""Build the HIP library (gfx950) in-tree: bullet-physics_amd/libbullet_collision_hip.so.

-ffp-contract=off keeps every f32 multiply and add separately rounded (the
AABB bits, hence the pair set, must match the CPU arithmetic exactly);
packed-fp32 instructions are disabled for the device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libbullet_collision_hip.so")
SOURCES = [os.path.join(HERE, "csrc", "bch_lib.hip")]
# measurement probes for the roofline figures (bench.py --valu-calibration); not the collision path
PROBE = os.path.join(HERE, "libbch_probe.so")
PROBE_SOURCES = [os.path.join(HERE, "csrc", "bch_probe.hip")]
ARCH = os.environ.get("BCH_OFFLOAD_ARCH", "gfx950")


def hipcc_command(out: str = LIB, sources=None) -> list[str]:
    return ["hipcc", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument",
            "-I", os.path.join(ROOT, "include"), "-o", out] + (sources or SOURCES)


def _compile(out: str, sources, verbose: bool) -> None:
    tmp = out + ".tmp"
    cmd = hipcc_command(tmp, sources)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for " + " ".join(sources))
    os.replace(tmp, out)


def build(verbose: bool = False) -> str:
    _compile(PROBE, PROBE_SOURCES, verbose)
    _compile(LIB, SOURCES, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
