"""Differentiable signed distance over the HIP forward and adjoint kernels.

``signed_distance(world, position, orientation, pairs)`` returns, per pair
(i, j), the signed distance (> 0 separated), the unit normal from body i to
body j and the witness points on i and on j, as torch tensors on the GPU
that autograd differentiates with respect to the poses (position [N, 3],
orientation [N, 4] as x, y, z, w).  The forward and backward run on
torch's current stream; the backward recomputes the walks (nothing is kept
between the two calls but the poses and the pair list).
"""
from __future__ import annotations

import torch


def _bind_stream(world):
    world.ctx.set_stream(torch.cuda.current_stream().cuda_stream)


class SignedDistance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, position, orientation, world, pairs):
        pos = position.detach().to(torch.float32).contiguous()
        quat = orientation.detach().to(torch.float32).contiguous()
        if pos.shape[0] != world.n or quat.shape[0] != world.n:
            raise ValueError(f"poses for {pos.shape[0]} bodies, the world holds {world.n}")
        if pairs.dtype != torch.int32 or not pairs.is_contiguous():
            raise ValueError("pairs must be a contiguous int32 [P, 2] tensor")
        P = pairs.shape[0]
        _bind_stream(world)
        world.set_poses_device(pos.data_ptr(), quat.data_ptr())
        out = torch.empty((P, 10), dtype=torch.float32, device=pos.device)
        status = torch.empty(P, dtype=torch.uint8, device=pos.device)
        if P:
            world.sd_forward(pairs.data_ptr(), P, out.data_ptr(), status.data_ptr())
        ctx.save_for_backward(pos, quat, pairs)
        ctx.world = world
        ctx.mark_non_differentiable(status)
        return out[:, 0], out[:, 1:4], out[:, 4:7], out[:, 7:10], status

    @staticmethod
    def backward(ctx, g_sd, g_n, g_pa, g_pb, _g_status):
        pos, quat, pairs = ctx.saved_tensors
        world = ctx.world
        P = pairs.shape[0]
        g = torch.zeros((P, 10), dtype=torch.float32, device=pos.device)
        if g_sd is not None:
            g[:, 0] = g_sd
        if g_n is not None:
            g[:, 1:4] = g_n
        if g_pa is not None:
            g[:, 4:7] = g_pa
        if g_pb is not None:
            g[:, 7:10] = g_pb
        g14 = torch.zeros((P, 14), dtype=torch.float32, device=pos.device)
        if P:
            _bind_stream(world)
            world.set_poses_device(pos.data_ptr(), quat.data_ptr())
            world.sd_backward(pairs.data_ptr(), P, g.data_ptr(), g14.data_ptr(), None)
        i = pairs[:, 0].long()
        j = pairs[:, 1].long()
        gpos = torch.zeros_like(pos).index_add_(0, i, g14[:, 0:3]).index_add_(0, j, g14[:, 7:10])
        gquat = torch.zeros_like(quat).index_add_(0, i, g14[:, 3:7]).index_add_(0, j, g14[:, 10:14])
        return gpos, gquat, None, None


def signed_distance(world, position, orientation, pairs):
    """(sd [P], normal [P, 3], witness_a [P, 3], witness_b [P, 3], status [P])."""
    return SignedDistance.apply(position, orientation, world, pairs)


class _DeviceArray:
    """A device buffer seen through __cuda_array_interface__ (zero copy)."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (ptr, False),
                                         "version": 2}


def world_pairs(world) -> torch.Tensor:
    """The last world pass's pair list as an int32 [P, 2] tensor (a copy: the
    next pass reuses the library's buffer)."""
    P = world.n_pairs
    if P == 0:
        return torch.zeros((0, 2), dtype=torch.int32, device="cuda")
    d_pairs, _, _ = world.results_device()
    return torch.as_tensor(_DeviceArray(d_pairs, (P, 2), "<i4"), device="cuda").clone()


def penetration_loss(world, position, orientation, aabb_margin: float = 0.04):
    """Penetration penalty sum(max(0, -sd)^2) over the pairs the world pass
    finds at these poses (broadphase and signed distance both on the GPU).
    Returns (loss, penetration [P], pairs [P, 2])."""
    pos = position.detach().to(torch.float32).contiguous()
    quat = orientation.detach().to(torch.float32).contiguous()
    _bind_stream(world)
    world.set_poses_device(pos.data_ptr(), quat.data_ptr())
    world.collide(aabb_margin)
    pairs = world_pairs(world)
    sd, _, _, _, _ = signed_distance(world, position, orientation, pairs)
    pen = torch.relu(-sd)
    return (pen * pen).sum(), pen, pairs


def depenetrate(world, position, orientation, steps: int = 100, lr: float = 0.2, aabb_margin: float = 0.04):
    """Gradient descent on the penetration penalty over positions and
    orientations (renormalised after each step).  Returns the final poses and
    per step (loss, max penetration, pairs)."""
    pos = position.detach().clone().requires_grad_(True)
    quat = orientation.detach().clone().requires_grad_(True)
    history = []
    for _ in range(steps):
        loss, pen, pairs = penetration_loss(world, pos, quat, aabb_margin)
        pos.grad = None
        quat.grad = None
        loss.backward()
        with torch.no_grad():
            pos -= lr * pos.grad
            quat -= lr * quat.grad
            quat /= quat.norm(dim=1, keepdim=True)
        history.append((float(loss.detach()), float(pen.detach().max()) if pen.numel() else 0.0,
                        int(pairs.shape[0])))
    return pos.detach(), quat.detach(), history


def depenetrate_graph(world, position, orientation, steps: int = 100, lr: float = 0.2, aabb_margin: float = 0.04,
                      headroom: float = 1.5, chunk: int = 25, pair_capacity: int | None = None,
                      info: dict | None = None):
    """:func:`depenetrate` with each step one captured HIP graph
    (bch_depen_capture): broadphase, signed distance, the adjoint of the
    penetrating pairs, the per-body gradient and the pose update all run on
    the device, with no host round trip inside a chunk of steps.  The
    history is read back every `chunk` steps; a chunk in which the pairs
    overflowed the capacity is rolled back and rerun after a capture with a
    larger one.  Returns the final poses and per step (loss, max
    penetration, pairs), like :func:`depenetrate`.  `pair_capacity` fixes the
    first capture's capacity (default: the first pass's pairs x headroom);
    `info`, when given, receives the number of recaptures."""
    import ctypes as C

    import numpy as np

    from . import BchError, _check, _lib

    L, h = world.L, world.ctx.h
    pos = position.detach().to(torch.float32).contiguous().clone()
    quat = orientation.detach().to(torch.float32).contiguous().clone()
    _bind_stream(world)
    world.set_poses_device(pos.data_ptr(), quat.data_ptr())
    P0, _ = world.collide(aabb_margin)
    cap = pair_capacity if pair_capacity is not None else max(1024, int(P0 * headroom) + 1024)
    found = C.c_uint64()
    recaptures = 0

    def capture():
        nonlocal cap
        for _ in range(4):
            rc = L.bch_depen_capture(h, aabb_margin, lr, pos.data_ptr(), quat.data_ptr(), cap, chunk,
                                     C.byref(found))
            if rc == _lib.BCH_OVERFLOW:
                cap = max(2 * cap, int(found.value * headroom) + 1024)
                continue
            _check(h, rc)
            return
        raise BchError(_lib.BCH_OVERFLOW, "de-penetration capture kept overflowing")

    capture()
    loss = np.zeros(chunk, np.float32)
    maxpen = np.zeros(chunk, np.float32)
    npairs = np.zeros(chunk, np.uint32)
    count = C.c_uint32()
    history = []
    try:
        while len(history) < steps:
            k = min(chunk, steps - len(history))
            saved = (pos.clone(), quat.clone())
            _check(h, L.bch_depen_replay(h, k))
            rc = L.bch_depen_history(h, loss.ctypes.data, maxpen.ctypes.data, npairs.ctypes.data, None, chunk,
                                     C.byref(count), 1)
            if rc == _lib.BCH_OVERFLOW:
                pos.copy_(saved[0])
                quat.copy_(saved[1])
                cap = max(2 * cap, int(npairs[:k].max() * headroom) + 1024)
                capture()
                recaptures += 1
                continue
            _check(h, rc)
            history += [(float(loss[s]), float(maxpen[s]), int(npairs[s])) for s in range(k)]
    finally:
        L.bch_depen_release(h)
    if info is not None:
        info["recaptures"] = recaptures
    return pos, quat, history
