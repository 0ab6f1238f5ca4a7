"""Helpers for the GPU parity tests: device setup, host<->device copies and
comparison statistics against the oracle."""
import numpy as np


def require_gpu():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device visible: GPU tests need an MI355X")
    return torch


def to_dev(a):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32).copy()).cuda()
    return torch.from_numpy(a.copy()).cuda()


def to_host(t, dtype=None):
    import torch

    torch.cuda.synchronize()
    a = t.cpu().numpy()
    if dtype is not None:
        a = a.view(dtype)
    return a


def ulp_diff(a, b):
    """|a - b| in units in the last place for float32 arrays of equal sign."""
    ai = a.view(np.int32).astype(np.int64)
    bi = b.view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return np.abs(ai - bi)


def hit_stats(gpu, ref, far=1e5):
    """Per-pixel agreement of two Hit buffers (float4 per pixel)."""
    g = gpu.reshape(-1, 4).astype(np.float64)
    r = ref.reshape(-1, 4).astype(np.float64)
    same_id = (g[:, 1] == r[:, 1]) & (g[:, 2] == r[:, 2])
    both = same_id & (g[:, 0] < far) & (r[:, 0] < far)
    rel = np.abs(g[both, 0] - r[both, 0]) / np.maximum(r[both, 0], 1e-30)
    objects = (r[:, 1] >= 0) | (g[:, 1] >= 0)
    return {
        "pixels": len(g),
        "id_match": float(same_id.mean()),
        "id_mismatch_px": int((~same_id).sum()),
        "object_px": int(objects.sum()),
        "object_id_match": float(same_id[objects].mean()) if objects.any() else 1.0,
        "t_rel_median": float(np.median(rel)) if rel.size else 0.0,
        "t_rel_p99": float(np.percentile(rel, 99)) if rel.size else 0.0,
        "t_rel_max": float(rel.max()) if rel.size else 0.0,
        "t_rel_le_1e5": float((rel <= 1e-5).mean()) if rel.size else 1.0,
        "same_id_mask": same_id,
    }
