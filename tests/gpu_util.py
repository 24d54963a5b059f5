"""Device-buffer helpers for GPU tests (torch is only the allocator)."""
import numpy as np
import torch

import aperture as ap


def to_dev(arr):
    """Copies a numpy (structured) array into a device byte tensor."""
    a = np.ascontiguousarray(arr)
    t = torch.empty(max(a.nbytes, 16), dtype=torch.uint8, device="cuda")
    if a.nbytes:
        t[: a.nbytes].copy_(torch.from_numpy(a.view(np.uint8).reshape(-1)))
    return t


def from_dev(t, dtype, n):
    nbytes = n * np.dtype(dtype).itemsize
    host = t[:nbytes].cpu().numpy()
    return host.view(dtype).copy()


def trace(scene, rays, any_hit=False, stats=False, mode=None):
    """Runs the GPU closest/any-hit traversal on host rays; returns host hits (and stats)."""
    if mode is not None:
        scene.set_traversal_mode(mode)
    n = len(rays)
    d_rays = to_dev(np.ascontiguousarray(rays, dtype=ap.RAY_DTYPE))
    d_hits = torch.empty(max(n * 8, 16), dtype=torch.uint8, device="cuda")
    d_stats = torch.empty(max(n * 8, 16), dtype=torch.uint8, device="cuda") if stats else None
    stream = torch.cuda.current_stream()
    if any_hit:
        scene.trace_any(d_rays, d_hits, n=n, stream=stream)
    else:
        scene.trace_closest(d_rays, d_hits, n=n, stream=stream, stats=d_stats)
    torch.cuda.synchronize()
    hits = from_dev(d_hits, ap.HIT_DTYPE, n)
    if stats:
        return hits, from_dev(d_stats, ap.STATS_DTYPE, n)
    return hits
