"""Device-buffer helpers for GPU tests (the library's own allocator)."""
import numpy as np

import aperture as ap


def trace(scene, rays, any_hit=False, stats=False, mode=None):
    """Runs the GPU closest/any-hit traversal on host rays; returns host hits (and stats)."""
    if mode is not None:
        scene.set_traversal_mode(mode)
    rays = np.ascontiguousarray(rays, dtype=ap.RAY_DTYPE)
    n = len(rays)
    d_rays = ap.DeviceBuffer.from_numpy(rays)
    d_hits = ap.DeviceBuffer(n * 8)
    d_stats = ap.DeviceBuffer(n * 8) if stats else None
    if any_hit:
        scene.trace_any(d_rays, d_hits, n=n)
    else:
        scene.trace_closest(d_rays, d_hits, n=n, stats=d_stats)
    ap.synchronize()
    hits = d_hits.to_numpy(ap.HIT_DTYPE, n)
    if stats:
        return hits, d_stats.to_numpy(ap.STATS_DTYPE, n)
    return hits
