"""Batched rendering through the C ABI (mini_raytrace_ext.h section 5):
many engines progress together in one HIP graph launch."""
from __future__ import annotations

import ctypes as C

from . import _lib as L


def render_batch(engines, frames) -> dict:
    """Adds frames[i] samples per pixel to engines[i] (clamped to each spp
    limit) in one graph launch.  Returns the batch statistics and the frames
    each request rendered."""
    n = len(engines)
    reqs = (L.mrtxBatchRequest * n)()
    for i, (e, f) in enumerate(zip(engines, frames)):
        reqs[i].engine = e.handle.value
        reqs[i].frames = int(f)
    st = L.mrtxBatchStats()
    L.check(L.lib().mrtxRenderBatch(reqs, n, C.byref(st)), "mrtxRenderBatch")
    return {"captureMs": st.captureMs, "launchMs": st.launchMs, "requests": st.requests, "frames": st.frames,
            "mode": st.mode, "extend": (st.extendBlock, st.extendStack), "shadow": (st.shadowBlock, st.shadowStack),
            "framesDone": [reqs[i].framesDone for i in range(n)]}
