"""Autotuner for the stencil kernels.

Times each candidate launch shape of the Wilson stencil on the resident gauge
field with HIP events on the caller's stream and installs the fastest in the
context.  The tunable is the number of x0 sites each thread marches over
(``march``: 1 selects the one-site-per-thread kernel) together with the
workgroup order (XCD-contiguous or hardware order).
"""

from __future__ import annotations

from ctypes import c_void_p
from typing import Dict, Iterable, Optional

import torch

from . import _native as nat


def time_dslash(ctx: nat.Context, gauge: nat.Gauge, prec: int, recon: int, fin: nat.Field, fout: nat.Field,
                nrhs: int, reps: int, stream: torch.cuda.Stream, op: int = nat.WL_OP_D) -> float:
    """Average microseconds per application, HIP events on ``stream``."""
    lib = ctx._lib
    h = c_void_p(stream.cuda_stream)
    for _ in range(2):
        nat.check(lib.wl_dslash(ctx.h, gauge.h, op, recon, fin.h, fout.h, nrhs, h))
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        nat.check(lib.wl_dslash(ctx.h, gauge.h, op, recon, fin.h, fout.h, nrhs, h))
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) * 1e3 / reps


def tune_march(ctx: nat.Context, gauge: nat.Gauge, prec: int, recon: int, fin: nat.Field, fout: nat.Field,
               nrhs: int = 1, candidates: Iterable[int] = (1, 2, 4, 8), reps: int = 20,
               stream: Optional[torch.cuda.Stream] = None) -> Dict[str, object]:
    """Pick the fastest (march, XCD order) for (prec, recon); returns
    {"best": m, "xcd_remap": r, "us": {"m<m>x<r>": t}}."""
    stream = stream or torch.cuda.current_stream()
    L0 = ctx.lattice_shape[0]
    gauge.prepare(prec, recon)
    times = {}
    for r in (1, 0):
        ctx.set_xcd_remap(prec, bool(r))
        for m in candidates:
            if L0 % m:
                continue
            ctx.set_march(prec, m)
            times[(m, r)] = time_dslash(ctx, gauge, prec, recon, fin, fout, nrhs, reps, stream)
    best = min(times, key=times.get)
    ctx.set_march(prec, best[0])
    ctx.set_xcd_remap(prec, bool(best[1]))
    return {"best": best[0], "xcd_remap": best[1], "us": {f"m{m}x{r}": t for (m, r), t in times.items()}}
