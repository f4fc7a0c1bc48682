"""Autotuner for the stencil kernels.

Times each candidate launch shape of the Wilson stencil on the resident gauge
field with HIP events on the caller's stream and installs the fastest in the
context.  The tunable is the number of x0 sites each thread marches over
(``march``: 1 selects the one-site-per-thread kernel), for multi-RHS
launches the number of right-hand sides sharing each link load in a thread,
and the workgroup order (XCD-contiguous or hardware order).
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


def tune(ctx: nat.Context, gauge: nat.Gauge, prec: int, recon: int, fin: nat.Field, fout: nat.Field,
         nrhs: int = 1, marches: Iterable[int] = (1, 2, 4, 8), reps: int = 20,
         stream: Optional[torch.cuda.Stream] = None) -> Dict[str, object]:
    """Pick the fastest launch shape for (prec, recon) at nrhs right-hand sides
    per configuration and install it.  Candidates: every march length dividing
    L0 with one right-hand side per thread, plus 2 and 4 right-hand sides per
    thread when nrhs > 1; each in XCD-contiguous and hardware workgroup order.
    Returns {"march", "rhs_per_thread", "xcd_remap", "us": {label: us}}."""
    stream = stream or torch.cuda.current_stream()
    L0 = ctx.lattice_shape[0]
    multi = nrhs > 1
    gauge.prepare(prec, recon)
    shapes = [(m, 1) for m in marches if L0 % m == 0]
    if multi:
        shapes += [(1, k) for k in (2, 4) if nrhs % k == 0]
    times = {}
    for r in (1, 0):
        for m, k in shapes:
            ctx.set_tuning(prec, multi, m, k if multi else 0, bool(r))
            times[(m, k, r)] = time_dslash(ctx, gauge, prec, recon, fin, fout, nrhs, reps, stream)
    best = min(times, key=times.get)
    ctx.set_tuning(prec, multi, best[0], best[1] if multi else 0, bool(best[2]))
    return {"march": best[0], "rhs_per_thread": best[1], "xcd_remap": best[2],
            "us": {f"m{m}k{k}x{r}": t for (m, k, r), t in times.items()}}


def tune_march(ctx, gauge, prec, recon, fin, fout, nrhs: int = 1, candidates=(1, 2, 4, 8), reps: int = 20,
               stream=None) -> Dict[str, object]:
    """Single-RHS form of `tune`: {"best": march, "xcd_remap": r, "us": ...}."""
    t = tune(ctx, gauge, prec, recon, fin, fout, nrhs, candidates, reps, stream)
    return {"best": t["march"], "xcd_remap": t["xcd_remap"], "us": t["us"]}
