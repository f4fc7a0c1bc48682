"""Autotuner for the stencil kernels.

Times each candidate launch shape of the Wilson stencil on the resident gauge
field with HIP events on the caller's stream and installs the fastest in the
context.  The tunable is the number of x0 sites each thread marches over
(``march``: 1 selects the one-site-per-thread kernel), for multi-RHS
launches the number of right-hand sides sharing each link load in a thread,
the site order of the one-site kernels (flat, x1 slabs or (x1, x2) tiles
walked x0-fastest, or compact 4D workgroup boxes) and the workgroup order
(XCD-contiguous or hardware order).
"""

from __future__ import annotations

from ctypes import byref, c_void_p
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
         stream: Optional[torch.cuda.Stream] = None, slabs: Iterable[int] = (0, 2, 4, 8),
         cols_outer: Iterable[int] = (0, 2, 3), boxes: Iterable[int] = (1, 2, 7),
         mrhs_marches: Iterable[int] = (2, 4, 8, 16), planes: Iterable[int] = (2, 4, 8),
         lds_boxes: Iterable[int] = (1, 2, 3),
         tiles: Iterable = ()) -> Dict[str, object]:
    """Pick the fastest launch shape for (prec, recon) at nrhs right-hand sides
    per configuration and install it.  Candidates: every march length dividing
    L0 with one right-hand side per thread, plus 2 and 4 right-hand sides per
    thread when nrhs > 1 (for FP16 also two per thread marching the lengths of
    `mrhs_marches` along x0, and the LDS-staged plane kernel marching the
    lengths of `planes` along x1); the one-site-per-thread shapes in flat and x1-slab
    site order (slab extents dividing L1) and in the (x1, x2) tile orders of
    `tiles` ((slab, slab_x2) pairs dividing L1 and L2, walked x0-fastest so
    that each XCD sweeps x0 over a compact block); each in XCD-contiguous and hardware
    workgroup order, and for nrhs > 1 in each column order of `cols_outer`
    (0: the columns of a tile back to back; n: blocks of n columns, or column
    groups, sweeping the lattice in turn, which keeps a block's spinor
    neighbours in L2 while its links stay shared; sweeping one column at a
    time, n = 1, re-reads the links and measured slower at 32^4).
    Returns {"march", "rhs_per_thread", "xcd_remap", "slab", "slab_x2", "block_shape", "plane", "lds_box",
    "cols_outer", "us": {label: us}}."""
    stream = stream or torch.cuda.current_stream()
    L0, L1, L2 = ctx.lattice_shape[0], ctx.lattice_shape[1], ctx.lattice_shape[2]
    multi = nrhs > 1
    gauge.prepare(prec, recon)
    # site orders of the one-site kernels: ("s", (slab, slab_x2)) or ("b", box preset)
    orders1 = [("s", (s, 0)) for s in slabs if s == 0 or (s < L1 and L1 % s == 0)]
    orders1 += [("s", (int(a), int(c))) for a, c in tiles
                if a <= L1 and L1 % a == 0 and c < L2 and L2 % c == 0]
    orders1 += [("b", b) for b in boxes if _box_fits(ctx.lattice_shape, b)]
    shapes = []
    for m in marches:
        if L0 % m:
            continue
        shapes += [(m, 1, o) for o in (orders1 if m == 1 else [("s", (0, 0))])]
    if prec == nat.WL_F16:
        # the LDS-staged plane kernel, marching the lengths of `planes` along x1
        L1v, L3v = ctx.lattice_shape[1], ctx.lattice_shape[3]
        if 2 * L3v <= 256:
            shapes += [(1, 1, ("p", p)) for p in planes if L1v % p == 0]
    if multi:
        shapes += [(1, k, o) for k in (2, 4) if nrhs % k == 0 for o in orders1]
        if prec == nat.WL_F16 and nrhs % 2 == 0:
            # the FP16 two-RHS kernel marching along x0
            shapes += [(m, 2, ("s", (0, 0))) for m in mrhs_marches if m > 1 and L0 % m == 0]
    if prec == nat.WL_F16 and ctx.nc == 3 and recon == 12:
        # the LDS box kernel (two columns per pass; 12-real links)
        shapes += [(1, 2 if multi else 1, ("l", b)) for b in lds_boxes if _lds_box_fits(ctx.lattice_shape, b)]
    corders = [int(c) for c in cols_outer] if multi else [0]

    eo_slab = ctx.get_tuning(prec, nrhs)["eo_slab"]  # kept: the even-odd hop is tuned on its own

    def install(m, k, r, o, co):
        ctx.set_tuning(prec, multi, m, k if multi else 0, bool(r), o[1][0] if o[0] == "s" else 0, co, 0,
                       o[1] if o[0] == "b" else 0, o[1] if o[0] == "p" else 0, eo_slab, o[1] if o[0] == "l" else 0,
                       o[1][1] if o[0] == "s" else 0)

    times = {}
    for co in corders:
        for r in (1, 0):
            for m, k, o in shapes:
                if o[0] == "l" and (co != corders[0] or r != 1):
                    continue  # the box kernel ignores the workgroup and column orders
                install(m, k, r, o, co)
                times[(m, k, r, o, co)] = time_dslash(ctx, gauge, prec, recon, fin, fout, nrhs, reps, stream)
    def as_dict(m, k, r, o, co):
        return {"march": m, "rhs_per_thread": k, "xcd_remap": r, "slab": o[1][0] if o[0] == "s" else 0,
                "slab_x2": o[1][1] if o[0] == "s" else 0,
                "block_shape": o[1] if o[0] == "b" else 0, "plane": o[1] if o[0] == "p" else 0,
                "lds_box": o[1] if o[0] == "l" else 0, "cols_outer": co}

    ranked = sorted(times, key=times.get)
    best = ranked[0]
    install(*best)
    m, k, r, o, co = best
    return {**as_dict(*best),
            "ranked": [as_dict(*c) for c in ranked],
            "us": {f"m{m}k{k}x{r}{_label(o)}" + (f"c{co}" if co else ""): t for (m, k, r, o, co), t in times.items()}}


def _label(o) -> str:
    """Short name of a site order: s<slab>[t<slab_x2>], b<box>, p<plane>, l<lds box>."""
    if o[0] == "s":
        return f"s{o[1][0]}" + (f"t{o[1][1]}" if o[1][1] else "")
    return f"{o[0]}{o[1]}"


def install_tuning(ctx: nat.Context, prec: int, multi: bool, t: Dict[str, object]) -> None:
    """Install a tuning dict as returned in `tune(...)["ranked"]` (the even-odd slab is kept)."""
    eo = ctx.get_tuning(prec, 2 if multi else 1)["eo_slab"]
    ctx.set_tuning(prec, multi, int(t["march"]), int(t["rhs_per_thread"]) if multi else 0, bool(t["xcd_remap"]),
                   int(t["slab"]), int(t["cols_outer"]), 0, int(t["block_shape"]), int(t["plane"]), eo,
                   int(t.get("lds_box", 0)), int(t.get("slab_x2", 0)))


def set_eo_slab(ctx: nat.Context, prec: int, multi: bool, eo_slab: int) -> None:
    """Set the x1 slab of the even-odd hop, keeping the other launch-shape settings."""
    ctx.set_eo_slab(prec, multi, eo_slab)


def tune_eo(ctx: nat.Context, gauge: nat.Gauge, b: torch.Tensor, x: torch.Tensor, nrhs: int, prec: int, recon: int,
            slabs: Iterable[int] = (0, 2, 4, 8), iters: int = 24, delta: float = 0.01,
            stream: Optional[torch.cuda.Stream] = None,
            update_slabs: Iterable[int] = (0, 2, 4, 8), update_iters: int = 8) -> Dict[str, object]:
    """Pick the x1 slab order of the even-odd hop for inner precision prec at
    nrhs right-hand sides per configuration and install it.  Each candidate
    (flat, or slab extents dividing L1) is timed on a solve of the even-odd
    normal equations (wl_cg_eo_normal) capped at `iters` iterations, with
    b and x the caller's complex128 reference-layout source and solution
    arrays; the fused hop, Schur combine and CG-step launches it times are
    the ones a full solve repeats; the solves capture HIP graphs, so they run
    on `stream` or, by default, on a side stream ordered after the current
    one.  For a mixed-precision prec the slab of the FP64 hops (the reliable
    updates, plain FP64 even-odd solves and FP64 even-odd operators at this
    nrhs share it) is then picked among `update_slabs`, with the inner slab
    installed (an empty `update_slabs` leaves it alone).  Those candidates
    are timed on an FP64-inner solve capped at `update_iters` iterations, in
    which every hop is an FP64 hop: in the mixed solve the update hops are
    idle (gated off) in all but a few iterations, so its time would rank
    them by dispatch noise.  Returns {"eo_slab", "ms": {slab: ms}} and, when
    the FP64 slab was tuned, {"update_eo_slab", "update_ms"}."""
    if stream is None:
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
    L1 = ctx.lattice_shape[1]
    multi = nrhs > 1
    ncol = ctx.batch * nrhs
    for name, t in (("b", b), ("x", x)):
        if (t.dtype != torch.complex128 or not t.is_cuda or not t.is_contiguous()
                or t.numel() != ncol * ctx.volume * 4 * ctx.nc):
            raise ValueError(f"{name} must be a contiguous complex128 CUDA tensor of {ncol} spinor fields")
    h = c_void_p(stream.cuda_stream)
    st = (nat.WlSolveStats * ncol)()

    def sweep(p, cands, solve_prec, n_iter):
        sp, keep = nat.solve_params(1e-30, n_iter, nat.WL_STOP_TRUE_RESIDUAL, solve_prec, recon, delta, 16, True)

        def run():
            nat.check(ctx._lib.wl_cg_eo_normal(ctx.h, gauge.h, c_void_p(b.data_ptr()), c_void_p(x.data_ptr()),
                                               c_void_p(0), nrhs, nat.WL_C128, byref(sp), st, h))

        times = {}
        for s in cands:
            if s and L1 % s:
                continue
            set_eo_slab(ctx, p, multi, s)
            run()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            best = None
            for _ in range(2):
                e0.record(stream)
                run()
                e1.record(stream)
                e1.synchronize()
                t = e0.elapsed_time(e1)
                best = t if best is None else min(best, t)
            times[s] = best
        del keep
        pick = min(times, key=times.get)
        set_eo_slab(ctx, p, multi, pick)
        return pick, {str(s): round(t, 3) for s, t in times.items()}

    pick, ms = sweep(prec, slabs, prec, iters)
    out = {"eo_slab": pick, "ms": ms}
    update_slabs = tuple(update_slabs)
    if prec != nat.WL_F64 and update_slabs:
        out["update_eo_slab"], out["update_ms"] = sweep(nat.WL_F64, update_slabs, nat.WL_F64, update_iters)
    torch.cuda.current_stream().wait_stream(stream)
    return out


# 4D workgroup box presets (log2 extents per direction), as in the library
BOX_SHAPES = {1: (3, 2, 0, 3), 2: (2, 2, 1, 3), 3: (1, 2, 2, 3), 4: (2, 1, 2, 3), 5: (3, 1, 1, 3),
              6: (4, 1, 0, 3), 7: (2, 2, 0, 4), 8: (3, 1, 0, 4)}


# LDS box kernel presets (log2 extents per direction), as in the library
LDS_BOX_SHAPES = {1: (1, 1, 2, 4), 2: (1, 2, 2, 3), 3: (2, 1, 2, 3)}


def _lds_box_fits(L, b: int) -> bool:
    sh = LDS_BOX_SHAPES.get(int(b))
    return sh is not None and all(int(L[mu]) % (1 << sh[mu]) == 0 for mu in range(4))


def _box_fits(L, b: int) -> bool:
    sh = BOX_SHAPES.get(int(b))
    return sh is not None and all(int(L[mu]) % (1 << sh[mu]) == 0 for mu in range(4))


def tune_march(ctx, gauge, prec, recon, fin, fout, nrhs: int = 1, candidates=(1, 2, 4, 8), reps: int = 20,
               stream=None) -> Dict[str, object]:
    """Single-RHS form of `tune`: {"best": march, "xcd_remap": r, "us": ...}."""
    t = tune(ctx, gauge, prec, recon, fin, fout, nrhs, candidates, reps, stream)
    return {"best": t["march"], "xcd_remap": t["xcd_remap"], "slab": t["slab"], "slab_x2": t["slab_x2"],
            "block_shape": t["block_shape"],
            "plane": t["plane"], "us": t["us"]}
