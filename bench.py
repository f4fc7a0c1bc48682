"""Benchmark: CG solves/s and Dslash HBM GB/s on a 32^4 SU(3) Wilson lattice,
mixed-precision multi-RHS CG vs the plain FP64 single-RHS path, true residual
1e-10 (recomputed in FP64).

Workload (one step): one batched solve of D^dag D x = b for k = 12 complex
Gaussian right-hand sides on a resident 32^4 gauge field (eps = 0.6,
m = 0.02, boundary phases (1,1,1,-1)), inner operator FP16 storage with
12-real link reconstruction, FP64 reliable updates; inputs already in HBM.
``value`` = right-hand sides solved per second.  Beside it: the plain FP64
single-RHS solve rate, every Dslash storage variant (autotuned launch shape),
the 24^3x48 multi-RHS case k in {1, 4, 12, 24}, and the CPU baseline.
Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from ctypes import byref, c_int, c_void_p
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "CG solves/s + Dslash HBM GB/s, 32^4 SU(3) Wilson; mixed vs FP64 at 1e-10 res"
PEAK_HBM_GBS = 8000.0
FLOP_PER_SITE = 1320


def dslash_bytes_per_site(prec: str, recon: int, nrhs: int = 1, nc: int = 3) -> int:
    """Compulsory HBM bytes per site of one application: links read once,
    each right-hand side's input read once and output written once (FP16
    adds the per-site FP32 scale of input and output)."""
    spin_reals = 24 if nc == 3 else 8
    link_reals = (12 if recon == 12 else 18) if nc == 3 else 2
    b = {"f64": 8, "f32": 4, "f16": 2}[prec]
    spin = 2 * spin_reals * b + (8 if prec == "f16" else 0)
    return 4 * link_reals * b + nrhs * spin


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lattice", type=str, default="32,32,32,32")
    ap.add_argument("--eps", type=float, default=0.6)
    ap.add_argument("--mass", type=float, default=0.02)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--nrhs", type=int, default=12)
    ap.add_argument("--inner", type=str, default="f16", help="inner precision of the mixed solve")
    ap.add_argument("--recon", type=int, default=12)
    ap.add_argument("--delta", type=float, default=0.01, help="reliable-update trigger of the headline solve")
    ap.add_argument("--eo-delta", type=float, default=0.005,
                    help="reliable-update trigger of the even-odd solves (the Schur system's residual falls faster per "
                         "iteration; at 32^4 0.005 saves one update against 0.01 at the same 27 iterations)")
    ap.add_argument("--accum", type=str, default="f16", help="inner solution accumulator of the headline solve: f16 | auto (f32) | f64")
    ap.add_argument("--search", type=str, default="recurrence",
                    help="search-direction update of the mixed solves: recurrence (t = D r + beta t, two launches "
                         "per iteration) | xpay (p = r + beta p in its own pass, three)")
    ap.add_argument("--poll", type=int, default=16)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--dslash-reps", type=int, default=50)
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--no-config4", action="store_true", help="skip the 24^3x48 multi-RHS case")
    ap.add_argument("--no-config2", action="store_true", help="skip the 16^4 FP64 single-RHS case")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving scenario")
    ap.add_argument("--no-propagator", action="store_true",
                    help="skip the point-source propagator and chronological-guess measurements")
    ap.add_argument("--serving-requests", type=int, default=60, help="requests of the 200/s overload point")
    ap.add_argument("--serving-sweep-requests", type=int, default=200,
                    help="requests per point of the sustainable-load sweep (0 skips the sweep)")
    ap.add_argument("--serving-sweep-loads", type=str, default="0.25,0.5,0.75",
                    help="arrival rates of the sweep as fractions of the request rate served at overload")
    ap.add_argument("--detail-out", type=str, default="gpurun_out/bench_detail.json",
                    help="file for the full record (every variant, tuning table and sub-benchmark); the last "
                         "stdout line is the compact record")
    ap.add_argument("--no-stream", action="store_true", help="skip the HBM streaming probes")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lime", action="store_true", help="skip the SciDAC payload decode")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the 8^3x16 workloads of the reference's published CPU timings")
    ap.add_argument("--profile-only", action="store_true",
                    help="run only the Dslash variants and a calibration copy (for rocprofv3 counter passes)")
    ap.add_argument("--solve-trace", action="store_true",
                    help="tune as the full run does, then run the headline solve once more after a 1 s pause "
                         "(for a kernel trace of that solve alone) and stop")
    ap.add_argument("--eo-trace", action="store_true",
                    help="tune the even-odd hop as the full run does, run the mixed k-RHS even-odd solve, pause 1 s, "
                         "run it again (for a kernel trace of that solve alone) and stop")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Bench:
    def __init__(self, args):
        import torch

        from jaxqft_amd import _native as nat

        self.torch = torch
        self.nat = nat
        self.lib = nat.load()
        self.args = args
        self.stream = torch.cuda.Stream()  # non-default stream (graph capture needs one)
        self.h = c_void_p(self.stream.cuda_stream)

    def solve(self, ctx, gauge, b, x, nrhs, prec, recon, masses=None, delta=None):
        nat = self.nat
        a = self.args
        sp, keep = nat.solve_params(a.tol, 20000, nat.WL_STOP_TRUE_RESIDUAL, nat.PREC[prec], recon,
                                    a.delta if delta is None else delta, a.poll, bool(a.graph), masses,
                                    accum=a.accum if delta is None else "auto", search=a.search)
        st = (nat.WlSolveStats * b.ncol)()
        nat.check(self.lib.wl_cg_normal_field(ctx.h, gauge.h, b.h, x.h, 0, nrhs, byref(sp), st, self.h))
        del keep
        return nat.stats_list(st, b.ncol)

    def solve_eo(self, ctx, gauge, bt, xt, nrhs, prec, recon):
        """Even-odd normal equations (wl_cg_eo_normal) on reference-layout device arrays."""
        nat = self.nat
        a = self.args
        sp, keep = nat.solve_params(a.tol, 20000, nat.WL_STOP_TRUE_RESIDUAL, nat.PREC[prec], recon, a.eo_delta, a.poll,
                                    bool(a.graph), None, accum=a.accum, search=a.search)
        st = (nat.WlSolveStats * nrhs)()
        nat.check(self.lib.wl_cg_eo_normal(ctx.h, gauge.h, c_void_p(bt.data_ptr()), c_void_p(xt.data_ptr()), c_void_p(0),
                                           nrhs, nat.WL_C128, byref(sp), st, self.h))
        del keep
        return nat.stats_list(st, nrhs)

    def timed(self, n, fn, *fargs):
        torch = self.torch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            st = fn(*fargs)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n, st

    def timed_solves(self, n, *sargs):
        torch = self.torch
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            st = self.solve(*sargs)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n, st


PROFILE_TRAFFIC = Path(__file__).resolve().parent / "profiles" / "traffic_dominant_kernel.json"
PROFILE_CANDIDATES = 8  # leading tuner candidates of the dominant kernel covered by the counter runs


def measured_traffic(args, k, L, tuning):
    """HBM bytes per launch of the dominant kernel (FETCH_SIZE + WRITE_SIZE,
    calibrated against a device copy) from the committed rocprofv3 summary,
    when that summary is of this workload and launch shape; else None."""
    try:
        d = json.loads(PROFILE_TRAFFIC.read_text())
    except (OSError, ValueError):
        return None, None
    want = {"lattice": list(L), "nrhs": k, "inner": args.inner, "recon": args.recon}
    if any(d.get(key) != val for key, val in want.items()):
        return None, None
    keys = ("march", "rhs_per_thread", "xcd_remap", "slab", "slab_x2", "block_shape", "cols_outer", "plane")
    for e in d.get("entries", []):
        t = e.get("tuning", {})
        if all(int(t.get(key, 0)) == int(tuning.get(key, 0)) for key in keys):
            return int(e["fetch_bytes"] + e["write_bytes"]), "profiles/" + PROFILE_TRAFFIC.name
    return None, None


PROFILE_TRAFFIC_CG = Path(__file__).resolve().parent / "profiles" / "traffic_cg_step.json"


def measured_traffic_cg(args, k, L):
    """HBM bytes per launch of the D^dag + CG-step stencil (FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy of known size) from the committed
    rocprofv3 summary, when that summary is of this workload; else None."""
    try:
        d = json.loads(PROFILE_TRAFFIC_CG.read_text())
    except (OSError, ValueError):
        return None, None
    want = {"lattice": list(L), "nrhs": k, "inner": args.inner, "recon": args.recon, "accumulator": args.accum,
            "search": args.search}
    if any(d.get(key) != val for key, val in want.items()):
        return None, None
    return int(d["fetch_bytes_per_launch"] + d["write_bytes_per_launch"]), "profiles/" + PROFILE_TRAFFIC_CG.name


def summarize(stats):
    return {
        "iters_max": max(s["iters"] for s in stats),
        "iters_mean": round(float(np.mean([s["iters"] for s in stats])), 1),
        "status": sorted(set(s["status"] for s in stats)),
        "true_resid_max": max(s["true_resid"] for s in stats),
        "reliable_updates_max": max(s["reliable_updates"] for s in stats),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}): the Wilson solver here runs on a "
              "single GPU, and sharding the lattice over GPUs (a halo exchange per stencil application) "
              "is not built.", file=sys.stderr)
        sys.exit(2)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields
    from jaxqft_amd.autotune import install_tuning, time_dslash, tune, tune_eo, tune_march

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    t_start = time.perf_counter()
    B = Bench(args)
    stream = B.stream
    L = tuple(int(v) for v in args.lattice.split(","))
    V = int(np.prod(L))
    k = args.nrhs
    bc = (1, 1, 1, -1)

    t0 = time.perf_counter()
    rng = fields.rng_for(fields.SEED_BASE + 2)
    U = fields.su3_hot_start(rng, 1, L, args.eps)
    b_host = fields.random_fermion(rng, (k, *L, 4, 3))
    log(f"generated fields {L} in {time.perf_counter() - t0:.1f}s")

    with torch.cuda.stream(stream):
        ctx = nat.Context(L, 3, 1, args.mass, 1.0, bc)
        gauge = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
        bk_t = torch.from_numpy(b_host).cuda()
        b1_t = bk_t[:1].contiguous()
        torch.cuda.synchronize()

        # ------------------------------------------------ Dslash (one RHS)
        dslash = {}
        for prec, recon in [("f64", 18), ("f64", 12), ("f32", 18), ("f32", 12), ("f16", 18), ("f16", 12)]:
            pc = nat.PREC[prec]
            fi, fo = nat.Field(ctx, pc, 1), nat.Field(ctx, pc, 1)
            fi.import_(b1_t)
            t1 = tune_march(ctx, gauge, pc, recon, fi, fo, reps=max(3, args.dslash_reps // 5), stream=stream)
            us = time_dslash(ctx, gauge, pc, recon, fi, fo, 1, args.dslash_reps, stream)
            bps = dslash_bytes_per_site(prec, recon)
            gbs = bps * V / (us * 1e-6) / 1e9
            dslash[f"{prec}_r{recon}"] = {
                "us": round(us, 2), "GB/s": round(gbs, 1), "frac_hbm_peak": round(gbs / PEAK_HBM_GBS, 4),
                "GFLOP/s": round(FLOP_PER_SITE * V / (us * 1e-6) / 1e9, 1), "bytes_per_site": bps,
                "march": t1["best"], "xcd_remap": t1["xcd_remap"], "slab": t1["slab"], "block_shape": t1["block_shape"],
                "plane": t1["plane"], "tune_us": {m: round(t, 2) for m, t in t1["us"].items()},
            }
            if prec == "f64" and recon == 18:
                # context only: the no-reuse convention (every neighbour spinor and link read, 360 reals/site)
                dslash[f"{prec}_r{recon}"]["GB/s_no_reuse_2880B"] = round(2880 * V / (us * 1e-6) / 1e9, 1)
            log(f"dslash {prec} r{recon}: {us:.1f} us {gbs:.0f} GB/s march {t1['best']} slab {t1['slab']} "
                f"box {t1['block_shape']}")
            del fi, fo

        if args.profile_only:
            # the headline's inner stencil: k right-hand sides, tuned as in the
            # full run; the last dslash_reps stencil launches before the copy
            # are the timed ones
            pc = nat.PREC[args.inner]
            fi, fo = nat.Field(ctx, pc, k), nat.Field(ctx, pc, k)
            fi.import_(bk_t)
            gauge.prepare(pc, args.recon)
            tk = tune(ctx, gauge, pc, args.recon, fi, fo, nrhs=k, reps=3, stream=stream)
            # the tuner's leading candidates, each timed in turn (2 warm-up and
            # dslash_reps timed launches apiece), so that the counter runs cover
            # whichever of them a full run picks
            cands = []
            for t in tk["ranked"][:PROFILE_CANDIDATES]:
                install_tuning(ctx, pc, True, t)
                cands.append({"tuning": ctx.get_tuning(pc, k),
                              "us": round(time_dslash(ctx, gauge, pc, args.recon, fi, fo, k, args.dslash_reps,
                                                      stream), 2)})
            install_tuning(ctx, pc, True, tk["ranked"][0])
            dslash[f"{args.inner}_r{args.recon}_k{k}"] = {"candidates": cands, "timed_launches": args.dslash_reps}
            del fi, fo
            lime_prof = lime_decode(B, args, L, False)
            src = torch.empty(32 << 20, dtype=torch.float64, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(5):
                dst.copy_(src)
            torch.cuda.synchronize()
            print(json.dumps({"profile_only": True, "dslash": dslash, "lime_decode": lime_prof,
                              "copy_bytes_each_way": src.numel() * 8}))
            return

        # ---------------------------------------- Dslash (k RHS, shared links)
        dslash_k = {}
        for prec, recon in [("f32", 12), ("f16", 12), ("f64", 18)]:
            pc = nat.PREC[prec]
            fi, fo = nat.Field(ctx, pc, k), nat.Field(ctx, pc, k)
            fi.import_(bk_t)
            gauge.prepare(pc, recon)
            tk = tune(ctx, gauge, pc, recon, fi, fo, nrhs=k, reps=max(3, args.dslash_reps // 10), stream=stream)
            us = time_dslash(ctx, gauge, pc, recon, fi, fo, k, max(5, args.dslash_reps // 5), stream)
            best = (tk["rhs_per_thread"], us)
            tried = {m: round(t, 2) for m, t in tk["us"].items()}
            bps = dslash_bytes_per_site(prec, recon, k)
            gbs = bps * V / (best[1] * 1e-6) / 1e9
            dslash_k[f"{prec}_r{recon}"] = {
                "nrhs": k, "rhs_per_thread": best[0], "march": tk["march"], "xcd_remap": tk["xcd_remap"],
                "slab": tk["slab"], "slab_x2": tk["slab_x2"], "block_shape": tk["block_shape"],
                "us": round(best[1], 2), "us_per_rhs": round(best[1] / k, 2),
                "GB/s": round(gbs, 1), "frac_hbm_peak": round(gbs / PEAK_HBM_GBS, 4), "bytes_per_site": bps,
                "tune_us": tried,
            }
            log(f"dslash {prec} r{recon} k={k}: {best[1]:.1f} us ({best[1]/k:.1f}/rhs) {gbs:.0f} GB/s, "
                f"{len(tried)} shapes tuned (table in the full record)")
            del fi, fo

        if args.eo_trace:
            # the mixed k-RHS even-odd solve of the full run (tuned the same
            # way), once warm, then after a pause that marks the trace
            xek = torch.empty_like(bk_t)
            eo_tune = tune_eo(ctx, gauge, bk_t, xek, k, nat.PREC[args.inner], args.recon, delta=args.eo_delta,
                              stream=stream)
            B.solve_eo(ctx, gauge, bk_t, xek, k, args.inner, args.recon)
            torch.cuda.synchronize()
            time.sleep(1.0)
            t, st = B.timed(1, B.solve_eo, ctx, gauge, bk_t, xek, k, args.inner, args.recon)
            log(f"traced mixed {k}-RHS even-odd solve {t*1e3:.1f} ms {summarize(st)}")
            print(json.dumps({"eo_trace": True, "ms": round(t * 1e3, 2), "inner_hop_eo_slab": eo_tune["eo_slab"],
                              "update_hop_eo_slab": eo_tune.get("update_eo_slab"), **summarize(st)}))
            return

        if args.solve_trace:
            # one warm headline solve; the calibration copy (1 GiB read and
            # written by the streaming probe, for counter runs); a pause that
            # marks the trace; the traced solve
            bk = nat.Field(ctx, nat.WL_F64, k)
            bk.import_(bk_t)
            xk = nat.Field(ctx, nat.WL_F64, k)
            B.solve(ctx, gauge, bk, xk, k, args.inner, args.recon)
            torch.cuda.synchronize()
            calib = torch.empty(2 << 27, dtype=torch.float64, device="cuda")
            nat.check(B.lib.wl_stream_probe(nat.WL_STREAM_COPY, 1, 1, c_void_p(calib.data_ptr()),
                                            c_void_p(calib.data_ptr() + (1 << 30)), 1 << 30, 0, B.h))
            torch.cuda.synchronize()
            del calib
            time.sleep(1.0)
            t, st = B.timed_solves(1, ctx, gauge, bk, xk, k, args.inner, args.recon)
            log(f"traced mixed {k}-RHS solve {t*1e3:.1f} ms {summarize(st)}")
            print(json.dumps({"solve_trace": True, "ms": round(t * 1e3, 2), "calibration_copy_bytes_each_way": 1 << 30,
                              **summarize(st)}))
            return

        # ------------------------------------------------------------ solves
        b1 = nat.Field(ctx, nat.WL_F64, 1)
        b1.import_(b1_t)
        x1 = nat.Field(ctx, nat.WL_F64, 1)
        B.solve(ctx, gauge, b1, x1, 1, "f64", 18)
        fp64_s, st64 = B.timed_solves(args.fp64_steps, ctx, gauge, b1, x1, 1, "f64", 18)
        log(f"fp64 single-RHS solve {fp64_s*1e3:.1f} ms {summarize(st64)}")
        x1m = nat.Field(ctx, nat.WL_F64, 1)
        B.solve(ctx, gauge, b1, x1m, 1, args.inner, args.recon)
        mix1_s, stm1 = B.timed_solves(args.fp64_steps, ctx, gauge, b1, x1m, 1, args.inner, args.recon)
        log(f"mixed single-RHS solve {mix1_s*1e3:.1f} ms {summarize(stm1)}")
        xref = torch.empty_like(b1_t)
        x1.export(xref)

        # config 3: FP32 / FP16 inner, reconstruction on/off, delta = 0.1
        config3 = {}
        xc = nat.Field(ctx, nat.WL_F64, 1)
        xct = torch.empty_like(b1_t)
        for prec, recon in [("f32", 18), ("f32", 12), ("f16", 18), ("f16", 12)]:
            B.solve(ctx, gauge, b1, xc, 1, prec, recon, delta=0.1)
            t, st = B.timed_solves(1, ctx, gauge, b1, xc, 1, prec, recon, None, 0.1)
            xc.export(xct)
            torch.cuda.synchronize()
            err = float(torch.linalg.norm(xct - xref) / torch.linalg.norm(xref))
            config3[f"{prec}_r{recon}"] = {**summarize(st), "s_per_solve": round(t, 4), "rel_error_vs_fp64": err}
            log(f"config3 {prec} r{recon} delta 0.1: {t*1e3:.1f} ms {summarize(st)} err {err:.2e}")
        del xc

        bk = nat.Field(ctx, nat.WL_F64, k)
        bk.import_(bk_t)
        xk = nat.Field(ctx, nat.WL_F64, k)
        for _ in range(args.warmup):
            B.solve(ctx, gauge, bk, xk, k, args.inner, args.recon)
        mixk_s, stmk = B.timed_solves(args.steps, ctx, gauge, bk, xk, k, args.inner, args.recon)
        log(f"mixed {k}-RHS solve {mixk_s*1e3:.1f} ms {summarize(stmk)}")

        # the same solve once more with kernel timing: host-driven iterations
        # with HIP events on the bench stream around each iteration's launches
        ctx.set_profiling(True)
        xp = nat.Field(ctx, nat.WL_F64, k)
        stp = B.solve(ctx, gauge, bk, xp, k, args.inner, args.recon)
        prof = ctx.profile()
        ctx.set_profiling(False)
        del xp
        log(f"profiled {k}-RHS solve {summarize(stp)}: {prof}")

        # error of the mixed solutions against the plain FP64 solve
        xa = torch.empty_like(b1_t)
        xb = torch.empty_like(b1_t)
        xkt = torch.empty_like(bk_t)
        x1m.export(xa)
        x1.export(xb)
        xk.export(xkt)
        torch.cuda.synchronize()
        err1 = float(torch.linalg.norm(xa - xb) / torch.linalg.norm(xb))
        errk = float(torch.linalg.norm(xkt[:1] - xb) / torch.linalg.norm(xb))

        # even-odd preconditioned normal equations (solve_eop_even_normal): the
        # same sources' even parts, FP64 and mixed, one and k right-hand sides
        eo = {}
        with torch.cuda.stream(stream):
            xe1 = torch.empty_like(b1_t)
            xek = torch.empty_like(bk_t)
        # x1 slab order of the inner-precision hop, timed on capped solves
        eo_tune = tune_eo(ctx, gauge, bk_t, xek, k, nat.PREC[args.inner], args.recon, delta=args.eo_delta,
                          stream=stream)
        eo["inner_hop_eo_slab"] = eo_tune["eo_slab"]
        eo["inner_hop_tuning_ms"] = eo_tune["ms"]
        eo["update_hop_eo_slab"] = eo_tune.get("update_eo_slab")
        eo["update_hop_tuning_ms"] = eo_tune.get("update_ms")
        log(f"even-odd inner hop: eo_slab {eo_tune['eo_slab']} {eo_tune['ms']}; "
            f"FP64 update hop: eo_slab {eo_tune.get('update_eo_slab')} {eo_tune.get('update_ms')}")
        # the single-RHS launches have settings of their own: tuned the same way on one column
        eo_tune1 = tune_eo(ctx, gauge, b1_t, xe1, 1, nat.PREC[args.inner], args.recon, delta=args.eo_delta,
                           stream=stream)
        eo["single_rhs_tuning"] = eo_tune1
        log(f"even-odd single-RHS: {eo_tune1}")
        for name, bt, xt, nr, prec, recon, n in (("fp64_single_rhs", b1_t, xe1, 1, "f64", 18, args.fp64_steps),
                                                ("mixed_single_rhs", b1_t, xe1, 1, args.inner, args.recon,
                                                 args.fp64_steps),
                                                ("mixed_multi_rhs", bk_t, xek, k, args.inner, args.recon, args.steps)):
            B.solve_eo(ctx, gauge, bt, xt, nr, prec, recon)
            t, st = B.timed(n, B.solve_eo, ctx, gauge, bt, xt, nr, prec, recon)
            eo[name] = {**summarize(st), "nrhs": nr, "s_per_call": round(t, 4), "solves_per_s": round(nr / t, 3)}
            log(f"even-odd {name}: {t*1e3:.1f} ms {summarize(st)}")
        eo["speedup_vs_full_fp64_single"] = round(eo["fp64_single_rhs"]["solves_per_s"] * fp64_s, 3)
        eo["speedup_vs_full_mixed_multi"] = round(eo["mixed_multi_rhs"]["solves_per_s"] * mixk_s / k, 3)
        eo["system"] = "S_e^dag S_e x_e = b_e, normalized even Schur complement, FP64 true residual %g" % args.tol
        eo["reliable_delta"] = args.eo_delta
        del xe1, xek

        # analytic pseudofermion force from a solve (X, Y resident, FP64,
        # reference layouts): HIP events around back-to-back launches
        with torch.cuda.stream(stream):
            Xf = bk_t[:1].contiguous()
            Yf = bk_t[1:2].contiguous()
            Ff = torch.empty((1, 4, *L, 3, 3), dtype=torch.complex128, device="cuda")
            for _ in range(2):
                nat.check(B.lib.wl_pf_force(ctx.h, gauge.h, c_void_p(Xf.data_ptr()), c_void_p(Yf.data_ptr()),
                                            c_void_p(Ff.data_ptr()), nat.WL_LAYOUT_BMXYIJ, nat.WL_C128, B.h))
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(stream)
            nrep = max(5, args.dslash_reps // 5)
            for _ in range(nrep):
                nat.check(B.lib.wl_pf_force(ctx.h, gauge.h, c_void_p(Xf.data_ptr()), c_void_p(Yf.data_ptr()),
                                            c_void_p(Ff.data_ptr()), nat.WL_LAYOUT_BMXYIJ, nat.WL_C128, B.h))
            ev1.record(stream)
            ev1.synchronize()
        force_us = ev0.elapsed_time(ev1) * 1e3 / nrep
        force_bytes = (2 * 192 + 4 * 144 + 4 * 144) * V  # X, Y, links read once; F written once
        force = {"us": round(force_us, 2), "GB/s": round(force_bytes / (force_us * 1e-6) / 1e9, 1),
                 "frac_hbm_peak": round(force_bytes / (force_us * 1e-6) / 1e9 / PEAK_HBM_GBS, 4),
                 "bytes_per_site": 2 * 192 + 8 * 144}
        log(f"force: {force_us:.1f} us {force['GB/s']} GB/s")
        del Xf, Yf, Ff

        # inner stencil of the headline solve, timed on its own (HIP events)
        pc = nat.PREC[args.inner]
        fi, fo = nat.Field(ctx, pc, k), nat.Field(ctx, pc, k)
        fi.import_(bk_t)
        dom_us = time_dslash(ctx, gauge, pc, args.recon, fi, fo, k, args.dslash_reps, stream)
        dom_bytes = dslash_bytes_per_site(args.inner, args.recon, k) * V
        dom_gbs = dom_bytes / (dom_us * 1e-6) / 1e9
        dom_tuning = ctx.get_tuning(pc, k)
        dom_traffic, dom_traffic_src = measured_traffic(args, k, L, dom_tuning)
        del fi, fo, bk, xk, b1, x1, x1m
        torch.cuda.synchronize()
        copy = device_copy_rate(stream)
        log(f"device copy: {copy['GB/s']} GB/s ({copy['frac_hbm_peak']} of peak)")
        for entry in list(dslash.values()) + list(dslash_k.values()):
            entry["frac_of_device_copy"] = round(entry["GB/s"] / copy["GB/s"], 4)

    cg_step = cg_step_roofline(args, V, k, prof, stp)
    cg_step["traffic"], cg_step["traffic_source"] = measured_traffic_cg(args, k, L)
    detail = {
        "metric": METRIC,
        "value": round(k / mixk_s, 4),
        "unit": "solves/s",
        "ms_per_step": round(mixk_s * 1e3, 3),
        "config": {"lattice": list(L), "nrhs": k, "inner": args.inner, "recon": args.recon, "delta": args.delta,
                   "accumulator": args.accum, "graph": bool(args.graph)},
        "cg_step_kernels": cg_step,
        "dslash_launch": {
            "kernel": f"stencil {args.inner} r{args.recon}, {k} RHS, store epilogue, tuning {dom_tuning}",
            "us": round(dom_us, 2), "GB/s": round(dom_gbs, 1), "frac_hbm_peak": round(dom_gbs / PEAK_HBM_GBS, 4),
            "algorithmic_bytes": dom_bytes, "traffic": dom_traffic, "traffic_source": dom_traffic_src,
            "frac_of_device_copy": round(dom_gbs / copy["GB/s"], 4),
        },
        "device_copy": copy,
        "solve": {
            "fp64_single_rhs": {**summarize(st64), "s_per_solve": round(fp64_s, 4), "solves_per_s": round(1 / fp64_s, 3),
                                # whole solve / iterations (set-up and the final true residual included)
                                "us_per_iteration": round(fp64_s / st64[0]["iters"] * 1e6, 1),
                                "cg_iteration_GB/s": round(3840 * V / (fp64_s / st64[0]["iters"]) / 1e9, 1),
                                "cg_iteration_frac_hbm_peak": round(3840 * V / (fp64_s / st64[0]["iters"]) / 1e9
                                                                    / PEAK_HBM_GBS, 4),
                                "cg_iteration_bytes_per_site": 3840},
            "mixed_single_rhs": {**summarize(stm1), "s_per_solve": round(mix1_s, 4),
                                 "solves_per_s": round(1 / mix1_s, 3), "rel_error_vs_fp64": err1},
            "mixed_multi_rhs": {**summarize(stmk), "nrhs": k, "s_per_batch": round(mixk_s, 4),
                                "solves_per_s": round(k / mixk_s, 3), "rel_error_vs_fp64_col0": errk},
            "speedup_mixed_multi_vs_fp64_single": round((k / mixk_s) / (1 / fp64_s), 3),
            "speedup_mixed_single_vs_fp64_single": round(fp64_s / mix1_s, 3),
        },
        "even_odd": eo,
        "pseudofermion_force_fp64": force,
        "config3_single_rhs_delta0.1": config3,
        "dslash_single_rhs": dslash,
        "dslash_multi_rhs": dslash_k,
    }
    if not args.no_stream:
        detail["device_stream"] = device_stream(B)
    if not args.no_config2:
        detail["fp64_16x16x16x16"] = config2(B, args, not args.no_cpu)
    if not args.no_config4:
        detail["multi_rhs_24x24x24x48"] = config4(B, args)
    if not args.no_propagator:
        detail.update(propagator_chrono(B, args, U, L))
    if not args.no_serving:
        detail["serving"] = serving(B, args)
    if not args.no_published:
        detail["published_workloads_8x8x8x16"] = published_workloads(B, args)
    if not args.no_lime:
        detail["lime_decode"] = lime_decode(B, args, L, not args.no_cpu)
    if not args.no_cpu:
        detail["cpu_baseline"] = cpu_baseline(U, b_host[:1], L, args, st64[0]["iters"])
    detail["bench_wall_s"] = round(time.perf_counter() - t_start, 1)
    try:
        outp = Path(args.detail_out)
        outp.parent.mkdir(parents=True, exist_ok=True)
        outp.write_text(json.dumps(detail, indent=1))
        log(f"full record written to {outp}")
    except OSError as e:
        log(f"could not write the full record: {e}")
    line = compact_record(args, detail, L, k)
    text = json.dumps(line, separators=(",", ":"))
    log(f"compact record: {len(text)} bytes")
    print(text, flush=True)


def cg_step_roofline(args, V, k, prof, stats):
    """Algorithmic bytes and kernel times of the two stencil launches of the
    headline CG iteration, from the profiled solve (HIP events on the bench
    stream around every launch of its iterations).

    Per site, links counted once (12-real FP16: 96 B) and, per right-hand
    side, each spinor field moved once in the storage precision (FP16: 48 B
    of halves + 4 B scale).  Search direction by recurrence (the default):
      D with |t|^2 (t = D r + beta t_prev):  links + k (r read + t_prev read + t written)
      D^dag with the CG step:                links + k (t read + r read + r written
                                             + p_prev read + p written)
                                             plus, every second iteration (pairwise y),
                                             k (y read + y written)
    With p = r + beta p in a pass of its own (--search xpay):
      D with |t|^2 (t = D p):   links + k (p read + t written)
      D^dag with the CG step:   links + k (t read + r read + r written)
                                plus, every second iteration (pairwise y),
                                k (p read + p_prev read + y read + y written)
    The D^dag launch is reported over one pair of iterations (both parities
    timed), which is how its average over a solve is made up."""
    spin = dslash_bytes_per_site(args.inner, args.recon, 1) - dslash_bytes_per_site(args.inner, args.recon, 0)
    spin1 = spin // 2  # one field of the inner precision
    links = dslash_bytes_per_site(args.inner, args.recon, 0)
    acc = {"f16": "f16", "f32": "f32", "f64": "f64", "auto": "f32"}[args.accum]
    spin_y = (dslash_bytes_per_site(acc, 18, 1) - dslash_bytes_per_site(acc, 18, 0)) // 2
    rec = args.search == "recurrence"
    if rec:
        op_b = (links + k * 3 * spin1) * V
        dag_b0 = (links + k * 5 * spin1) * V
        dag_b1 = (links + k * (5 * spin1 + 2 * spin_y)) * V
    else:
        op_b = (links + k * 2 * spin1) * V
        dag_b0 = (links + k * 3 * spin1) * V
        dag_b1 = (links + k * (5 * spin1 + 2 * spin_y)) * V
    t_op = prof["us_op"]
    t_dag = prof["us_op_dag"]
    n = prof["iterations"]
    if min(n) == 0:
        return {"error": "no timed iterations", "profile": prof}
    # which parity added to y: the slower launch of the pair
    y_par = 1 if t_dag[1] >= t_dag[0] else 0
    pair_b = dag_b0 + dag_b1
    pair_us = t_dag[0] + t_dag[1]
    op_us = (t_op[0] * n[0] + t_op[1] * n[1]) / (n[0] + n[1])
    dag_gbs = pair_b / (pair_us * 1e-6) / 1e9
    op_gbs = op_b / (op_us * 1e-6) / 1e9
    return {
        "profiled_solve": summarize(stats),
        "timed_iterations": n,
        "dag_cg_step": {
            "kernel": f"D^dag {args.inner} r{args.recon}, {k} RHS, fused CG step "
                      f"({'p = r + beta p_prev, ' if rec else ''}y in {acc}, pairwise)",
            "us_mean": round(pair_us / 2, 2), "us_by_parity": [round(t, 2) for t in t_dag], "y_parity": y_par,
            "bytes_per_site_without_y": dag_b0 // V, "bytes_per_site_with_y": dag_b1 // V,
            "GB/s": round(dag_gbs, 1), "frac_hbm_peak": round(dag_gbs / PEAK_HBM_GBS, 4),
        },
        "d_norm": {
            "kernel": f"D {args.inner} r{args.recon}, {k} RHS, {'t = D r + beta t_prev, ' if rec else ''}|t|^2 partials",
            "us_mean": round(op_us, 2), "bytes_per_site": op_b // V,
            "GB/s": round(op_gbs, 1), "frac_hbm_peak": round(op_gbs / PEAK_HBM_GBS, 4),
        },
        "search": args.search,
        "xpay_us": round(prof["us_xpay"], 2) if not rec else None,
    }


def compact_record(args, d, L, k):
    """The one-line stdout record: the headline, the roofline of the launch
    that dominates the timed step (D^dag with the fused CG step), the
    standalone Dslash launch beside it, the streaming ceiling, the CPU
    baseline and one summary figure per other configuration."""
    cg = d["cg_step_kernels"]
    dag = cg.get("dag_cg_step", {})
    dn = cg.get("d_norm", {})
    dl = d["dslash_launch"]
    line = {
        "metric": METRIC,
        "value": d["value"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": d["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.inner}-storage/f32-arith inner, f64 reliable updates",
        "data": "synthetic (seeded PCG64 SU(3) links eps=%g, complex Gaussian sources)" % args.eps,
        "config": {
            "workload": "SU(3) Wilson D^dag D CG, %s, m=%g, bc=(1,1,1,-1), %d RHS per solve, %s inner (recon %d, "
                        "per-site FP32 scales, FP32 arithmetic), FP64 reliable updates delta=%g, FP64 true "
                        "residual %g" % ("x".join(map(str, L)), args.mass, k, args.inner, args.recon, args.delta,
                                         args.tol),
            "lattice": list(L), "nrhs": k, "search": args.search,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dag.get("kernel"),
            "achieved": dag.get("GB/s"),
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": dag.get("frac_hbm_peak"),
            "traffic": cg.get("traffic"),
            "kernel_us": dag.get("us_mean"),
            "bytes_per_site": [dag.get("bytes_per_site_without_y"), dag.get("bytes_per_site_with_y")],
        },
        "roofline_d_norm": {"achieved": dn.get("GB/s"), "frac": dn.get("frac_hbm_peak"), "kernel_us": dn.get("us_mean"),
                            "bytes_per_site": dn.get("bytes_per_site")},
        "roofline_dslash": {"kernel": "D %s r%d, %d RHS (store)" % (args.inner, args.recon, k), "achieved": dl["GB/s"],
                            "frac": dl["frac_hbm_peak"], "kernel_us": dl["us"], "traffic": dl["traffic"],
                            "algorithmic_bytes": dl["algorithmic_bytes"]},
        "device_copy_GB/s": d["device_copy"]["GB/s"],
    }
    if "device_stream" in d:
        line["device_stream"] = {kk: {"GB/s": v["GB/s"], "frac": v["frac_hbm_peak"], "variant": v["variant"]}
                                 for kk, v in d["device_stream"]["best"].items()}
    sv = d["solve"]
    line["fp64_single_rhs"] = {"solves_per_s": sv["fp64_single_rhs"]["solves_per_s"],
                               "iters": sv["fp64_single_rhs"]["iters_max"]}
    line["speedup_vs_fp64_single"] = sv["speedup_mixed_multi_vs_fp64_single"]
    line["true_resid_max"] = sv["mixed_multi_rhs"]["true_resid_max"]
    line["rel_error_vs_fp64"] = sv["mixed_multi_rhs"]["rel_error_vs_fp64_col0"]
    eo = d.get("even_odd", {}).get("mixed_multi_rhs")
    if eo:
        line["even_odd_mixed_12rhs_solves_per_s"] = eo["solves_per_s"]
    c3 = d.get("config3_single_rhs_delta0.1")
    if c3:
        line["config3"] = {kk: {"ms": round(v["s_per_solve"] * 1e3, 1), "iters": v["iters_max"],
                                "err_vs_fp64": float("%.2e" % v["rel_error_vs_fp64"])} for kk, v in c3.items()}
    c2 = d.get("fp64_16x16x16x16")
    if c2:
        line["config2_16^4"] = {"dslash_GB/s": c2["dslash_f64"]["GB/s"], "GFLOP/s": c2["dslash_f64"]["GFLOP/s"],
                                "solves_per_s": c2["fp64_solve"]["solves_per_s"], "cache_resident": True}
    c4 = d.get("multi_rhs_24x24x24x48")
    if c4:
        line["config4_24^3x48"] = {kk: v["solves_per_s"] for kk, v in c4.items()}
    pw = d.get("published_workloads_8x8x8x16")
    if pw:
        line["published_cpu_workloads_8^3x16"] = {k: {"ms": v["ms"], "reference_cpu_ms": v["reference_cpu_ms"]}
                                                  for k, v in pw["cases"].items()}
    srv = d.get("serving")
    if srv:
        cb = srv["continuous_batching"]
        line["serving"] = {"overload_200/s": {"rhs_per_s": cb["rhs_per_s"], "p50_ms": cb["latency_p50_ms"],
                                              "p99_ms": cb["latency_p99_ms"]},
                           "fp64_queue_rhs_per_s": srv["fp64_unbatched_queue"]["rhs_per_s"],
                           "sweep": [{"req/s": p["offered_req_per_s"], "p50_ms": p["latency_p50_ms"],
                                      "p99_ms": p["latency_p99_ms"], "rhs_per_s": p["rhs_per_s"]}
                                     for p in srv.get("sustainable_load_sweep", [])]}
    if "cpu_baseline" in d:
        cbl = d["cpu_baseline"]
        line["cpu_baseline"] = {kk: cbl[kk] for kk in ("value", "unit", "cores", "kind", "sample")}
    line["detail"] = args.detail_out
    return line


def config2(B, args, with_cpu):
    """16^4, eps=0.5, m=0.05, one complex Gaussian source: the plain FP64
    single-RHS path.  The FP64 Dslash (tuned launch shape) in GB/s of
    algorithmic bytes (960 B/site) and GFLOP/s (1320 flop/site), the FP64 CG
    solve rate, and the C/OpenMP oracle solving the same system on the host
    cores (the whole solve, iterated-residual rule at 1e-10).  The 16^4
    working set (FP64 links 38 MB) fits the 256 MB Infinity Cache, so its
    GB/s is cache-resident, not an HBM roofline figure."""
    torch, nat = B.torch, B.nat
    from jaxqft_amd import fields
    from jaxqft_amd.autotune import time_dslash, tune_march

    L = (16, 16, 16, 16)
    V = int(np.prod(L))
    bc = (1, 1, 1, -1)
    rng = fields.rng_for(fields.SEED_BASE + 1)
    U = fields.su3_hot_start(rng, 1, L, 0.5)
    b = fields.random_fermion(rng, (1, *L, 4, 3))
    with torch.cuda.stream(B.stream):
        ctx = nat.Context(L, 3, 1, 0.05, 1.0, bc)
        gauge = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
        bt = torch.from_numpy(b).cuda()
        fi, fo = nat.Field(ctx, nat.WL_F64, 1), nat.Field(ctx, nat.WL_F64, 1)
        fi.import_(bt)
        t1 = tune_march(ctx, gauge, nat.WL_F64, 18, fi, fo, reps=10, stream=B.stream)
        us = time_dslash(ctx, gauge, nat.WL_F64, 18, fi, fo, 1, args.dslash_reps, B.stream)
        gbs = 960 * V / (us * 1e-6) / 1e9
        bf, xf = nat.Field(ctx, nat.WL_F64, 1), nat.Field(ctx, nat.WL_F64, 1)
        bf.import_(bt)
        B.solve(ctx, gauge, bf, xf, 1, "f64", 18)
        t, st = B.timed_solves(3, ctx, gauge, bf, xf, 1, "f64", 18)
    res = {
        "lattice": list(L), "eps": 0.5, "mass": 0.05, "cache_resident": True,
        "dslash_f64": {"us": round(us, 2), "GB/s": round(gbs, 1), "frac_hbm_peak": round(gbs / PEAK_HBM_GBS, 4),
                       "GFLOP/s": round(FLOP_PER_SITE * V / (us * 1e-6) / 1e9, 1), "bytes_per_site": 960,
                       "march": t1["best"], "slab": t1["slab"], "block_shape": t1["block_shape"]},
        "fp64_solve": {**summarize(st), "s_per_solve": round(t, 5), "solves_per_s": round(1 / t, 2)},
    }
    log(f"16^4: dslash {us:.1f} us {gbs:.0f} GB/s, FP64 solve {t*1e3:.2f} ms {summarize(st)}")
    if with_cpu:
        from oracle import oracle_c

        threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
        oracle_c.cg(U[0], b[0], L, 3, 0.05, bc, 1e-10, 5)  # warm-up
        t0 = time.perf_counter()
        _, it = oracle_c.cg(U[0], b[0], L, 3, 0.05, bc, 1e-10, 20000)
        cs = time.perf_counter() - t0
        res["cpu_baseline"] = {"s_per_solve": round(cs, 3), "solves_per_s": round(1 / cs, 4), "iters": it,
                               "cores": threads, "kind": "port",
                               "sample": "the whole FP64 CG solve with the C/OpenMP oracle (iterated residual 1e-10)"}
        res["gpu_over_cpu"] = round(cs / t, 1)
        log(f"16^4 CPU oracle solve {cs:.2f} s ({it} iterations, {threads} threads)")
    return res


def propagator_chrono(B, args, U32, L32):
    """The point-source propagator driver (measurements.py) at 32^4 through the
    theory object, tuned as a user would: D^dag on the 12 unit point sources
    and one 12-column mixed solve.  Chronological guesses at 16^4 along a
    smooth gauge trajectory U_k = exp(k eps H) U (eps = 0.01), phi_k =
    D_k^dag eta: iterations and time of each solve from a cold start and from
    the MRE guess over the last four solutions (chrono.py)."""
    torch, nat = B.torch, B.nat
    from jaxqft_amd import fields
    from jaxqft_amd.autotune import tune
    from jaxqft_amd.chrono import chrono_predict_mre
    from jaxqft_amd.measurements import full_point_source_propagator
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    pol = dict(cg_tol=args.tol, cg_maxiter=20000, solver_precision=args.inner, solver_recon=args.recon,
               reliable_delta=args.delta, solver_accum=args.accum, use_graph=bool(args.graph), poll_every=args.poll)
    out = {}
    with torch.cuda.stream(B.stream):
        th = SU3WilsonNf2(lattice_shape=L32, mass=args.mass, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, **pol)
        rq = th.resident_gauge(torch.from_numpy(U32).cuda())
        ctx = th._context()
        probe = torch.from_numpy(fields.random_fermion(fields.rng_for(fields.SEED_BASE + 7), (12, *L32, 4, 3))).cuda()
        for prec, recon in ((args.inner, args.recon), ("f64", 18)):
            pc = nat.PREC[prec]
            fi, fo = nat.Field(ctx, pc, 12), nat.Field(ctx, pc, 12)
            fi.import_(probe)
            rq.gauge.prepare(pc, recon)
            tune(ctx, rq.gauge, pc, recon, fi, fo, nrhs=12, reps=3, stream=B.stream)
            del fi, fo
        del probe
        src = (3, 5, 7, 11)
        full_point_source_propagator(rq, th, src, as_numpy=False)  # graphs, workspaces
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prop, info = full_point_source_propagator(rq, th, src, as_numpy=False)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = th.last_solve_stats
        out["propagator_32x32x32x32"] = {
            "wall_s": round(wall, 4), "solve_s": round(info["inv_solve_total_sec_step"], 4), "columns": 12,
            "iters_max": max(c["iters"] for c in st), "true_resid_max": max(c["true_resid"] for c in st),
            "status": sorted({c["status"] for c in st}),
            "workload": "D^dag on the 12 unit point sources at (3,5,7,11), one 12-column mixed solve of D^dag D, "
                        "the propagator reshaped to (B, *L, 4, 3, 4, 3) on the device",
        }
        log(f"propagator 32^4: {wall*1e3:.1f} ms (solve {info['inv_solve_total_sec_step']*1e3:.1f} ms)")
        del prop, rq, th

        L = (16, 16, 16, 16)
        rng = fields.rng_for(fields.SEED_BASE + 8)
        U0 = torch.from_numpy(fields.su3_hot_start(rng, 1, L, 0.5)).cuda()
        a = torch.from_numpy(rng.standard_normal((1, 4, *L, 3, 3)) + 1j * rng.standard_normal((1, 4, *L, 3, 3))).cuda()
        H = 0.5 * (a - a.conj().transpose(-1, -2))
        H = H - (torch.diagonal(H, dim1=-2, dim2=-1).sum(-1) / 3.0)[..., None, None] * torch.eye(
            3, dtype=H.dtype, device="cuda")
        eta = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3))).cuda()
        th = SU3WilsonNf2(lattice_shape=L, mass=0.05, fermion_bc=(1, 1, 1, -1), dtype=np.complex128, **pol)
        xs, ys, cold, warm = [], [], [], []
        for kk in range(8):
            Uk = torch.linalg.matrix_exp((0.01 * kk) * H.reshape(-1, 3, 3)).reshape(H.shape) @ U0
            rq = th.resident_gauge(Uk)
            phi = th.apply_Ddag(rq, eta)
            th.solve_normal(rq, phi)  # warm-up of this gauge's encodings and graphs
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x = th.solve_normal(rq, phi)
            torch.cuda.synchronize()
            cold.append((th.last_solve_stats[0]["iters"], time.perf_counter() - t0))
            if xs:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                guess = chrono_predict_mre(xs, ys, eta, 4)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                x = th.solve_normal_with_guess(rq, phi, guess)
                torch.cuda.synchronize()
                warm.append((th.last_solve_stats[0]["iters"], time.perf_counter() - t1, t1 - t0))
            xs.append(x)
            ys.append(th.apply_D(rq, x))
            del rq
        out["chrono_16x16x16x16"] = {
            "solves": len(warm), "cold_iters_mean": round(float(np.mean([c[0] for c in cold[1:]])), 2),
            "mre_iters_mean": round(float(np.mean([w[0] for w in warm])), 2),
            "cold_ms_mean": round(1e3 * float(np.mean([c[1] for c in cold[1:]])), 3),
            "mre_solve_ms_mean": round(1e3 * float(np.mean([w[1] for w in warm])), 3),
            "mre_prediction_ms_mean": round(1e3 * float(np.mean([w[2] for w in warm])), 3),
            "workload": "U_k = exp(0.01 k H) U, phi_k = D_k^dag eta, mixed solves to 1e-10; MRE guess from the "
                        "last 4 (x, D x) pairs with target eta",
        }
        log(f"chrono 16^4: cold {out['chrono_16x16x16x16']['cold_iters_mean']} iters, "
            f"MRE {out['chrono_16x16x16x16']['mre_iters_mean']} iters")
    return out


def config4(B, args):
    """24^3x48, eps=0.6, m=0.03: k in {1, 4, 12, 24} right-hand sides (k=12:
    D^dag applied to the 12 unit point sources at a random site, as the
    propagator loop of measurements.py:333-345 does; otherwise complex
    Gaussian), FP32 inner stencil, FP64 reliable updates.  Before each k the
    autotuner picks the FP32 inner and FP64 update launch shapes for k
    right-hand sides on this lattice."""
    torch, nat = B.torch, B.nat
    from jaxqft_amd import fields
    from jaxqft_amd.autotune import tune

    L = (24, 24, 24, 48)
    rng = fields.rng_for(fields.SEED_BASE + 3)
    U = fields.su3_hot_start(rng, 1, L, 0.6)
    res = {}
    with torch.cuda.stream(B.stream):
        ctx = nat.Context(L, 3, 1, 0.03, 1.0, (1, 1, 1, -1))
        gauge = nat.Gauge(ctx, torch.from_numpy(U).cuda(), nat.WL_LAYOUT_BMXYIJ)
        for k in (1, 4, 12, 24):
            if k == 12:
                site = tuple(int(v) for v in rng.integers(0, L))
                src = torch.from_numpy(fields.point_sources(L, site)).cuda()
                bt = torch.empty_like(src)
                nat.check(B.lib.wl_apply(ctx.h, gauge.h, nat.WL_OP_DDAG, c_void_p(src.data_ptr()),
                                         c_void_p(bt.data_ptr()), 12, nat.WL_C128, B.h))
                kind = "D^dag point sources"
            else:
                bt = torch.from_numpy(fields.random_fermion(rng, (k, *L, 4, 3))).cuda()
                kind = "Gaussian"
            b = nat.Field(ctx, nat.WL_F64, k)
            b.import_(bt)
            x = nat.Field(ctx, nat.WL_F64, k)
            tuned = {}
            for prec in ("f32", "f64"):
                pc = nat.PREC[prec]
                fi, fo = nat.Field(ctx, pc, k), nat.Field(ctx, pc, k)
                fi.import_(bt)
                tk = tune(ctx, gauge, pc, 12, fi, fo, nrhs=k, reps=3, stream=B.stream)
                tuned[prec] = {"rhs_per_thread": tk["rhs_per_thread"], "march": tk["march"], "slab": tk["slab"],
                               "cols_outer": tk["cols_outer"], "block_shape": tk["block_shape"]}
                del fi, fo
            B.solve(ctx, gauge, b, x, k, "f32", 12)
            t, st = B.timed_solves(1, ctx, gauge, b, x, k, "f32", 12)
            res[str(k)] = {**summarize(st), "sources": kind, "s_per_batch": round(t, 4),
                           "solves_per_s": round(k / t, 3), "tuning": tuned}
            log(f"24^3x48 k={k}: {t*1e3:.1f} ms, {k/t:.2f} solves/s {summarize(st)}")
            del b, x, bt
    return res


# The reference's published timings (HANDOFF.md:2161-2179: CPU, 20 threads, JAX
# jit, complex64; its `--tests perf` / `eoperf` / `eopforce` runs at 8^3x16,
# mass 0.05, weak field hot_start(0.05), su3_wilson_nf2.py:1719-1721,
# :2602-2650), in ms per call.
PUBLISHED_CPU_MS = {"D": 0.85, "DdagD": 1.64, "cg_243_iters_tol0": 1450.0, "eo_cg_86_iters_tol0": 400.0,
                    "eo_analytic_force_tol1e-7": 68.2}


def published_workloads(B, args):
    """The workloads behind the reference's published CPU timings, run through
    the drop-in (SU3WilsonNf2, complex64 tensors resident on the device, as the
    reference's default dtype): D and D^dag D per call, CG on D^dag D for a
    fixed 243 iterations and even-odd Schur CG for a fixed 86 (tol 0, FP32
    inner with FP64 reliable updates, the arithmetic of a complex64 solve kept
    honest by the FP64 updates), and the even-odd monomial's analytic force
    with CG tol 1e-7.  Each figure is the median wall time per call over
    repeated calls, host overhead included, beside the published CPU figure."""
    torch = B.torch
    from jaxqft_amd import fields
    from jaxqft_amd.su3_wilson_nf2 import SU3WilsonNf2

    L = (8, 8, 8, 16)
    rng = fields.rng_for(fields.SEED_BASE + 9)
    U = torch.from_numpy(fields.su3_hot_start(rng, 1, L, 0.05).astype(np.complex64)).cuda()
    psi = torch.from_numpy(fields.random_fermion(rng, (1, *L, 4, 3)).astype(np.complex64)).cuda()
    base = dict(lattice_shape=L, mass=0.05, fermion_bc=(1, 1, 1, -1), dtype=np.complex64, solver_precision="f32",
                solver_recon=12, use_graph=bool(args.graph), poll_every=args.poll)

    def med(fn, n):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    cases = {}
    with torch.cuda.stream(B.stream):
        th = SU3WilsonNf2(cg_tol=0.0, cg_maxiter=243, **base)
        G = th.resident_gauge(U)
        cases["D"] = {"ms": med(lambda: th.apply_D(G, psi), 50)}
        cases["DdagD"] = {"ms": med(lambda: th.apply_normal(G, psi), 50)}
        cases["cg_243_iters_tol0"] = {"ms": med(lambda: th.solve_normal(G, psi), 5),
                                      "iters": th.last_solve_stats[0]["iters"]}
        the = SU3WilsonNf2(cg_tol=0.0, cg_maxiter=86, **base)
        Ge = the.resident_gauge(U)
        rhs_e = the.project_even(psi)
        cases["eo_cg_86_iters_tol0"] = {"ms": med(lambda: the.solve_eop_even_normal(Ge, rhs_e), 5),
                                        "iters": the.last_solve_stats[0]["iters"]}
        thf = SU3WilsonNf2(cg_tol=1e-7, cg_maxiter=1000, **base)
        Gf = thf.resident_gauge(U)
        phi_e = thf.apply_eo_schur_even_dagger(Gf, rhs_e, normalized=True)
        cases["eo_analytic_force_tol1e-7"] = {"ms": med(lambda: thf.eop_pseudofermion_force(Gf, phi_e), 5),
                                              "iters": thf.last_solve_stats[0]["iters"]}
    for k, v in cases.items():
        v["ms"] = round(v["ms"], 4)
        v["reference_cpu_ms"] = PUBLISHED_CPU_MS[k]
        v["ratio"] = round(PUBLISHED_CPU_MS[k] / v["ms"], 1)
        log(f"8^3x16 {k}: {v['ms']:.4f} ms per call, published CPU {PUBLISHED_CPU_MS[k]} ms ({v['ratio']}x)")
    return {"lattice": list(L), "links": "hot_start(0.05), complex64", "mass": 0.05, "cases": cases,
            "source": "the reference's published CPU timings (20 threads, JAX jit, complex64) for these workloads"}


def serving(B, args):
    """Poisson requests (mean 200/s) over {8^4, 16^4, 24^3x48, 32^4}, 4
    device-resident fields per size (eps = 0.5), m ~ U[0.02, 0.2], k ~ U{1..12}
    right-hand sides; continuous batching with FP16 inner / FP64 reliable
    updates against an unbatched FP64 queue on the same requests."""
    torch = B.torch
    from jaxqft_amd import fields
    from jaxqft_amd.serving import Request, Server, fp64_queue, poisson_requests

    lattices = [(8, 8, 8, 8), (16, 16, 16, 16), (24, 24, 24, 48), (32, 32, 32, 32)]
    t0 = time.perf_counter()
    U = {}
    for i, L in enumerate(lattices):
        for f in range(4):
            U[(L, f)] = fields.su3_hot_start_device(fields.SEED_BASE + 100 + 10 * i + f, 1, L, 0.5)
    torch.cuda.synchronize()
    log(f"serving: generated {len(U)} gauge fields in {time.perf_counter() - t0:.1f}s")
    free, _ = torch.cuda.mem_get_info()
    budget = 0.6 * free
    reqs = poisson_requests(200.0, args.serving_requests, lattices, 4, seed=fields.SEED_BASE + 5)
    srv = Server(U, capacity=12, inner="f16", recon=12, tol=args.tol, delta=args.delta, poll_every=8,
                 hbm_budget_bytes=budget)
    del U
    # warm the lanes (graph capture, gauge encodings) on a short burst
    warm = poisson_requests(1000.0, 8, lattices, 4, seed=1)
    srv.run(warm, realtime=False)
    res = srv.run(reqs, realtime=True)
    log(f"serving continuous batching: {res}")
    base = [Request(rid=r.rid, key=r.key, mass=r.mass, rhs=r.rhs, arrival=r.arrival) for r in reqs]
    res64 = fp64_queue(base, srv.backend.gauges, srv.backend.ctxs, tol=args.tol, realtime=True)
    log(f"serving unbatched fp64 queue: {res64}")
    diff = max(float(torch.linalg.norm(a.x - b.x) / torch.linalg.norm(b.x)) for a, b in zip(reqs, base))
    del reqs, base
    torch.cuda.synchronize()
    # sustainable loads: Poisson arrivals at fractions of the request rate
    # served at overload, so that p50 / p99 are service latencies and not
    # queue growth; right-hand sides are made at admission and dropped when
    # a request finishes (stats kept)
    sweep = []
    if args.serving_sweep_requests > 0:
        cap = res["requests_per_s"]
        for frac in (float(v) for v in args.serving_sweep_loads.split(",")):
            rate = frac * cap
            rq = poisson_requests(rate, args.serving_sweep_requests, lattices, 4,
                                  seed=fields.SEED_BASE + 50 + int(100 * frac), lazy=True)
            out = srv.run(rq, realtime=True, release=True)
            out["offered_req_per_s"] = round(rate, 3)
            out["offered_fraction_of_overload_rate"] = frac
            sweep.append(out)
            log(f"serving at {rate:.2f} req/s ({frac} of the overload rate): {out}")
            del rq
    return {
        "sustainable_load_sweep": sweep,
        "workload": f"{args.serving_requests} Poisson requests at 200/s (overload: arrivals outpace service, "
                    "so latencies include queue growth) over 8^4,16^4,24^3x48,32^4 (4 fields each, eps=0.5), "
                    "m~U[0.02,0.2], k~U{1..12}",
        "hbm_budget_GB": round(budget / 1e9, 1),
        "continuous_batching": res,
        "fp64_unbatched_queue": res64,
        "throughput_ratio": round(res["rhs_per_s"] / res64["rhs_per_s"], 3),
        "max_rel_diff_vs_fp64_queue": diff,
    }


def lime_decode(B, args, L, with_cpu):
    """SciDAC gauge payload decode at the bench lattice: precision D, four
    links per site, sites lexicographic and big-endian, decoded by
    wl_lime_decode into BM...IJ complex128 (byte swap + axis reversal + the
    direction axis moved first).  Algorithmic bytes: the payload read once
    and the links written once, 576 + 576 B per site.  Also times the
    payload staged from pinned host memory, and the NumPy oracle decode of
    the same payload on one host core."""
    import torch

    nat = B.nat
    V = int(np.prod(L))
    stream = B.stream
    with torch.cuda.stream(stream):
        g = torch.Generator(device="cuda").manual_seed(11)
        U = torch.randn((4, *L, 3, 3, 2), generator=g, device="cuda", dtype=torch.float64)
        pay = U.permute(4, 3, 2, 1, 0, 5, 6, 7).contiguous().view(torch.uint8).view(-1, 8).flip(-1).reshape(-1)
        pay = pay.contiguous()
        Uc = torch.view_as_complex(U).reshape(4, *L, 9)
        out = torch.empty_like(Uc)
        cd = (c_int * 4)(*L)

        def run(src):
            nat.check(B.lib.wl_lime_decode(c_void_p(src.data_ptr()), src.numel(), 8, cd, 4, 4, 9, nat.WL_LIME_COUNT,
                                           c_void_p(out.data_ptr()), nat.WL_C128, B.h))

        run(pay)
        stream.synchronize()
        exact = bool(torch.equal(out, Uc))
        nrep = max(5, args.dslash_reps // 2)
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(nrep):
            run(pay)
        ev1.record(stream)
        ev1.synchronize()
        us = ev0.elapsed_time(ev1) * 1e3 / nrep
        nb = pay.numel()
        host = torch.empty(nb, dtype=torch.uint8, pin_memory=True)
        host.copy_(pay)
        dev = torch.empty_like(pay)
        ev0.record(stream)
        for _ in range(3):
            dev.copy_(host, non_blocking=True)
            run(dev)
        ev1.record(stream)
        ev1.synchronize()
        staged_ms = ev0.elapsed_time(ev1) / 3
        exact = exact and bool(torch.equal(out, Uc))
    res = {
        "payload_bytes": nb, "kernel_us": round(us, 2), "launches": nrep,
        "GB/s": round(2 * nb / (us * 1e-6) / 1e9, 1), "frac_hbm_peak": round(2 * nb / (us * 1e-6) / 1e9 / PEAK_HBM_GBS, 4),
        "bytes_per_site": 2 * 576, "bit_exact": exact,
        "pinned_host_to_links_ms": round(staged_ms, 3),
        "pinned_host_to_links_GB/s": round(nb / (staged_ms * 1e-3) / 1e9, 2),
    }
    log(f"lime decode: {us:.1f} us {res['GB/s']} GB/s, from pinned host {staged_ms:.2f} ms, exact {exact}")
    if with_cpu:
        from oracle import lime_oracle as lo

        raw = host.numpy()
        t0 = time.perf_counter()
        lat = lo.decode_payload(raw, "D", 144, 4, L, "QDP_D3_ColorMatrix", 3, 4)["lattice_data"]
        Uh = np.ascontiguousarray(lo.gauge_from_lattice(lat, 4))
        cpu_s = time.perf_counter() - t0
        res["cpu_oracle"] = {"s": round(cpu_s, 3), "GB/s": round(nb / cpu_s / 1e9, 3), "cores": 1,
                             "sample": "the whole payload decoded to a contiguous BM...IJ array by the NumPy oracle",
                             "matches_gpu": bool(np.array_equal(Uh[0].reshape(4, *L, 9), out.cpu().numpy()))}
        log(f"lime decode cpu oracle: {cpu_s:.2f} s")
        del lat, Uh
    del U, pay, Uc, out, host, dev
    torch.cuda.synchronize()
    return res


def device_stream(B, mib: int = 1024, reps: int = 10):
    """Hand-written HBM streaming kernels (wl_stream_probe), 1 GiB per
    buffer: read-only, write-only and copy, 16 B per lane, plain or
    non-temporal accesses, 1, 2 or 4 accesses in flight per lane, one pass
    over the buffer or a 2048 / 8192-workgroup grid-stride loop.  HIP events
    on the bench stream around `reps` launches; bytes counted are those read
    (read), written (write) or read + written (copy).  The fastest variant of
    each kind is the measured ceiling the stencils are read against."""
    import torch

    nat = B.nat
    stream = B.stream
    nbytes = mib << 20
    res = {"bytes_per_buffer": nbytes, "variants": {}, "best": {}}
    with torch.cuda.stream(stream):
        src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
        src.fill_(1.0)
        dst = torch.empty_like(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for kind, name, moved in ((nat.WL_STREAM_READ, "read", nbytes), (nat.WL_STREAM_WRITE, "write", nbytes),
                                  (nat.WL_STREAM_COPY, "copy", 2 * nbytes)):
            for ntm in (0, 1):
                for unroll in (1, 2, 4):
                    for blocks in (0, 2048, 8192):
                        def run():
                            nat.check(B.lib.wl_stream_probe(kind, ntm, unroll, c_void_p(src.data_ptr()),
                                                            c_void_p(dst.data_ptr()), nbytes, blocks, B.h))
                        run()
                        run()
                        e0.record(stream)
                        for _ in range(reps):
                            run()
                        e1.record(stream)
                        e1.synchronize()
                        us = e0.elapsed_time(e1) * 1e3 / reps
                        gbs = moved / (us * 1e-6) / 1e9
                        tag = f"{name}/{'nt' if ntm else 'plain'}/u{unroll}/{'pass' if blocks == 0 else f'g{blocks}'}"
                        res["variants"][tag] = round(gbs, 1)
                        if name not in res["best"] or gbs > res["best"][name]["GB/s"]:
                            res["best"][name] = {"GB/s": round(gbs, 1), "frac_hbm_peak": round(gbs / PEAK_HBM_GBS, 4),
                                                 "us": round(us, 2), "variant": tag}
        del src, dst
    torch.cuda.synchronize()
    log("device stream: " + ", ".join(f"{k} {v['GB/s']} GB/s ({v['variant']})" for k, v in res["best"].items()))
    return res


def device_copy_rate(stream, mib: int = 1024, reps: int = 10):
    """Plain device-to-device copy of `mib` MiB on `stream` (HIP events):
    read + write bytes per second, the practical HBM rate the stencils are
    compared with."""
    import torch

    with torch.cuda.stream(stream):
        src = torch.empty(mib << 17, dtype=torch.float64, device="cuda")
        src.fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            dst.copy_(src)
        e1.record(stream)
        e1.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    nbytes = 2 * src.numel() * 8
    gbs = nbytes / (us * 1e-6) / 1e9
    return {"GB/s": round(gbs, 1), "frac_hbm_peak": round(gbs / PEAK_HBM_GBS, 4), "us": round(us, 2),
            "bytes_read_plus_written": nbytes}


def cpu_baseline(U, b, L, args, gpu_iters):
    """The C/OpenMP oracle (the reference algorithm restated) timed on this
    host's cores: FP64 CG iterations on D^dag D on the same fields for about
    args.cpu_seconds, extrapolated to the GPU FP64 iteration count."""
    from oracle import oracle_c

    threads = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    Ub, bb = U[0], b[0]
    t0 = time.perf_counter()
    oracle_c.cg(Ub, bb, L, 3, args.mass, (1, 1, 1, -1), 0.0, 1)
    one = time.perf_counter() - t0
    n = max(1, int(args.cpu_seconds / max(one, 1e-3)))
    t0 = time.perf_counter()
    _, it = oracle_c.cg(Ub, bb, L, 3, args.mass, (1, 1, 1, -1), 0.0, n)
    s_per_iter = (time.perf_counter() - t0) / it
    return {
        "value": round(1.0 / (s_per_iter * max(1, gpu_iters)), 6),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{it} FP64 CG iterations on D^dag D at {'x'.join(map(str, L))} with the C/OpenMP oracle "
                  f"({threads} threads): {s_per_iter*1e3:.1f} ms/iteration, extrapolated to the GPU FP64 "
                  f"single-RHS iteration count {gpu_iters}",
    }


if __name__ == "__main__":
    main()
