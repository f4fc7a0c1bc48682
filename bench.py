"""Benchmark: CG solves/s and Dslash HBM GB/s on a 32^4 SU(3) Wilson lattice,
mixed-precision reliable-update CG vs plain FP64, true residual 1e-10.

One step = one solve of D^dag D x = b on the resident 32^4 gauge field
(eps = 0.6, m = 0.02, boundary phases (1,1,1,-1), complex Gaussian source),
inputs already in HBM.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from ctypes import byref, c_void_p
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "CG solves/s + Dslash HBM GB/s, 32^4 SU(3) Wilson; mixed vs FP64 at 1e-10 res"
PEAK_HBM_GBS = 8000.0
FLOP_PER_SITE = 1320

# compulsory HBM bytes per site of one Dslash application (links read once,
# input spinor read once, output written once)
def dslash_bytes_per_site(prec: str, recon: int, nc: int = 3) -> int:
    spin_reals = 24 if nc == 3 else 8
    link_reals = (12 if recon == 12 else 18) if nc == 3 else 2
    if prec == "f64":
        return (4 * link_reals + 2 * spin_reals) * 8
    if prec == "f32":
        return (4 * link_reals + 2 * spin_reals) * 4
    return (4 * link_reals + 2 * spin_reals) * 2 + 2 * 4  # + per-site scale in and out


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--lattice", type=str, default="32,32,32,32")
    ap.add_argument("--eps", type=float, default=0.6)
    ap.add_argument("--mass", type=float, default=0.02)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--inner", type=str, default="f16", help="inner precision of the mixed solve")
    ap.add_argument("--recon", type=int, default=12)
    ap.add_argument("--delta", type=float, default=0.1)
    ap.add_argument("--poll", type=int, default=16)
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--dslash-reps", type=int, default=50)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-only", action="store_true",
                    help="run only the Dslash variants and a calibration copy (for rocprofv3 counter passes)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is a design "
              "(see the domain-decomposition section of DESIGN.md) and is not run.", file=sys.stderr)
        sys.exit(2)
    import torch

    from jaxqft_amd import _native as nat
    from jaxqft_amd import fields

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    L = tuple(int(v) for v in args.lattice.split(","))
    V = int(np.prod(L))
    bc = (1, 1, 1, -1)
    stream = torch.cuda.Stream()  # non-default stream (graph capture needs one)
    sp_handle = stream.cuda_stream

    t0 = time.perf_counter()
    rng = fields.rng_for(fields.SEED_BASE + 2)
    U = fields.su3_hot_start(rng, 1, L, args.eps)
    b_host = fields.random_fermion(rng, (1, *L, 4, 3))
    t_gen = time.perf_counter() - t0
    print(f"[bench] generated fields {L} in {t_gen:.1f}s", file=sys.stderr, flush=True)

    lib = nat.load()
    with torch.cuda.stream(stream):
        ctx = nat.Context(L, 3, 1, args.mass, 1.0, bc)
        Ut = torch.from_numpy(U).cuda()
        gauge = nat.Gauge(ctx, Ut, nat.WL_LAYOUT_BMXYIJ)
        del Ut
        bt = torch.from_numpy(b_host).cuda()
        b = nat.Field(ctx, nat.WL_F64, 1)
        b.import_(bt)
        x = nat.Field(ctx, nat.WL_F64, 1)
        x64 = nat.Field(ctx, nat.WL_F64, 1)
        torch.cuda.synchronize()

        # ---------------------------------------------------------- Dslash
        # autotune the march length per (precision, recon), then time the best
        from jaxqft_amd.autotune import time_dslash, tune_march

        dslash = {}
        variants = [("f64", 18), ("f64", 12), ("f32", 18), ("f32", 12), ("f16", 18), ("f16", 12)]
        for prec, recon in variants:
            pc = nat.PREC[prec]
            fi = nat.Field(ctx, pc, 1)
            fo = nat.Field(ctx, pc, 1)
            fi.import_(bt)
            tune = tune_march(ctx, gauge, pc, recon, fi, fo, reps=max(3, args.dslash_reps // 5), stream=stream)
            us = time_dslash(ctx, gauge, pc, recon, fi, fo, 1, args.dslash_reps, stream)
            bps = dslash_bytes_per_site(prec, recon)
            gbs = bps * V / (us * 1e-6) / 1e9
            dslash[f"{prec}_r{recon}"] = {
                "us": round(us, 2), "GB/s": round(gbs, 1), "frac_hbm_peak": round(gbs / PEAK_HBM_GBS, 4),
                "GFLOP/s": round(FLOP_PER_SITE * V / (us * 1e-6) / 1e9, 1), "bytes_per_site": bps,
                "march": tune["best"], "tune_us": {str(k): round(v, 2) for k, v in tune["us"].items()},
            }
            print(f"[bench] dslash {prec} r{recon}: {us:.1f} us  {gbs:.0f} GB/s  march {tune}", file=sys.stderr,
                  flush=True)
            del fi, fo

        if args.profile_only:
            # calibration: a plain device copy of known size (256 MiB read + 256 MiB written)
            src = torch.empty(32 << 20, dtype=torch.float64, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(5):
                dst.copy_(src)
            torch.cuda.synchronize()
            print(json.dumps({"profile_only": True, "dslash": dslash, "copy_bytes_each_way": src.numel() * 8}))
            return

        # ---------------------------------------------------------- solves
        def solve(field_x, prec, recon, graph):
            sp, keep = nat.solve_params(args.tol, 20000, nat.WL_STOP_TRUE_RESIDUAL, nat.PREC[prec], recon, args.delta,
                                        args.poll, bool(graph))
            st = (nat.WlSolveStats * 1)()
            nat.check(lib.wl_cg_normal_field(ctx.h, gauge.h, b.h, field_x.h, 0, 1, byref(sp), st, c_void_p(sp_handle)))
            return nat.stats_list(st, 1)[0]

        # plain FP64 reference solve
        st64 = solve(x64, "f64", 18, args.graph)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            st64 = solve(x64, "f64", 18, args.graph)
        torch.cuda.synchronize()
        fp64_s = (time.perf_counter() - t0) / args.fp64_steps
        print(f"[bench] fp64 solve {fp64_s:.3f}s {st64}", file=sys.stderr, flush=True)

        for _ in range(args.warmup):
            stm = solve(x, args.inner, args.recon, args.graph)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            stm = solve(x, args.inner, args.recon, args.graph)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        ms_per_step = elapsed / args.steps * 1e3
        print(f"[bench] mixed solve {ms_per_step:.1f} ms {stm}", file=sys.stderr, flush=True)

        xa = torch.empty_like(bt)
        xb = torch.empty_like(bt)
        x.export(xa)
        x64.export(xb)
        torch.cuda.synchronize()
        err_vs_fp64 = float(torch.linalg.norm(xa - xb) / torch.linalg.norm(xb))

    inner_key = f"{args.inner}_r{args.recon}"
    dom = dslash[inner_key]
    out = {
        "metric": METRIC,
        "value": round(1e3 / ms_per_step, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.inner}-inner/f64-reliable",
        "data": "synthetic (seeded PCG64 SU(3) links eps=%g, complex Gaussian source)" % args.eps,
        "config": {"workload": "SU(3) Wilson D^dag D CG, %s, m=%g, bc=(1,1,1,-1), true residual %g" % (
            "x".join(map(str, L)), args.mass, args.tol), "lattice": list(L), "nrhs": 1,
                   "inner": args.inner, "recon": args.recon, "delta": args.delta, "graph": bool(args.graph)},
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_dslash {inner_key}",
            "achieved": dom["GB/s"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": dom["frac_hbm_peak"],
            "traffic": None,
        },
        "solve": {
            "mixed": {**stm, "s_per_solve": round(ms_per_step / 1e3, 4)},
            "fp64": {**st64, "s_per_solve": round(fp64_s, 4), "solves_per_s": round(1.0 / fp64_s, 4)},
            "speedup_mixed_vs_fp64": round(fp64_s / (ms_per_step / 1e3), 3),
            "rel_error_mixed_vs_fp64": err_vs_fp64,
        },
        "dslash": dslash,
    }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(U, b_host, L, args, st64["iters"])
    print(json.dumps(out))


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
    _, k = oracle_c.cg(Ub, bb, L, 3, args.mass, (1, 1, 1, -1), 0.0, n)
    s_per_iter = (time.perf_counter() - t0) / k
    return {
        "value": round(1.0 / (s_per_iter * max(1, gpu_iters)), 6),
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{k} FP64 CG iterations on D^dag D at {'x'.join(map(str, L))} with the C/OpenMP oracle "
                  f"({threads} threads): {s_per_iter*1e3:.1f} ms/iteration, extrapolated to the GPU FP64 "
                  f"iteration count {gpu_iters}",
    }


if __name__ == "__main__":
    main()
