"""Point-source quark propagator through one multi-RHS solve.

Mirrors ``_full_point_source_propagator`` (core/measurements.py:277-371) for
theories built on this library with ``solver_kind="cg"``: the reference makes
Ns * Nc sequential calls ``solve_normal(U, D^dag e_{s,c})`` on the unit point
sources at ``source`` (:333-345); here the Ns * Nc right-hand sides are one
batched D^dag application and one multi-RHS solve (the columns share each link
load and converge independently).  Returns ``(prop, info)`` with
``prop[b, *x, s, c, s0, c0]`` the solution for the source (s0, c0), as in the
reference, and the reference's timing keys in ``info``.
"""

from __future__ import annotations

import time
from typing import Dict, Sequence, Tuple

import numpy as np
import torch

from . import _native as nat
from .theory import ResidentGauge


def _coerce_source(source: Sequence[int], lattice_shape: Sequence[int]) -> Tuple[int, ...]:
    """measurements.py:_coerce_source: one coordinate per axis, wrapped."""
    src = tuple(int(v) for v in source)
    if len(src) != len(lattice_shape):
        raise ValueError(f"source must have {len(lattice_shape)} coordinates, got {len(src)}")
    return tuple(int(s) % int(n) for s, n in zip(src, lattice_shape))


def full_point_source_propagator(q, theory, source: Sequence[int], *, backend: str = "iterative",
                                 as_numpy: bool = True) -> Tuple[object, Dict[str, float]]:
    key = str(backend).strip().lower()
    if key in ("iter", "cg", "point"):
        key = "iterative"
    if key != "iterative":
        raise ValueError(f"Unsupported propagator backend for the native path: {backend!r} (use 'iterative')")
    if str(getattr(theory, "solver_kind", "cg")).lower() != "cg":
        raise ValueError("the native propagator path solves the normal equations (solver_kind='cg')")
    L = tuple(int(v) for v in theory.lattice_shape)
    src = _coerce_source(source, L)
    bs, ns, nc = int(theory.Bs), 4, int(theory.Nc)
    k = ns * nc
    t0 = time.perf_counter()
    rhs = torch.zeros((bs, k, *L, ns, nc), dtype=torch.complex128, device="cuda")
    for s in range(ns):
        for c in range(nc):
            rhs[(slice(None), s * nc + c, *src, s, c)] = 1.0
    rq = q if isinstance(q, ResidentGauge) else theory.resident_gauge(q)  # upload the links once
    ctx, g = theory._gauge(rq)
    flat = rhs.reshape((bs * k, *L, ns, nc))
    rhs_n = torch.empty_like(flat)
    nat.check(ctx._lib.wl_apply(ctx.h, g.h, nat.WL_OP_DDAG, nat.c_void_p(flat.data_ptr()),
                                nat.c_void_p(rhs_n.data_ptr()), k, nat.WL_C128, nat.c_void_p(nat.current_stream())))
    ts = time.perf_counter()
    sol = theory.solve_normal_multi(rq, rhs_n.reshape((bs, k, *L, ns, nc)))
    torch.cuda.synchronize()
    dt = time.perf_counter() - ts
    sol_t = sol if isinstance(sol, torch.Tensor) else torch.as_tensor(np.asarray(sol))
    # (b, s0*nc + c0, *x, s, c) -> (b, *x, s, c, s0, c0)
    nd = len(L)
    prop = sol_t.reshape((bs, ns, nc, *L, ns, nc)).permute(0, *range(3, 3 + nd), 3 + nd, 4 + nd, 1, 2).contiguous()
    build = time.perf_counter() - t0
    per = dt / k
    info = {
        "inv_backend_dense": 0.0,
        "inv_n_solves": float(k),
        "inv_solve_total_sec_step": float(dt),
        "inv_solve_mean_sec": float(per),
        "inv_solve_min_sec": float(per),
        "inv_solve_max_sec": float(per),
        "inv_prop_build_wall_sec": float(build),
        "inv_used_normal_equations": 1.0,
        "inv_cache_hit": 0.0,
        "inv_solve_total_sec_this_call": float(dt),
        "inv_multi_rhs_columns": float(k),
    }
    return (prop.cpu().numpy() if as_numpy else prop), info
