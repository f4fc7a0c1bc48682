"""ctypes binding of libwilson_amd.so (the C ABI declared in include/wilson_amd.h).

The library must be built in-tree (``__graft_entry__.build()`` or
``python -m jaxqft_amd.build_native``).  There is no CPU fallback: every
operator call runs the HIP kernels, and a missing library or GPU raises.
"""

from __future__ import annotations

import ctypes
import re
from ctypes import POINTER, byref, c_double, c_int, c_size_t, c_void_p
from pathlib import Path
from typing import List, Optional

import torch  # noqa: F401  (loads torch's HIP runtime first, so the library binds to the same one)

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libwilson_amd.so"
HEADER_PATH = Path(__file__).resolve().parent.parent / "include" / "wilson_amd.h"

WL_OK, WL_ERR_BAD_ARG, WL_ERR_HIP, WL_ERR_OOM, WL_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
WL_F64, WL_F32, WL_F16 = 0, 1, 2
WL_C128, WL_C64 = 0, 1
WL_LAYOUT_BMXYIJ, WL_LAYOUT_BXYMIJ = 0, 1
WL_OP_D, WL_OP_DDAG, WL_OP_G5D, WL_OP_K, WL_OP_KDAG, WL_OP_DDAGD = 0, 1, 2, 3, 4, 5
(WL_EO_K_OE, WL_EO_K_EO, WL_EO_KDAG_OE, WL_EO_KDAG_EO, WL_EO_SCHUR_E, WL_EO_SCHUR_E_DAG, WL_EO_SCHUR_O,
 WL_EO_SCHUR_O_DAG) = range(8)
WL_STOP_TRUE_RESIDUAL, WL_STOP_ITERATED = 0, 1
WL_LIME_LEXI, WL_LIME_SITE, WL_LIME_COUNT = 0, 1, 2
WL_ACCUM_AUTO, WL_ACCUM_F64, WL_ACCUM_F32, WL_ACCUM_F16 = 0, 1, 2, 3
ACCUM = {None: WL_ACCUM_AUTO, "auto": WL_ACCUM_AUTO, "f64": WL_ACCUM_F64, "f32": WL_ACCUM_F32, "f16": WL_ACCUM_F16}
WL_SEARCH_RECURRENCE, WL_SEARCH_XPAY = 0, 1
SEARCH = {"recurrence": WL_SEARCH_RECURRENCE, "xpay": WL_SEARCH_XPAY}
STATUS_NAMES = {-1: "running", 0: "converged", 1: "maxiter", 2: "nan", 3: "breakdown"}
PREC = {"f64": WL_F64, "f32": WL_F32, "f16": WL_F16}


class WlGeom(ctypes.Structure):
    _fields_ = [("nd", c_int), ("L", c_int * 4), ("nc", c_int), ("batch", c_int)]


class WlParams(ctypes.Structure):
    _fields_ = [("mass", c_double), ("wilson_r", c_double), ("bc_re", c_double * 4), ("bc_im", c_double * 4)]


class WlSolveParams(ctypes.Structure):
    _fields_ = [
        ("tol", c_double),
        ("maxiter", c_int),
        ("stop_rule", c_int),
        ("inner_prec", c_int),
        ("inner_recon", c_int),
        ("delta", c_double),
        ("poll_every", c_int),
        ("use_graph", c_int),
        ("mass", POINTER(c_double)),
        ("accum_prec", c_int),
        ("search_update", c_int),
    ]


class WlSolveStats(ctypes.Structure):
    _fields_ = [
        ("iters", c_int),
        ("status", c_int),
        ("reliable_updates", c_int),
        ("pad", c_int),
        ("true_resid", c_double),
        ("iter_resid", c_double),
    ]


class WlTuning(ctypes.Structure):
    _fields_ = [("march", c_int), ("rhs_per_thread", c_int), ("xcd_remap", c_int), ("slab", c_int),
                ("cols_outer", c_int), ("block_shape", c_int), ("blocks_per_cu", c_int), ("plane", c_int),
                ("eo_slab", c_int), ("lds_box", c_int), ("slab_x2", c_int)]


class WlProfile(ctypes.Structure):
    _fields_ = [("iterations", c_int * 2), ("us_op", c_double * 2), ("us_op_dag", c_double * 2), ("us_xpay", c_double),
                ("iterations_without_update", c_int), ("pad", c_int)]


WL_STREAM_READ, WL_STREAM_WRITE, WL_STREAM_COPY = 0, 1, 2
WL_NOALLOC_SCRATCH_BYTES = 64 << 20


class NativeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libwilson_amd error {code}: {msg}")
        self.code = code


_lib: Optional[ctypes.CDLL] = None


def declared_symbols() -> List[str]:
    """Function names declared in include/wilson_amd.h."""
    txt = HEADER_PATH.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wl_[a-z0-9_]+)\s*\(", txt)))


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the Wilson operator has no CPU fallback)"
        )
    lib = ctypes.CDLL(str(LIB_PATH))
    vp = c_void_p
    sig = {
        "wl_version": (c_int, []),
        "wl_last_error": (ctypes.c_char_p, []),
        "wl_ctx_create": (c_int, [POINTER(WlGeom), POINTER(WlParams), POINTER(vp)]),
        "wl_ctx_destroy": (None, [vp]),
        "wl_ctx_set_mass": (c_int, [vp, c_double]),
        "wl_ctx_set_tuning": (c_int, [vp, c_int, c_int, POINTER(WlTuning)]),
        "wl_ctx_get_tuning": (c_int, [vp, c_int, c_int, POINTER(WlTuning)]),
        "wl_ctx_set_eo_slab": (c_int, [vp, c_int, c_int, c_int]),
        "wl_gauge_upload": (c_int, [vp, vp, c_int, c_int, c_int, POINTER(vp)]),
        "wl_gauge_release": (None, [vp]),
        "wl_gauge_prepare": (c_int, [vp, c_int, c_int, vp]),
        "wl_apply": (c_int, [vp, vp, c_int, vp, vp, c_int, c_int, vp]),
        "wl_apply_noalloc": (c_int, [vp, vp, c_int, vp, vp, c_int, c_int, vp]),
        "wl_field_create": (c_int, [vp, c_int, c_int, POINTER(vp)]),
        "wl_field_destroy": (None, [vp]),
        "wl_field_import": (c_int, [vp, vp, c_int, vp]),
        "wl_field_export": (c_int, [vp, vp, c_int, vp]),
        "wl_field_bytes": (c_size_t, [vp]),
        "wl_dslash": (c_int, [vp, vp, c_int, c_int, vp, vp, c_int, vp]),
        "wl_cg_normal": (c_int, [vp, vp, vp, vp, vp, c_int, c_int, POINTER(WlSolveParams), POINTER(WlSolveStats), vp]),
        "wl_cg_normal_field": (c_int, [vp, vp, vp, vp, c_int, c_int, POINTER(WlSolveParams), POINTER(WlSolveStats), vp]),
        "wl_eo_apply": (c_int, [vp, vp, c_int, c_int, vp, vp, c_int, c_int, vp]),
        "wl_pf_force": (c_int, [vp, vp, vp, vp, vp, c_int, c_int, vp]),
        "wl_pf_force_field": (c_int, [vp, vp, vp, vp, vp, c_int, c_int, vp]),
        "wl_cg_eo_normal": (c_int, [vp, vp, vp, vp, vp, c_int, c_int, POINTER(WlSolveParams), POINTER(WlSolveStats), vp]),
        "wl_lime_decode": (c_int, [vp, c_size_t, c_int, POINTER(c_int), c_int, c_int, c_int, c_int, vp, c_int, vp]),
        "wl_ctx_set_profiling": (c_int, [vp, c_int]),
        "wl_ctx_get_profile": (c_int, [vp, POINTER(WlProfile)]),
        "wl_stream_probe": (c_int, [c_int, c_int, c_int, vp, vp, c_size_t, c_int, vp]),
    }
    sig.update({
        "wl_batch_bytes": (c_size_t, [vp, c_int, c_int]),
        "wl_batch_create": (c_int, [vp, vp, c_int, POINTER(WlSolveParams), vp, POINTER(vp)]),
        "wl_batch_destroy": (None, [vp]),
        "wl_batch_admit": (c_int, [vp, c_int, vp, c_int, c_double]),
        "wl_batch_step": (c_int, [vp]),
        "wl_batch_query": (c_int, [vp, POINTER(WlSolveStats), POINTER(c_int)]),
        "wl_batch_retire": (c_int, [vp, c_int, vp, c_int]),
    })
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != WL_OK:
        lib = load()
        msg = lib.wl_last_error()
        raise NativeError(rc, msg.decode() if msg else "")


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("the native Wilson operator needs a HIP device (no CPU fallback)")


def current_stream() -> int:
    return torch.cuda.current_stream().cuda_stream


class Context:
    """Owns a wl_ctx (geometry, mass, boundary phases) on the current device."""

    def __init__(self, lattice_shape, nc: int, batch: int, mass: float, wilson_r: float, bc):
        require_gpu()
        lib = load()
        L = tuple(int(v) for v in lattice_shape)
        if len(L) != 4:
            raise ValueError(f"the native operator supports Nd=4 lattices only (got {L})")
        g = WlGeom(4, (c_int * 4)(*L), int(nc), int(batch))
        bc = [complex(v) for v in bc]
        p = WlParams(float(mass), float(wilson_r), (c_double * 4)(*[z.real for z in bc]),
                     (c_double * 4)(*[z.imag for z in bc]))
        h = c_void_p()
        check(lib.wl_ctx_create(byref(g), byref(p), byref(h)))
        self.h = h
        self.lattice_shape = L
        self.nc = int(nc)
        self.batch = int(batch)
        self.mass = float(mass)
        self.volume = L[0] * L[1] * L[2] * L[3]
        self._lib = lib

    def set_tuning(self, prec: int, multi_rhs: bool, march: int = 0, rhs_per_thread: int = 0,
                   xcd_remap: bool = True, slab: int = 0, cols_outer: int = 0, blocks_per_cu: int = 0,
                   block_shape: int = 0, plane: int = 0, eo_slab: int = 0, lds_box: int = 0,
                   slab_x2: int = 0) -> None:
        t = WlTuning(int(march), int(rhs_per_thread), 1 if xcd_remap else 0, int(slab), int(cols_outer),
                     int(block_shape), int(blocks_per_cu), int(plane), int(eo_slab), int(lds_box), int(slab_x2))
        check(self._lib.wl_ctx_set_tuning(self.h, int(prec), 1 if multi_rhs else 0, byref(t)))

    def set_eo_slab(self, prec: int, multi_rhs: bool, eo_slab: int) -> None:
        check(self._lib.wl_ctx_set_eo_slab(self.h, int(prec), 1 if multi_rhs else 0, int(eo_slab)))

    def get_tuning(self, prec: int, nrhs: int = 1):
        t = WlTuning()
        check(self._lib.wl_ctx_get_tuning(self.h, int(prec), int(nrhs), byref(t)))
        return {"march": t.march, "rhs_per_thread": t.rhs_per_thread, "xcd_remap": bool(t.xcd_remap),
                "slab": t.slab, "cols_outer": int(t.cols_outer), "blocks_per_cu": t.blocks_per_cu,
                "block_shape": t.block_shape, "plane": t.plane, "eo_slab": t.eo_slab, "lds_box": t.lds_box,
                "slab_x2": t.slab_x2}

    # single-RHS shortcuts
    def set_march(self, prec: int, march: int) -> None:
        cur = self.get_tuning(prec, 1)
        self.set_tuning(prec, False, march, 0, cur["xcd_remap"], cur["slab"] if march == 1 else 0, cur["cols_outer"],
                        cur["blocks_per_cu"], cur["block_shape"] if march == 1 else 0, eo_slab=cur["eo_slab"])

    def get_march(self, prec: int) -> int:
        return self.get_tuning(prec, 1)["march"]

    def set_mass(self, mass: float) -> None:
        check(self._lib.wl_ctx_set_mass(self.h, float(mass)))
        self.mass = float(mass)

    def set_profiling(self, on: bool) -> None:
        """Kernel timing of CG iterations (host-driven iterations with HIP
        events between launches); switching resets the sums."""
        check(self._lib.wl_ctx_set_profiling(self.h, 1 if on else 0))

    def profile(self) -> dict:
        p = WlProfile()
        check(self._lib.wl_ctx_get_profile(self.h, byref(p)))
        return {"iterations": [p.iterations[0], p.iterations[1]], "us_op": [p.us_op[0], p.us_op[1]],
                "us_op_dag": [p.us_op_dag[0], p.us_op_dag[1]], "us_xpay": p.us_xpay,
                "iterations_without_update": p.iterations_without_update}

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self._lib.wl_ctx_destroy(h)
            except Exception:
                pass
            self.h = None


class Gauge:
    """Device-resident gauge configurations (library-owned handle)."""

    def __init__(self, ctx: Context, U: torch.Tensor, layout: int):
        if not U.is_cuda:
            raise ValueError("gauge tensor must be on the GPU")
        U = U.contiguous()
        dtype = WL_C128 if U.dtype == torch.complex128 else WL_C64
        h = c_void_p()
        check(ctx._lib.wl_gauge_upload(ctx.h, c_void_p(U.data_ptr()), int(layout), dtype, 0, byref(h)))
        self.h = h
        self.ctx = ctx

    def prepare(self, prec: int, recon: int) -> None:
        check(self.ctx._lib.wl_gauge_prepare(self.h, int(prec), int(recon), c_void_p(current_stream())))

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.ctx._lib.wl_gauge_release(h)
            except Exception:
                pass
            self.h = None


class Field:
    """Internal-layout (structure of arrays) spinor columns on the device."""

    def __init__(self, ctx: Context, prec: int, ncol: int):
        h = c_void_p()
        check(ctx._lib.wl_field_create(ctx.h, int(prec), int(ncol), byref(h)))
        self.h = h
        self.ctx = ctx
        self.prec = int(prec)
        self.ncol = int(ncol)

    @property
    def nbytes(self) -> int:
        return int(self.ctx._lib.wl_field_bytes(self.h))

    def import_(self, src: torch.Tensor) -> None:
        src = src.contiguous()
        dt = WL_C128 if src.dtype == torch.complex128 else WL_C64
        check(self.ctx._lib.wl_field_import(self.h, c_void_p(src.data_ptr()), dt, c_void_p(current_stream())))

    def export(self, dst: torch.Tensor) -> None:
        dt = WL_C128 if dst.dtype == torch.complex128 else WL_C64
        check(self.ctx._lib.wl_field_export(self.h, c_void_p(dst.data_ptr()), dt, c_void_p(current_stream())))

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                self.ctx._lib.wl_field_destroy(h)
            except Exception:
                pass
            self.h = None


def solve_params(tol: float, maxiter: int, stop_rule: int = WL_STOP_TRUE_RESIDUAL, inner_prec: int = WL_F64,
                 inner_recon: int = 18, delta: float = 0.1, poll_every: int = 8, use_graph: bool = False,
                 masses=None, accum: str = "auto", search: str = "recurrence"):
    sp = WlSolveParams()
    sp.accum_prec = ACCUM[accum]
    sp.search_update = SEARCH[search]
    sp.tol = float(tol)
    sp.maxiter = int(maxiter)
    sp.stop_rule = int(stop_rule)
    sp.inner_prec = int(inner_prec)
    sp.inner_recon = int(inner_recon)
    sp.delta = float(delta)
    sp.poll_every = int(poll_every)
    sp.use_graph = 1 if use_graph else 0
    keep = None
    if masses is not None:
        keep = (c_double * len(masses))(*[float(m) for m in masses])
        sp.mass = ctypes.cast(keep, POINTER(c_double))
    else:
        sp.mass = POINTER(c_double)()
    return sp, keep


def stats_list(stats_arr, n: int):
    return [
        {
            "iters": int(stats_arr[i].iters),
            "status": STATUS_NAMES.get(int(stats_arr[i].status), str(stats_arr[i].status)),
            "reliable_updates": int(stats_arr[i].reliable_updates),
            "true_resid": float(stats_arr[i].true_resid),
            "iter_resid": float(stats_arr[i].iter_resid),
        }
        for i in range(n)
    ]
