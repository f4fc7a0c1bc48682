"""ctypes loader of the C/OpenMP oracle (test infrastructure / CPU baseline only)."""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libwilson_oracle.so"
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
        lib = ctypes.CDLL(str(LIB))
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.wo_apply.argtypes = [dp, dp, dp, ip, ctypes.c_int, ctypes.c_double, dp, dp, ctypes.c_int]
        lib.wo_apply.restype = None
        lib.wo_cg.argtypes = [dp, dp, dp, ip, ctypes.c_int, ctypes.c_double, dp, dp, ctypes.c_double, ctypes.c_int]
        lib.wo_cg.restype = ctypes.c_int
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _args(L, bc):
    Lc = np.asarray(L, dtype=np.int32)
    bc = np.asarray([complex(v) for v in bc])
    return Lc, np.ascontiguousarray(bc.real), np.ascontiguousarray(bc.imag)


def apply(U, psi, L, nc, mass, bc, sign):
    """U: (4, *L, nc, nc) or (4, *L) complex128 BMXYIJ without batch; psi (*L, 4, nc)."""
    lib = load()
    U = np.ascontiguousarray(U, dtype=np.complex128)
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.empty_like(psi)
    Lc, br, bi = _args(L, bc)
    lib.wo_apply(_p(U), _p(psi), _p(out), Lc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(nc), float(mass),
                 _p(br), _p(bi), int(sign))
    return out


def cg(U, b, L, nc, mass, bc, tol, maxiter):
    lib = load()
    U = np.ascontiguousarray(U, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.empty_like(b)
    Lc, br, bi = _args(L, bc)
    k = lib.wo_cg(_p(U), _p(b), _p(x), Lc.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(nc), float(mass),
                  _p(br), _p(bi), float(tol), int(maxiter))
    return x, int(k)
