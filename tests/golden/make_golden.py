"""Write the golden fixtures of the Wilson path from the CPU oracle.

The reference holds no stored vectors for this path (its checks are runtime
invariants), so the fixtures are produced by the NumPy restatement, which is
itself pinned by tests/test_oracle.py.  Run from the repo root:

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from jaxqft_amd import fields  # noqa: E402
from oracle import wilson_oracle as wo  # noqa: E402

OUT = Path(__file__).resolve().parent


def operator_case(name, L, nc, mass, eps, bc, seed):
    rng = fields.rng_for(seed)
    U = fields.su3_hot_start(rng, 1, L, eps) if nc == 3 else fields.u1_hot_start(rng, 1, L, eps)
    psi = fields.random_fermion(rng, (1, *L, 4, nc))
    ora = wo.WilsonOracle(L, nc, mass, fermion_bc=bc)
    np.savez_compressed(
        OUT / f"{name}.npz",
        U=U, psi=psi, mass=mass, bc=np.asarray(bc, dtype=np.complex128), L=np.asarray(L),
        D=ora.apply_D(U, psi), Ddag=ora.apply_Ddag(U, psi), g5D=ora.apply_g5D(U, psi),
        DdagD=ora.apply_normal(U, psi), K=ora.apply_offdiag(U, psi), Kdag=ora.apply_offdiag_dagger(U, psi),
    )


def cg_case():
    """Plain FP64 SU(3) D^dag D CG at 4^4: eps=0.3, m=0.1, (1,1,1,-1), true residual 1e-10."""
    L = (4, 4, 4, 4)
    rng = fields.rng_for(fields.SEED_BASE + 0)
    U = fields.su3_hot_start(rng, 1, L, 0.3)
    b = fields.random_fermion(rng, (1, *L, 4, 3))
    ora = wo.WilsonOracle(L, 3, 0.1, fermion_bc=(1, 1, 1, -1))
    x, iters = wo.cg(lambda v: ora.apply_normal(U, v), b, tol=1e-12, maxiter=5000)
    res = np.linalg.norm(ora.apply_normal(U, x) - b) / np.linalg.norm(b)
    assert res <= 1e-11, res
    np.savez_compressed(OUT / "cg_su3_4x4x4x4.npz", U=U, b=b, x=x, iters=iters, true_resid=res, mass=0.1,
                        bc=np.asarray((1, 1, 1, -1), dtype=np.complex128))


def index_case(L):
    t = wo.checkerboard_neighbor_maps(L)
    np.savez_compressed(OUT / f"index_{'x'.join(map(str, L))}.npz", **t)


def main():
    operator_case("op_su3_2x2x2x2", (2, 2, 2, 2), 3, 0.1, 0.3, (1, 1, 1, -1), 101)
    operator_case("op_su3_4x4x4x4", (4, 4, 4, 4), 3, 0.1, 0.3, (1, 1, 1, -1), 102)
    operator_case("op_su3_2x3x4x5", (2, 3, 4, 5), 3, 0.05, 0.5, (np.exp(0.4j), 1, -1, -1), 103)
    operator_case("op_u1_2x2x2x2", (2, 2, 2, 2), 1, 0.1, 0.3, (1, 1, 1, -1), 104)
    operator_case("op_u1_4x4x4x4", (4, 4, 4, 4), 1, 0.2, 0.5, (1, 1, 1, -1), 105)
    cg_case()
    for L in [(4, 4, 4, 4), (8, 8, 8, 8), (4, 4, 4, 8)]:
        index_case(L)


if __name__ == "__main__":
    main()
