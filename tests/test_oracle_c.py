"""The C/OpenMP oracle (CPU baseline) agrees with the NumPy oracle."""

import numpy as np
import pytest

from jaxqft_amd import fields
from oracle import oracle_c
from oracle import wilson_oracle as wo


@pytest.mark.parametrize("nc", [3, 1])
def test_c_oracle_matches_numpy(nc):
    L = (2, 3, 4, 5)
    bc = (np.exp(0.4j), 1, -1, -1)
    rng = fields.rng_for(900 + nc)
    U = fields.su3_hot_start(rng, 1, L, 0.6) if nc == 3 else fields.u1_hot_start(rng, 1, L, 0.6)
    psi = fields.random_fermion(rng, (1, *L, 4, nc))
    ora = wo.WilsonOracle(L, nc, 0.07, fermion_bc=bc)
    for sign, ref in ((-1, ora.apply_D(U, psi)), (+1, ora.apply_Ddag(U, psi))):
        got = oracle_c.apply(U[0], psi[0], L, nc, 0.07, bc, sign)
        assert np.linalg.norm(got - ref[0]) / np.linalg.norm(ref) < 1e-14


def test_c_oracle_cg():
    z = np.load(__import__("pathlib").Path(__file__).resolve().parent / "golden" / "cg_su3_4x4x4x4.npz")
    x, k = oracle_c.cg(z["U"][0], z["b"][0], (4, 4, 4, 4), 3, 0.1, (1, 1, 1, -1), 1e-12, 5000)
    assert np.linalg.norm(x - z["x"][0]) / np.linalg.norm(z["x"]) < 1e-10
    assert k > 0
