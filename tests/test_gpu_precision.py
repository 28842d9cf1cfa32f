"""Precision forms on a converged DMRG run: the N=20 open Heisenberg chain
(10 sweeps, bond dimension ramped 16, 32, 48, then 64 for the last seven
sweeps, cutoff 1e-10) reaches the same energy per site in FP32 and
split-BF16 as in FP64 within 1e-10, and the exact-diagonalisation energy
(tests/golden/ed_energies.json) within 3e-10 (the m = 64 truncation leaves
1.74e-10 per site in the oracle's run).  The reduced forms end
every Davidson call with an FP64 product (Rayleigh quotient)."""
import json
import os

import pytest

from ttgpu import host as P
from ttgpu import lib as T

pytestmark = pytest.mark.gpu

ED = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))


def run(prec):
    P.set_precision(prec)
    try:
        N = 20
        psi = P.mps_product([i % 2 for i in range(N)])
        r = P.dmrg(psi, P.mpo_heisenberg_nn(N), 10, [16, 32, 48, 64, 64, 64, 64, 64, 64, 64], [1e-10], [2], 1e-10)
    finally:
        P.set_precision(T.FP64)
    return r["E"] / 20


@pytest.mark.parametrize("prec", [T.FP32, T.BF16X3])
def test_reduced_precision_dmrg_energy(prec):
    e64 = run(T.FP64)
    e = run(prec)
    assert abs(e - e64) < 1e-10
    assert abs(e - ED["heisenberg_obc_N20"] / 20) < 3e-10
