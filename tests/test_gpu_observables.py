"""<psi|H|phi> and <psi|phi> (mps/observables.cpp:64-105, 415-450) on the
device contraction engine, pinned by closed forms: the Neel product state of
the open Heisenberg chain has <H> = -(N-1)/4 and norm 1; a DMRG ground state
has <psi|H|psi> / <psi|psi> equal to the DMRG energy and to exact
diagonalisation (tests/golden/ed_energies.json), and a random MPS's
<psi|H|psi> / <psi|psi> is the energy DMRG's initial H_eff expectation
reports (same state, same operator)."""
import json
import os

import numpy as np
import pytest

from ttgpu import host as P
from ttgpu.workloads import random_mps

pytestmark = pytest.mark.gpu

ED = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))


def test_neel_state_closed_form():
    N = 12
    H = P.mpo_heisenberg_nn(N)
    psi = P.mps_product([i % 2 for i in range(N)])
    assert P.psiphi(psi, psi) == pytest.approx(1.0, abs=1e-14)
    assert P.psiHphi(psi, H, psi) == pytest.approx(-(N - 1) / 4.0, abs=1e-13)


def test_dmrg_ground_state_expectation():
    N = 10
    H = P.mpo_heisenberg_nn(N)
    psi = P.mps_product([i % 2 for i in range(N)])
    r = P.dmrg(psi, H, 6, [16, 32], [1e-12], [2], 1e-12)
    nrm = P.psiphi(psi, psi)
    e = P.psiHphi(psi, H, psi) / nrm
    assert nrm == pytest.approx(1.0, abs=1e-10)
    assert e == pytest.approx(r["E"], abs=1e-10)
    assert abs(e - ED["heisenberg_obc_N10"]) < 1e-9


def test_random_mps_expectation_matches_initial_heff():
    N = 8
    H = P.mpo_heisenberg_nn(N)
    psi = P.mps_product([i % 2 for i in range(N)])
    for i, t in enumerate(random_mps(N, 12, 2.0, seed=3)):
        psi.set_tensor(i, P.PTensor.from_flat(t))
    psi.set_center(-1)
    nrm = P.psiphi(psi, psi)
    e = P.psiHphi(psi, H, psi) / nrm
    # dmrg() canonicalises, normalises and reports <psi|H_eff|psi> at site 0
    # before sweeping: the same number
    r = P.dmrg(psi, H, 1, [12], [0.0], [1], 1e-12)
    assert e == pytest.approx(r["E0"], rel=1e-11)


def test_contract_with_different_block_lists_matches_oracle():
    """qtensor::contract (qtensor.cpp:1095-1119) over two random subsets of
    the legal blocks: only blocks both hold contribute."""
    import oracle as O
    from ttgpu import lib as T
    from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor
    rng = np.random.default_rng(5)
    a = make_index(INWARD, "a", LINK, 0, [((q, 0), d) for q, d in zip([-2, 0, 2], [3, 5, 2])])
    s = make_index(INWARD, "s", SITE, 0, [((-1, 0), 1), ((1, 0), 1)])
    b = make_index(OUTWARD, "b", LINK, 0, [((q, 0), d) for q, d in zip([-3, -1, 1, 3], [2, 4, 4, 1])])
    X = random_tensor([a, s, b], rng, keep=0.7)
    Y = random_tensor([a, s, b], rng, keep=0.6)
    ref = O.OTensor.from_flat(X).inner(O.OTensor.from_flat(Y))
    got = T.DTensor.from_flat(X).dot_matched(T.DTensor.from_flat(Y))
    assert got == pytest.approx(ref, rel=1e-13, abs=1e-13)
