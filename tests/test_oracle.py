"""Pins the oracle (CPU FP64 restatement) to what mathematics fixes: dense
numpy contractions, Kronecker-product Hamiltonians, numpy SVD / eigh, and
exact-diagonalisation energies (tests/golden/ed_energies.json, made by
tests/golden/make_ed.py).  A dropped term, a wrong sign or a transposed
operand in the restatement fails one of these."""
import json
import os

import numpy as np
import pytest

import oracle as O
from dense_ref import einsum_contract, heisenberg_dense, mpo_to_dense
from scenarios import heff_inputs, mpo_bond, sector_dims
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, legal_blocks, make_index, random_tensor

ED = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))


def sec(qs, ds):
    return [((q, 0), d) for q, d in zip(qs, ds)]


def test_init_block_matches_bruteforce_enumeration():
    a = make_index(INWARD, "a", LINK, 0, sec([-2, 0, 2], [1, 2, 1]))
    b = make_index(INWARD, "b", SITE, 0, sec([-1, 1], [1, 1]))
    c = make_index(OUTWARD, "c", LINK, 0, sec([-3, -1, 1, 3], [1, 3, 3, 1]))
    t = O.init_block([a, b, c]).to_flat()
    assert np.array_equal(t.block_qi, legal_blocks([a, b, c]))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_contract_matches_dense_einsum(seed):
    rng = np.random.default_rng(seed)
    a = make_index(INWARD, "a", LINK, 0, sec([-1, 0, 1, 2], [3, 1, 4, 2]))
    b = make_index(OUTWARD, "b", LINK, 0, sec([-2, 0, 2], [2, 3, 2]))
    c = make_index(INWARD, "c", LINK, 0, sec([-1, 1, 3], [5, 2, 1]))
    d = make_index(OUTWARD, "d", LINK, 2, sec([0, 1, 2], [2, 2, 3]))
    e = make_index(INWARD, "e", LINK, 0, sec([-2, -1, 0, 1], [1, 2, 3, 2]))
    A = random_tensor([a, b, c, d], rng, keep=0.7)
    B = random_tensor([e, d.dag(), b.dag()], rng, keep=0.7)
    oA, oB = O.OTensor.from_flat(A), O.OTensor.from_flat(B)
    pairs = oA.pairs(oB)
    assert pairs == [(3, 1), (1, 2)]          # paired in B's order
    Cd = (oA * oB).to_flat().to_dense()
    assert np.allclose(Cd, einsum_contract(A, B, pairs), atol=1e-13)


def test_result_block_order_is_q_right_left():
    # independent statement of the reference order: sort (mid QN, right
    # tuple, left tuple) over every (left, right) combination of equal flux
    rng = np.random.default_rng(5)
    l = make_index(INWARD, "l", LINK, 0, sec([-2, 0, 2], [2, 3, 1]))
    w = make_index(OUTWARD, "w", LINK, 0, sec([-2, 0, 2], [1, 2, 1]))
    s = make_index(INWARD, "s", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "r", LINK, 0, sec([-3, -1, 1, 3], [1, 2, 2, 1]))
    lp = l.copy(); lp.level = 1
    L = random_tensor([lp, w, l.dag()], rng)
    x = random_tensor([l, s, r], rng)
    C = (O.OTensor.from_flat(L) * O.OTensor.from_flat(x)).to_flat()
    lefts = {tuple(b[:2]) for b in L.block_qi}
    rights = {tuple(b[1:]) for b in x.block_qi}
    flux_l = lambda t: lp.qn[t[0]][0] - w.qn[t[1]][0]
    flux_r = lambda t: -s.qn[t[0]][0] + r.qn[t[1]][0]
    combos = [(flux_l(lt), rt, lt) for lt in lefts for rt in rights if flux_l(lt) == flux_r(rt)]
    expect = [list(lt) + list(rt) for _, rt, lt in sorted(combos)]
    assert C.block_qi.tolist() == expect


def test_nn_mpo_reproduces_heisenberg_chain():
    for N in (4, 6):
        H = O.mpo_heisenberg_nn(N)
        D = mpo_to_dense([H.tensor(i).to_flat().to_dense() for i in range(N)])
        bonds = [(i, i + 1) for i in range(N - 1)]
        assert np.abs(D - heisenberg_dense(N, bonds, [1.0] * len(bonds))).max() == 0.0


@pytest.mark.parametrize("kind,Nx,Ny,yper,J2", [(1, 7, 1, False, 0.5), (0, 3, 4, True, 0.0),
                                                (1, 3, 3, True, 0.5)])
def test_autompo_reproduces_lattice_hamiltonian(kind, Nx, Ny, yper, J2):
    N = Nx * Ny
    terms = O.heisenberg_terms(kind, Nx, Ny, yper, 1.0, J2)
    H = O.mpo_autompo(N, terms)
    D = mpo_to_dense([H.tensor(i).to_flat().to_dense() for i in range(N)])
    bonds = list(zip(terms[2][::3].tolist(), terms[4][::3].tolist()))
    assert np.abs(D - heisenberg_dense(N, bonds, terms[0][::3].tolist())).max() < 1e-14


def test_dense_svd_matches_numpy():
    rng = np.random.default_rng(3)
    for shape in [(7, 4), (4, 7), (6, 6), (1, 2), (5, 1)]:
        A = rng.standard_normal(shape)
        U, S, Vt = O.dense_svd(A)
        assert np.allclose(S, np.linalg.svd(A, compute_uv=False), atol=1e-13)
        assert np.allclose((U * S) @ Vt, A, atol=1e-13)
        assert np.allclose(U.T @ U, np.eye(U.shape[1]), atol=1e-13)
    # rank-deficient: U still an isometry
    A = np.outer(rng.standard_normal(5), rng.standard_normal(4))
    U, S, Vt = O.dense_svd(A)
    assert np.allclose(U.T @ U, np.eye(4), atol=1e-12)


def test_diag_matches_eigh():
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 4):
        A = rng.standard_normal((n, n))
        A = A + A.T
        ev, V = O.diag(A)
        assert np.allclose(ev, np.linalg.eigvalsh(A), atol=1e-13)
        assert np.allclose(A @ V, V * ev, atol=1e-12)


def test_heff_product_matches_dense_operator():
    W1, W2 = mpo_bond("nn_chain", 4)
    L, R, x = heff_inputs(W1, W2, 4, 12, 2, 1.0, seed=9)
    oL, oR, ox = (O.OTensor.from_flat(t) for t in (L, R, x))
    big = O.OBig(oL, O.OTensor.from_flat(W1), O.OTensor.from_flat(W2), oR)
    y = big.product(ox).to_flat().to_dense()
    Ld, Rd, xd = L.to_dense(), R.to_dense(), x.to_dense()
    # y[l', s', t', r'] = sum L[l', w, l] x[l, s, t, r] W1[w, s, s', c] W2[c, t, t', w'] R[r', w', r]
    ref = np.einsum("pwl,lstr,wsSc,ctTv,qvr->pSTq", Ld, xd, W1.to_dense(), W2.to_dense(), Rd)
    assert np.allclose(y, ref, atol=1e-12)


def test_dmrg_n10_matches_exact_diagonalisation():
    N = 10
    psi = O.mps_product([i % 2 for i in range(N)])
    H = O.mpo_heisenberg_nn(N)
    E, hs, e0, _ = O.dmrg(psi, H, 5, [16, 32, 64], [1e-10], [2], 1e-12)
    assert abs(E - ED["heisenberg_obc_N10"]) < 1e-9
    assert e0 == pytest.approx(-2.25)   # Neel product state: (N-1) * (-1/4)


@pytest.mark.slow
def test_dmrg_n20_matches_exact_diagonalisation():
    N = 20
    psi = O.mps_product([i % 2 for i in range(N)])
    H = O.mpo_heisenberg_nn(N)
    E, hs, _, _ = O.dmrg(psi, H, 10, [16, 32, 48, 64, 64, 64, 64, 64, 64, 64], [1e-10], [2], 1e-10)
    assert abs(E - ED["heisenberg_obc_N20"]) / N < 3e-10


def test_sector_dims_sum_and_shape():
    d = sector_dims(512, 20, 6.0)
    assert sum(x for _, x in d) == 512 and len(d) == 41 and min(x for _, x in d) >= 1
    assert [q for q, _ in d] == list(range(-40, 41, 2))
