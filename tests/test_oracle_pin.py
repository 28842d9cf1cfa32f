"""Pins the parts of the oracle that the dense checks of test_oracle.py
leave open: the truncation keep rule (qtensor_op.cpp:1289-1301) against
invariants of numpy singular values, and the block and task order of
operator* (qtensor.cpp:606-722) against an independent numpy statement
(tests/ref_order.py), on shapes of the three H_eff stages and of the
environment update, with missing blocks."""
import numpy as np
import pytest

import oracle as O
from dense_ref import check_keep_rule, sector_singular_values
from ref_order import contraction_order
from scenarios import heff_inputs, mpo_bond, two_site_case
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor


def sec(qs, ds):
    return [((q, 0), d) for q, d in zip(qs, ds)]


def kept_by_sector(S):
    f = S.to_flat()
    return {tuple(q): d for q, d in zip(f.idx[0].qn, f.idx[0].qdim)}


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("cutoff,K,zero", [
    (0.0, 0, None),        # everything kept
    (0.0, 0, 0),           # zero values kept too at cutoff 0
    (1e-2, 0, None),       # the cutoff ends the loop
    (2e-1, 0, 2),          # cutoff with a zero sector present
    (0.0, 9, None),        # K binds
    (1e-12, 9, None),      # K binds before the cutoff
    (1e-300, 0, 2),        # zero values never kept when cutoff != 0
    (1e-300, 200, -2),     # the zero stops the loop before K
])
def test_keep_rule_invariants(direction, cutoff, K, zero):
    A, B, x = two_site_case(3, zero_sector=zero)
    sv = sector_singular_values(x, 2)
    tr, kept, S = O.svd_bond(O.OTensor.from_flat(x), O.OTensor.from_flat(A), O.OTensor.from_flat(B),
                             "ID1Link4", direction, cutoff, K)
    check_keep_rule(sv, cutoff, K, kept, tr, kept_by_sector(S))
    # the kept values are the largest ones, sector by sector
    s = S.to_flat()
    for b in range(s.nblocks):
        q = tuple(s.idx[0].qn[s.block_qi[b, 0]])
        got = np.sort(s.block(b))[::-1]
        assert np.allclose(got, sv[q][:len(got)], rtol=1e-12, atol=1e-13)


def test_keep_rule_check_catches_an_off_by_one():
    # the invariants reject a kept count one off in either direction
    A, B, x = two_site_case(3)
    sv = sector_singular_values(x, 2)
    tr, kept, S = O.svd_bond(O.OTensor.from_flat(x), O.OTensor.from_flat(A), O.OTensor.from_flat(B),
                             "ID1Link4", 0, 1e-2, 0)
    allv = np.sort(np.concatenate(list(sv.values())))[::-1]
    for k in (kept - 1, kept + 1):
        trk = max(1.0 - float(np.sum(allv[:k] ** 2) / np.sum(allv ** 2)), 0.0)
        with pytest.raises(AssertionError):
            check_keep_rule(sv, 1e-2, 0, k, trk)


def assert_order(A, B):
    oA, oB = O.OTensor.from_flat(A), O.OTensor.from_flat(B)
    pairs = oA.pairs(oB)
    blocks, tasks = contraction_order(A, B, pairs)
    C = (oA * oB).to_flat()
    assert np.array_equal(C.block_qi, blocks)
    assert np.array_equal(O.last_tasks().astype(np.int64), tasks)
    return len(tasks)


@pytest.mark.parametrize("seed", [0, 1])
def test_task_order_heff_stages(seed):
    # the three stages of big_qtensor::product on a small bond, operands with
    # missing blocks (keep < 1), including the rank-5 x rank-6 stage 2
    rng = np.random.default_rng(seed)
    W1, W2 = mpo_bond("nn_chain", 6)
    L, R, x = heff_inputs(W1, W2, 6, 40, 4, 2.0, seed=seed)
    x = random_tensor(x.idx, rng, keep=0.8)
    L = random_tensor(L.idx, rng, keep=0.9)
    oL, ox = O.OTensor.from_flat(L), O.OTensor.from_flat(x)
    big = O.OBig(oL, O.OTensor.from_flat(W1), O.OTensor.from_flat(W2), O.OTensor.from_flat(R))
    W12 = big.mid().to_flat()
    assert W12.rank == 6
    n1 = assert_order(L, x)
    T1 = (oL * ox).to_flat()
    assert T1.rank == 5
    n2 = assert_order(T1, W12)
    T2 = (O.OTensor.from_flat(T1) * O.OTensor.from_flat(W12)).to_flat()
    n3 = assert_order(T2, R)
    assert min(n1, n2, n3) > 0


def test_task_order_env_update_and_multi_pair():
    # environment update shapes (psi^dag' * TL, * W, * psi) and a contraction
    # over two index pairs whose mid tuples have two components
    rng = np.random.default_rng(7)
    a = make_index(INWARD, "a", LINK, 0, sec([-1, 0, 1, 2], [3, 1, 4, 2]))
    b = make_index(OUTWARD, "b", LINK, 0, sec([-2, 0, 2], [2, 3, 2]))
    c = make_index(INWARD, "c", LINK, 0, sec([-1, 1, 3], [5, 2, 1]))
    d = make_index(OUTWARD, "d", LINK, 2, sec([0, 1, 2], [2, 2, 3]))
    e = make_index(INWARD, "e", LINK, 0, sec([-2, -1, 0, 1], [1, 2, 3, 2]))
    A = random_tensor([a, b, c, d], rng, keep=0.7)
    B = random_tensor([e, d.dag(), b.dag()], rng, keep=0.7)
    assert assert_order(A, B) > 0
    W1, _ = mpo_bond("nn_chain", 6)
    l = make_index(INWARD, "ID1Link6", LINK, 0, sec([-2, 0, 2], [3, 4, 3]))
    s = make_index(INWARD, "Site6", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "ID1Link7", LINK, 0, sec([-3, -1, 1, 3], [2, 4, 4, 2]))
    psi = random_tensor([l, s, r], rng)
    lp = l.copy(); lp.level = 1
    TL = random_tensor([lp, W1.idx[0].dag(), l.dag()], rng)
    psid = O.OTensor.from_flat(psi)
    psid.dag(); psid.prime()
    t1 = psid.to_flat()
    assert assert_order(t1, TL) > 0
    t4 = (O.OTensor.from_flat(t1) * O.OTensor.from_flat(TL)).to_flat()
    assert assert_order(t4, W1) > 0
    t5 = (O.OTensor.from_flat(t4) * O.OTensor.from_flat(W1)).to_flat()
    assert assert_order(t5, psi) > 0


def test_order_statement_rejects_a_swapped_loop():
    # the statement is sensitive to the loop nesting: swapping the right and
    # left loops gives a different block list on a case with several of each
    rng = np.random.default_rng(5)
    l = make_index(INWARD, "l", LINK, 0, sec([-2, 0, 2], [2, 3, 1]))
    w = make_index(OUTWARD, "w", LINK, 0, sec([-2, 0, 2], [1, 2, 1]))
    s = make_index(INWARD, "s", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "r", LINK, 0, sec([-3, -1, 1, 3], [1, 2, 2, 1]))
    lp = l.copy(); lp.level = 1
    L = random_tensor([lp, w, l.dag()], rng)
    x = random_tensor([l, s, r], rng)
    blocks, _ = contraction_order(L, x, [(2, 0)])
    swapped = sorted(blocks.tolist(), key=lambda t: (t[0], t[1]))
    assert blocks.tolist() != swapped


def _dense_state(psi):
    t = psi.tensor(0)
    for i in range(1, len(psi)):
        t = t * psi.tensor(i)
    return t.to_flat().to_dense().ravel()


@pytest.mark.parametrize("N,m,k", [(6, 8, 2), (6, 8, 0), (8, 12, 7), (8, 12, 3)])
def test_position_keeps_the_state_and_makes_it_canonical(N, m, k):
    """position(k) (qtt.cpp:932-994, through rc()/lc() at cutoff 0) leaves the
    state unchanged (the dense contraction of all site tensors, up to norm and
    sign) and makes every site left of k left-orthonormal and every site
    right of k right-orthonormal."""
    from ttgpu.workloads import random_mps
    psi = O.mps_product([i % 2 for i in range(N)])
    for i, t in enumerate(random_mps(N, m, 2.0, seed=5)):
        psi.set_tensor(i, O.OTensor.from_flat(t))
    psi.set_center(-1)
    before = _dense_state(psi)
    psi.position(k)
    assert psi.center() == k
    after = _dense_state(psi)
    b, a = before / np.linalg.norm(before), after / np.linalg.norm(after)
    assert min(np.linalg.norm(a - b), np.linalg.norm(a + b)) < 1e-13
    for i in range(N):
        A = psi.tensor(i).to_flat().to_dense()
        Dl, d, Dr = A.shape
        if i < k:
            M = A.reshape(Dl * d, Dr, order="F")
            G = M.T @ M
        elif i > k:
            M = A.reshape(Dl, d * Dr, order="F")
            G = M @ M.T
        else:
            continue
        assert np.abs(G - np.eye(G.shape[0])).max() < 1e-13, (i, k)


def _end_env(i0, w0, sign_flip=True):
    """The open-boundary end of an environment (dmrg.cpp:60-74): indices
    (psi link primed and dagged, MPO link, psi link), every legal block one,
    then every arrow flipped."""
    from ttgpu.flat import FlatTensor, legal_blocks
    a = i0.dag()
    a.level += 1
    idx = [a, w0.copy(), i0.copy()]
    bq = legal_blocks(idx)
    t = FlatTensor(idx, bq, np.zeros(0))
    t.data = np.ones(int(t.block_sizes().sum()))
    t.idx = [ix.dag() for ix in t.idx]
    return t


def test_davidson_reaches_the_lowest_eigenvalue_of_the_dense_operator():
    """tensor_davidsonIT (tensor_davidson.cpp:192-347) on a two-site H_eff
    built from environments of a real chain (Hermitian): with many restarts it
    converges to the lowest eigenvalue of the dense operator (numpy eigh),
    and with the reference's max_restart = 2 it stays between that value and
    the starting Rayleigh quotient (variational)."""
    from ttgpu.workloads import random_mps
    N = 8
    H = O.mpo_heisenberg_nn(N)
    psi = O.mps_product([i % 2 for i in range(N)])
    for i, t in enumerate(random_mps(N, 8, 2.0, seed=3)):
        psi.set_tensor(i, O.OTensor.from_flat(t))
    psi.set_center(-1)
    psi.position(3)
    A = [psi.tensor(i) for i in range(N)]
    W = [H.tensor(i) for i in range(N)]
    TL = O.OTensor.from_flat(_end_env(A[0].to_flat().idx[0], W[0].to_flat().idx[0]))
    for i in range(3):
        TL = O.env_step(A[i], TL, W[i])
    TR = O.OTensor.from_flat(_end_env(A[N - 1].to_flat().idx[2], W[N - 1].to_flat().idx[3]))
    for i in range(N - 1, 4, -1):
        TR = O.env_step(A[i], TR, W[i])
    big = O.OBig(TL, W[3], W[4], TR)
    x0 = A[3] * A[4]
    f = x0.to_flat()
    n = f.data.size
    cols = []
    for j in range(n):
        e = f.data * 0.0
        e[j] = 1.0
        u = O.OTensor.from_flat(type(f)(f.idx, f.block_qi, e))
        y = big.product(u).to_flat()
        pos = {tuple(q): b for b, q in enumerate(y.block_qi.tolist())}
        yo, fo = y.block_offsets(), f.block_offsets()
        col = np.zeros(n)
        for b, q in enumerate(f.block_qi.tolist()):
            if tuple(q) in pos:
                g = pos[tuple(q)]
                col[fo[b]:fo[b + 1]] = y.data[yo[g]:yo[g + 1]]
        cols.append(col)
    Hd = np.array(cols).T
    assert np.allclose(Hd, Hd.T, atol=1e-12 * np.abs(Hd).max())
    lam = np.linalg.eigvalsh(0.5 * (Hd + Hd.T))[0]
    rq = float(f.data @ Hd @ f.data / (f.data @ f.data))
    e_many, _ = big.davidson(x0.copy(), max_restart=60, tol=1e-12)
    assert abs(e_many - lam) < 1e-10 * abs(lam)
    e2, _ = big.davidson(x0.copy(), max_restart=2, tol=1e-12)
    assert lam - 1e-12 <= e2 <= rq + 1e-12
