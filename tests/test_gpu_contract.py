"""GPU parity of the block-sparse contraction (qtensor::operator*,
qtensor.cpp:523-734): the device-built block list, packed offsets and task
list must equal the oracle's bit for bit; values agree to 1e-12 relative
Frobenius error (FP64)."""
import numpy as np
import pytest

import oracle as O
from scenarios import heff_inputs, mpo_bond
from ttgpu import lib as T
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def check_contract(A, B, tol=1e-12, prec=T.FP64):
    oA, oB = O.OTensor.from_flat(A), O.OTensor.from_flat(B)
    pairs = oA.pairs(oB)
    oC = (oA * oB).to_flat()
    otasks = O.last_tasks()
    dA, dB = T.DTensor.from_flat(A), T.DTensor.from_flat(B)
    plan = T.Plan(dA, dB, pairs, prec)
    nb, nt, fl = plan.info()
    assert nb == oC.nblocks
    assert nt == len(otasks)
    dC = plan.new_output()
    plan.execute(dA, dB, dC)
    qi, off = dC.meta()
    assert np.array_equal(qi, oC.block_qi), "result block list differs from the reference order"
    assert np.array_equal(off, oC.block_offsets())
    assert np.array_equal(plan.tasks(), otasks), "task list differs from the reference loop order"
    flops = float(sum(2.0 * t[3] * t[4] * t[5] for t in otasks))
    assert fl == pytest.approx(flops, rel=1e-15)
    if len(oC.data):
        assert rel_err(dC.data(), oC.data) <= tol
    return oC, dC


def sec(qs, ds):
    return [((q, 0), d) for q, d in zip(qs, ds)]


def test_env_times_x_pattern():
    rng = np.random.default_rng(1)
    l = make_index(INWARD, "psiL", LINK, 0, sec([-4, -2, 0, 2, 4], [2, 5, 7, 3, 1]))
    w = make_index(OUTWARD, "HL", LINK, 0, sec([-2, 0, 2], [2, 4, 2]))
    s1 = make_index(INWARD, "S1", SITE, 0, sec([-1, 1], [1, 1]))
    s2 = make_index(INWARD, "S2", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "psiR", LINK, 0, sec([-4, -2, 0, 2, 4], [3, 6, 8, 2, 2]))
    lp = l.copy(); lp.level = 1
    L = random_tensor([lp, w, l.dag()], rng)
    x = random_tensor([l, s1, s2, r], rng, keep=0.8)
    check_contract(L, x)


def test_generic_two_pairs_permuted_and_ragged():
    rng = np.random.default_rng(2)
    a = make_index(INWARD, "a", LINK, 0, sec([-1, 0, 1, 2], [3, 1, 4, 2]))
    b = make_index(OUTWARD, "b", LINK, 0, sec([-2, 0, 2], [2, 3, 2]))
    c = make_index(INWARD, "c", LINK, 0, sec([-1, 1, 3], [5, 2, 1]))
    d = make_index(OUTWARD, "d", LINK, 2, sec([0, 1, 2], [2, 2, 3]))
    e = make_index(INWARD, "e", LINK, 0, sec([-2, -1, 0, 1], [1, 2, 3, 2]))
    # A = (a, b, c, d) ; B = (e, d^dag, b^dag): pairs in B order (d, b)
    A = random_tensor([a, b, c, d], rng, keep=0.7)
    B = random_tensor([e, d.dag(), b.dag()], rng, keep=0.7)
    check_contract(A, B)
    check_contract(B, A)


def test_outer_product_no_pairs():
    rng = np.random.default_rng(3)
    a = make_index(INWARD, "a", LINK, 0, sec([-1, 0, 1], [2, 3, 2]))
    b = make_index(OUTWARD, "b", LINK, 0, sec([-1, 0, 1], [1, 2, 2]))
    c = make_index(INWARD, "c", LINK, 0, sec([0, 2], [3, 1]))
    d = make_index(OUTWARD, "d", LINK, 0, sec([0, 2], [2, 2]))
    check_contract(random_tensor([a, b], rng), random_tensor([c, d], rng))


@pytest.mark.parametrize("site", [3, 10])
def test_heff_stage_chain_and_env_update(site):
    W1, W2 = mpo_bond("nn_chain", site)
    L, R, x = heff_inputs(W1, W2, site, 40, 4, 2.0, seed=11)
    oL, oR, ox = (O.OTensor.from_flat(t) for t in (L, R, x))
    oW1, oW2 = O.OTensor.from_flat(W1), O.OTensor.from_flat(W2)
    # W12 = W1 * W2 (big_qtensor::addMid)
    W12, _ = check_contract(W1, W2)
    T1, _ = check_contract(L, x)
    T2, _ = check_contract(T1, W12)
    check_contract(T2, R)
    # env update stages (dmrg.cpp:491-509): t1 = psi^dag', t4 = t1*TL, t5 = t4*W, TL' = t5*psi
    psi = L  # any rank-3 tensor with the TL index layout is a valid environment
    rng = np.random.default_rng(5)
    l = x.idx[0]
    s = x.idx[1]
    mid = make_index(OUTWARD, "ID1Link%d" % (site + 1), LINK, 0,
                     [((q, 0), d) for q, d in zip([qq[0] for qq in l.qn], l.qdim)])
    psi = random_tensor([l, s, mid], rng)
    t1 = psi.idx
    ot1 = O.OTensor.from_flat(psi)
    ot1.dag()
    ot1.prime()
    f1 = ot1.to_flat()
    t4, _ = check_contract(f1, L)
    t5, _ = check_contract(t4, W1)
    check_contract(t5, psi)
