"""Every grouped-GEMM work shape gives the oracle's contraction
(qtensor::operator*, qtensor.cpp:523-734): each MFMA tile shape, the skinny
vector-ALU kernel, and split-K (a long contraction walk over few output tiles,
cut into partial tiles that a second pass adds in fixed order).  The tile
shape is pinned through TTG_TILE_CFG, which the planner reads at plan time.
Tolerances: FP64 1e-12, FP32 2e-6, split-BF16 2e-5 relative Frobenius error
against the FP64 oracle."""
import numpy as np
import pytest

import oracle as O
from ttgpu import lib as T
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor

pytestmark = pytest.mark.gpu

TOL = {T.FP64: 1e-12, T.FP32: 2e-6, T.BF16X3: 2e-5}
NCFG = 12   # 11 MFMA tile shapes + the skinny kernel


def sec(qs, ds):
    return [((q, 0), d) for q, d in zip(qs, ds)]


def run(A, B, prec, cfg, monkeypatch):
    monkeypatch.setenv("TTG_TILE_CFG", str(cfg))
    oA, oB = O.OTensor.from_flat(A), O.OTensor.from_flat(B)
    pairs = oA.pairs(oB)
    oC = (oA * oB).to_flat()
    otasks = O.last_tasks()
    dA, dB = T.DTensor.from_flat(A), T.DTensor.from_flat(B)
    plan = T.Plan(dA, dB, pairs, prec)
    dC = plan.new_output()
    dC.fill(np.nan)          # every element must be written
    plan.execute(dA, dB, dC)
    qi, off = dC.meta()
    assert np.array_equal(qi, oC.block_qi)
    assert np.array_equal(off, oC.block_offsets())
    assert np.array_equal(plan.tasks(), otasks)
    got = dC.data()
    assert np.all(np.isfinite(got))
    err = np.linalg.norm(got - oC.data) / np.linalg.norm(oC.data)
    assert err <= TOL[prec], f"cfg {cfg} prec {prec}: rel err {err:.3e}"


def env_pattern(rng):
    """L(l', w, l) * x(l, s1, s2, r): ragged sectors, a missing x block."""
    l = make_index(INWARD, "psiL", LINK, 0, sec([-4, -2, 0, 2, 4], [2, 37, 45, 19, 1]))
    w = make_index(OUTWARD, "HL", LINK, 0, sec([-2, 0, 2], [3, 5, 3]))
    s1 = make_index(INWARD, "S1", SITE, 0, sec([-1, 1], [1, 1]))
    s2 = make_index(INWARD, "S2", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "psiR", LINK, 0, sec([-4, -2, 0, 2, 4], [3, 21, 50, 33, 2]))
    lp = l.copy()
    lp.level = 1
    return random_tensor([lp, w, l.dag()], rng), random_tensor([l, s1, s2, r], rng, keep=0.9)


def mpo_pattern(rng):
    """T1(l', w, s1, s2, r) * W(w, s1', s1, s2', s2, w'): K and N are MPO
    sector dims (<= 9), the skinny kernel's case."""
    l = make_index(INWARD, "psiL", LINK, 1, sec([-2, 0, 2], [30, 70, 41]))
    w = make_index(OUTWARD, "HL", LINK, 0, sec([-2, 0, 2], [7, 9, 7]))
    s1 = make_index(INWARD, "S1", SITE, 0, sec([-1, 1], [1, 1]))
    s2 = make_index(INWARD, "S2", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "psiR", LINK, 0, sec([-2, 0, 2], [25, 66, 40]))
    wr = make_index(OUTWARD, "HR", LINK, 0, sec([-2, 0, 2], [7, 9, 7]))
    s1p, s2p = s1.dag(), s2.dag()
    s1p.level = 1
    s2p.level = 1
    T1 = random_tensor([l, w, s1, s2, r], rng)
    W = random_tensor([w.dag(), s1p, s1.dag(), s2p, s2.dag(), wr], rng, keep=0.7)
    return T1, W


def long_k(rng):
    """Few output tiles, long contraction walk: split-K's case."""
    i = make_index(INWARD, "i", LINK, 0, sec([0, 2], [40, 17]))
    k = make_index(OUTWARD, "k", LINK, 0, sec([0, 2], [3000, 1100]))
    j = make_index(OUTWARD, "j", LINK, 0, sec([0, 2], [33, 70]))
    # A(i, k): flux i - k = 0; B(k, j) with k In, j Out
    A = random_tensor([i, k], rng)
    B = random_tensor([k.dag(), j], rng)
    return A, B


@pytest.mark.parametrize("cfg", range(NCFG))
@pytest.mark.parametrize("prec", [T.FP64, T.FP32, T.BF16X3])
@pytest.mark.parametrize("make", [env_pattern, mpo_pattern, long_k])
def test_tile_shapes(make, prec, cfg, monkeypatch):
    rng = np.random.default_rng(11)
    A, B = make(rng)
    run(A, B, prec, cfg, monkeypatch)
