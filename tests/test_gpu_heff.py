"""GPU parity of the lazy effective Hamiltonian (big_qtensor::product,
big_qtensor.cpp:23-40) against the oracle: FP64 to 1e-12 relative
Frobenius error; the reduced-precision forms to the bounds stated below."""
import numpy as np
import pytest

import oracle as O
from scenarios import heff_inputs, mpo_bond, WORKLOADS
from ttgpu import lib as T

pytestmark = pytest.mark.gpu

# relative Frobenius error bounds of one matvec
TOL = {T.FP64: 1e-12, T.FP32: 2e-5, T.BF16X3: 2e-4}


def build_case(kind, site, m, smax, sigma, seed=7):
    W1, W2 = mpo_bond(kind, site)
    L, R, x = heff_inputs(W1, W2, site, m, smax, sigma, seed)
    oL, oR, ox = (O.OTensor.from_flat(t) for t in (L, R, x))
    oW1, oW2 = O.OTensor.from_flat(W1), O.OTensor.from_flat(W2)
    big = O.OBig(oL, oW1, oW2, oR)
    W12 = big.mid()
    T1 = oL * ox
    T2 = T1 * W12
    pairs = [oL.pairs(ox), T1.pairs(W12), T2.pairs(oR)]
    return dict(L=L, R=R, x=x, W12=W12.to_flat(), big=big, ox=ox, pairs=pairs)


def run_heff(case, prec):
    dL, dR, dx = (T.DTensor.from_flat(case[k]) for k in ("L", "R", "x"))
    dW = T.DTensor.from_flat(case["W12"])
    h = T.Heff(dL, dW, dR, dx, case["pairs"], prec)
    y = h.new_output()
    h.apply(dx, y)
    return h, dx, y


def dense_y(case):
    oy = case["big"].product(case["ox"]).to_flat()
    return oy.to_dense()


@pytest.mark.parametrize("prec", [T.FP64, T.FP32, T.BF16X3])
def test_heff_small_chain(prec):
    case = build_case("nn_chain", 8, 60, 5, 2.0)
    h, dx, y = run_heff(case, prec)
    ref = dense_y(case)
    got = y.to_flat().to_dense()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= TOL[prec], err
    # repeated applies through the graph are identical
    y2 = h.new_output()
    h.apply(dx, y2)
    assert np.array_equal(y2.data(), y.data())


def test_heff_cfg1_fp64():
    kind, site, m, smax, sigma = WORKLOADS["cfg1_j1j2_chain_m512"]
    case = build_case(kind, site, m, smax, sigma)
    h, dx, y = run_heff(case, T.FP64)
    ref = dense_y(case)
    got = y.to_flat().to_dense()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-12
    fl, nt, sf = h.info()
    assert fl == pytest.approx(sf.sum())
    assert all(nt > 0)


def test_heff_linear():
    # H(a + 2b) == H a + 2 H b on the packed vectors (random environments are
    # not Hermitian, so linearity is the size-independent property here)
    case = build_case("nn_chain", 5, 40, 4, 2.0, seed=3)
    h, dx, y = run_heff(case, T.FP64)
    rng = np.random.default_rng(0)
    n = dx.info()[2]
    va, vb = rng.standard_normal(n), rng.standard_normal(n)
    a, b, c = dx.copy(), dx.copy(), dx.copy()
    a.upload(va)
    b.upload(vb)
    c.upload(va + 2 * vb)
    Ha, Hb, Hc = h.new_output(), h.new_output(), h.new_output()
    h.apply(a, Ha)
    h.apply(b, Hb)
    h.apply(c, Hc)
    lhs = Hc.data()
    rhs = Ha.data() + 2 * Hb.data()
    assert np.linalg.norm(lhs - rhs) <= 1e-13 * np.linalg.norm(rhs)


def test_small_eigh_matches_numpy():
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 4, 6):
        A = rng.standard_normal((n, n))
        A = A + A.T
        ev, V = T.small_eigh(A)
        ref = np.linalg.eigvalsh(A)
        assert np.allclose(ev, ref, atol=1e-13)
        assert np.allclose(A @ V, V * ev, atol=1e-12)


@pytest.mark.parametrize("src_prec,twin_prec", [(T.BF16X3, T.FP64), (T.FP32, T.FP64), (T.FP64, T.BF16X3)])
def test_heff_twin_matches_direct_operator(src_prec, twin_prec):
    """ttg_heff_create_twin: an operator at another precision on the source's
    plans (block and task lists copied, k-tile offsets recomputed when the
    k-tile depth differs) gives the oracle's product to the twin precision's
    bound, and each stage's block list, offsets and task tuples are those of
    the operator built directly at that precision."""
    case = build_case("j1j2_chain", 50, 160, 8, 3.0)
    h, dx, _ = run_heff(case, src_prec)
    tw = T.Heff.twin(h, twin_prec)
    y = tw.new_output()
    tw.apply(dx, y)
    ref = dense_y(case)
    err = np.linalg.norm(y.to_flat().to_dense() - ref) / np.linalg.norm(ref)
    assert err <= TOL[twin_prec], err
    direct, _, yd = run_heff(case, twin_prec)
    if twin_prec == T.FP64:
        d = np.linalg.norm(yd.to_flat().to_dense() - y.to_flat().to_dense()) / np.linalg.norm(ref)
        assert d <= 1e-13, d
    Li, xi = T.DTensor.from_flat(case["L"]).idx, dx.idx
    Wi, Ri = T.DTensor.from_flat(case["W12"]).idx, T.DTensor.from_flat(case["R"]).idx
    T1i = T.result_indices(Li, xi, case["pairs"][0])
    T2i = T.result_indices(T1i, Wi, case["pairs"][1])
    for st, (A_idx, B_idx) in enumerate([(Li, xi), (T1i, Wi), (T2i, Ri)]):
        pt = tw.stage_plan(st, A_idx, B_idx, case["pairs"][st])
        pd = direct.stage_plan(st, A_idx, B_idx, case["pairs"][st])
        assert np.array_equal(pt.tasks(), pd.tasks()), f"stage {st + 1}: task list"
        qt, ot = pt.new_output().meta()
        qd, od = pd.new_output().meta()
        assert np.array_equal(qt, qd) and np.array_equal(ot, od), f"stage {st + 1}: block list"
