"""GPU parity of the effective-Hamiltonian matvec at the batched workloads'
full sizes (big_qtensor::product, big_qtensor.cpp:23-40; each stage one
qtensor::operator*, qtensor.cpp:523-734), against the oracle:

* the 6x12 Heisenberg cylinder's two-site operator at m = 2048 (Sz sectors
  -30..30, Gaussian sigma 9, AutoMPO bond 35-36);
* the 8x16 J1-J2 cylinder's operator at m = 8192 (Sz sectors -40..40,
  Gaussian sigma 12, AutoMPO bond 63-64).

For each stage of the device operator: block list, packed offsets and the
(c, a, b, M, N, K) task list equal the oracle's bit for bit.  Values: the
product y is compared block by block, matched by sector tuple (no dense
form at m = 8192), FP64 within 1e-12 relative Frobenius error; FP32 and
split-BF16 within the bounds of TOL below."""
import numpy as np
import pytest

import oracle as O
from scenarios import WORKLOADS, heff_inputs, mpo_bond
from ttgpu import lib as T

pytestmark = pytest.mark.gpu

# relative Frobenius error of one matvec against the FP64 oracle: FP64 is
# the reference's arithmetic; FP32 (24-bit mantissa, K up to ~3000 terms)
# and split-BF16 (hi+lo bf16 operands, ~16-bit products, FP32 accumulation)
TOL = {T.FP64: 1e-12, T.FP32: 5e-6, T.BF16X3: 5e-5}

_cases = {}


def oracle_case(name):
    """Inputs, the oracle's three stage outputs and task lists (cached per
    module: the m = 8192 oracle run takes seconds)."""
    if name in _cases:
        return _cases[name]
    kind, site, m, smax, sigma = WORKLOADS[name]
    W1, W2 = mpo_bond(kind, site)
    L, R, x = heff_inputs(W1, W2, site, m, smax, sigma, seed=7)
    oL, oR, ox = (O.OTensor.from_flat(t) for t in (L, R, x))
    big = O.OBig(oL, O.OTensor.from_flat(W1), O.OTensor.from_flat(W2), oR)
    W12 = big.mid()
    pairs = [oL.pairs(ox)]
    T1 = oL * ox
    tasks = [O.last_tasks()]
    pairs.append(T1.pairs(W12))
    T2 = T1 * W12
    tasks.append(O.last_tasks())
    pairs.append(T2.pairs(oR))
    T3 = T2 * oR
    tasks.append(O.last_tasks())
    outs = []
    for t in (T1, T2):   # structures only
        f = t.to_flat()
        outs.append((f.idx, f.block_qi, f.block_offsets()))
    del T1, T2
    y = T3.to_flat()
    outs.append((y.idx, y.block_qi, y.block_offsets()))
    c = dict(L=L, R=R, x=x, W12=W12.to_flat(), pairs=pairs, tasks=tasks, outs=outs, y=y)
    _cases[name] = c
    return c


def blockwise_rel_err(got, ref):
    """||got - ref|| / ||ref|| over blocks matched by sector tuple; blocks of
    `got` that `ref` does not hold count against it in full."""
    pos = {tuple(q): b for b, q in enumerate(got.block_qi.tolist())}
    go, ro = got.block_offsets(), ref.block_offsets()
    seen = np.zeros(got.nblocks, bool)
    num = den = 0.0
    for b, q in enumerate(ref.block_qi.tolist()):
        r = ref.data[ro[b]:ro[b + 1]]
        den += float(r @ r)
        g = pos.get(tuple(q))
        assert g is not None, f"block {q} of the oracle's product missing on the device"
        seen[g] = True
        d = got.data[go[g]:go[g + 1]] - r
        num += float(d @ d)
    for g in np.flatnonzero(~seen):
        d = got.data[go[g]:go[g + 1]]
        num += float(d @ d)
    return np.sqrt(num / den)


def device_heff(c, prec):
    dL, dR, dx = (T.DTensor.from_flat(c[k]) for k in ("L", "R", "x"))
    dW = T.DTensor.from_flat(c["W12"])
    h = T.Heff(dL, dW, dR, dx, c["pairs"], prec)
    return h, dL, dW, dR, dx


@pytest.mark.parametrize("name", ["cfg2_cyl6x12_m2048", "cfg4_j1j2_cyl8x16_m8192"])
def test_heff_stage_bookkeeping_bit_exact(name):
    c = oracle_case(name)
    h, dL, dW, dR, dx = device_heff(c, T.FP64)
    lhs = [dL.idx, c["outs"][0][0], c["outs"][1][0]]
    rhs = [dx.idx, dW.idx, dR.idx]
    for s in range(3):
        p = h.stage_plan(s, lhs[s], rhs[s], c["pairs"][s])
        nb, nt, fl = p.info()
        oidx, oqi, ooff = c["outs"][s]
        assert nb == len(oqi), f"stage {s + 1}: block count"
        assert nt == len(c["tasks"][s]), f"stage {s + 1}: task count"
        assert np.array_equal(p.tasks(), c["tasks"][s]), f"stage {s + 1}: task list order"
        out = p.new_output()
        qi, off = out.meta()
        assert np.array_equal(qi, oqi), f"stage {s + 1}: result block list"
        assert np.array_equal(off, ooff), f"stage {s + 1}: packed offsets"
        t = c["tasks"][s].astype(np.float64)
        assert fl == pytest.approx(float(np.sum(2.0 * t[:, 3] * t[:, 4] * t[:, 5])), rel=1e-15)
        del out, p


@pytest.mark.parametrize("name", ["cfg2_cyl6x12_m2048", "cfg4_j1j2_cyl8x16_m8192"])
@pytest.mark.parametrize("prec", [T.FP64, T.FP32, T.BF16X3])
def test_heff_values_vs_oracle(name, prec):
    c = oracle_case(name)
    h, dL, dW, dR, dx = device_heff(c, prec)
    y = h.new_output()
    h.apply(dx, y)
    err = blockwise_rel_err(y.to_flat(), c["y"])
    print(f"{name} prec {prec}: matvec rel err vs oracle {err:.3e}", flush=True)
    assert err <= TOL[prec], err
    # a second apply through the graph gives the same bits
    y2 = h.new_output()
    h.apply(dx, y2)
    assert np.array_equal(y2.data(), y.data())
