"""Sector SVD (svd_bond, qtensor_op.cpp:1506-1544) on a state whose Gram
matrices are not finite: entries near 1e160 (their squares overflow FP64) or
a NaN.  svd_bond must fail with an error and leave the library usable; the
same bond with finite entries then matches the oracle."""
import numpy as np
import pytest

import oracle as O
from ttgpu import host as P
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor

pytestmark = pytest.mark.gpu


def sec(qs, ds):
    return [((q, 0), d) for q, d in zip(qs, ds)]


def two_site(seed):
    rng = np.random.default_rng(seed)
    l = make_index(INWARD, "ID1Link3", LINK, 0, sec([-2, 0, 2], [20, 90, 30]))
    s1 = make_index(INWARD, "Site3", SITE, 0, sec([-1, 1], [1, 1]))
    mo = make_index(OUTWARD, "ID1Link4", LINK, 0, sec([-3, -1, 1, 3], [10, 70, 80, 15]))
    s2 = make_index(INWARD, "Site4", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "ID1Link5", LINK, 0, sec([-2, 0, 2], [25, 85, 20]))
    return random_tensor([l, s1, mo], rng), random_tensor([mo.dag(), s2, r], rng)


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("cutoff", [1e-9, 0.0])
def test_nonfinite_gram_is_an_error(direction, cutoff):
    A, B = two_site(4)
    for case in ("overflow", "nan"):
        Ab = two_site(4)[0]
        if case == "overflow":
            Ab.data = Ab.data * 1e160
        else:
            Ab.data = Ab.data.copy()
            Ab.data[len(Ab.data) // 2] = np.nan
        pA, pB = P.PTensor.from_flat(Ab), P.PTensor.from_flat(B)
        with pytest.raises(RuntimeError, match="not finite"):
            P.svd_bond(pA * pB, pA, pB, "ID1Link4", direction, cutoff, 0)
    # the library is intact: the finite bond matches the oracle
    oA, oB = O.OTensor.from_flat(A), O.OTensor.from_flat(B)
    pA, pB = P.PTensor.from_flat(A), P.PTensor.from_flat(B)
    otr, okept, oS = O.svd_bond(oA * oB, oA, oB, "ID1Link4", direction, cutoff, 0)
    ptr, pkept, pS = P.svd_bond(pA * pB, pA, pB, "ID1Link4", direction, cutoff, 0)
    assert pkept == okept
    so = np.sort(oS.to_flat().data)
    assert np.allclose(np.sort(pS.to_flat().data), so, rtol=1e-9, atol=3e-8 * so.max())
