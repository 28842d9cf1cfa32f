"""GPU parity of qtensor::initBlock (qtensor.cpp:302-345) on the device
(ttg_tensor_init_blocks): the enumerated block list (index 0 fastest, zero
total flux with Inward +, Outward -), the packed offsets and the zeroed data
equal the brute-force enumeration (ttgpu.flat.legal_blocks) and the oracle's
initBlock bit for bit, for the index sets of the hot path at full size and
for ragged and empty cases."""
import numpy as np
import pytest

import oracle as O
from scenarios import link_index, sector_dims, site_index
from ttgpu import lib as T
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, FlatTensor, legal_blocks, make_index

pytestmark = pytest.mark.gpu


def check_init(idx):
    d = T.DTensor.init_blocks(idx)
    qi, off = d.meta()
    ref = legal_blocks(idx)
    assert np.array_equal(qi, ref), "device block list differs from the brute-force enumeration"
    f = FlatTensor(idx, ref, np.zeros(0))
    assert np.array_equal(off, f.block_offsets())
    o = O.init_block(idx).to_flat()
    assert np.array_equal(o.block_qi, qi), "device block list differs from the oracle's initBlock"
    assert np.array_equal(o.block_offsets(), off)
    data = d.data()
    assert len(data) == off[-1] and not np.any(data)
    return qi


def sec(qs, ds):
    return [((q, 0), d) for q, d in zip(qs, ds)]


def test_init_two_site_vector_m2048():
    # the two-site Davidson vector of the 6x12 cylinder operator (m = 2048,
    # Sz sectors -30..30, Gaussian sigma 9): (l In, s In, s In, r Out)
    secs = sector_dims(2048, 30, 9.0)
    l = link_index(INWARD, "ID1Link36", 0, secs)
    r = link_index(OUTWARD, "ID1Link38", 0, secs)
    qi = check_init([l, site_index(INWARD, "Site36", 0), site_index(INWARD, "Site37", 0), r])
    assert len(qi) > 100


def test_init_environment_and_mixed_arrows():
    rng = np.random.default_rng(5)
    l = make_index(INWARD, "l", LINK, 1, sec([-6, -4, -2, 0, 2, 4, 6], rng.integers(1, 9, 7)))
    w = make_index(OUTWARD, "w", LINK, 0, sec([-2, 0, 2], [3, 8, 3]))
    r = make_index(OUTWARD, "r", LINK, 0, sec([-5, -3, -1, 1, 3, 5], rng.integers(1, 9, 6)))
    check_init([l, w, l.dag()])          # environment shape (l' In, w Out, l Out)
    check_init([l, w, r])                # odd/even fluxes: no legal block
    s = make_index(INWARD, "s", SITE, 0, sec([-1, 1], [1, 1]))
    check_init([l, s, s.dag(), w.dag(), r])


def test_init_two_qn_components_and_single_index():
    a = make_index(INWARD, "a", LINK, 0, [((q0, q1), d) for (q0, q1), d in
                                          zip([(-1, 0), (0, -1), (0, 1), (1, 0)], [2, 3, 1, 4])])
    b = make_index(OUTWARD, "b", LINK, 0, [((q0, q1), d) for (q0, q1), d in
                                           zip([(-1, 0), (0, 1), (1, 0), (2, 2)], [5, 1, 2, 7])])
    check_init([a, b])
    check_init([make_index(INWARD, "z", LINK, 0, sec([-2, 0, 3], [2, 5, 1]))])


def test_init_no_legal_block():
    a = make_index(INWARD, "a", LINK, 0, sec([1, 3], [2, 2]))
    b = make_index(OUTWARD, "b", LINK, 0, sec([0, 2], [2, 2]))
    qi = check_init([a, b])
    assert len(qi) == 0
