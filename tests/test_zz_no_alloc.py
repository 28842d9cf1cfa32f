"""The allocation-free embed (ttg_embed_noalloc): its scratch is the
library's fixed device array, looked up at ttg_init, so the call makes no
device allocation.  Checked against ttg_embed (the allocating entry point it
mirrors) bit for bit, on two calls in a row (the second with fewer blocks,
so stale scratch from the first call must not leak into it), and on the
refusals: more blocks than the scratch holds, and a source block missing
from the destination."""
import ctypes as C

import numpy as np
import pytest

from ttgpu import lib as T
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor

pytestmark = pytest.mark.gpu


def sec(qs, ds):
    return [((q, 0), d) for q, d in zip(qs, ds)]


def noalloc(src, dst):
    T.embed_noalloc(src, dst)


def pool_in_use():
    v = [C.c_size_t(0) for _ in range(3)]
    T.check(T.lib().ttg_mem_info(*[C.byref(x) for x in v]))
    return v[0].value


def case(nl, seed):
    rng = np.random.default_rng(seed)
    qs = list(range(-2 * nl, 2 * nl + 1, 2))
    l = make_index(INWARD, "l", LINK, 0, sec(qs, rng.integers(1, 7, len(qs))))
    s = make_index(INWARD, "s", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "r", LINK, 0, sec([q + 1 for q in qs], rng.integers(1, 7, len(qs))))
    src = random_tensor([l, s, r], rng, keep=0.6)
    return src, T.DTensor.from_flat(src), T.DTensor.init_blocks([l, s, r])


@pytest.mark.parametrize("order", [(9, 3), (3, 9)])
def test_embed_noalloc_matches_embed(order):
    T.init()
    for k, nl in enumerate(order):
        f, src, full = case(nl, 11 + k)
        ref = full.copy()
        T.embed(src, ref)
        got = full.copy()
        got.fill(7.0)            # stale values must not survive the call
        before = pool_in_use()
        noalloc(src, got)
        assert pool_in_use() == before
        assert np.array_equal(got.data().view(np.uint64), ref.data().view(np.uint64))
        # the embedded blocks carry the source values; the others are zero
        assert np.isclose(got.norm(), np.linalg.norm(f.data), rtol=1e-14)


def test_embed_noalloc_refusals():
    T.init()
    f, src, full = case(4, 3)
    # destination lacking a source block: error after the kernels, no fault
    l, s, r = f.idx
    part = T.DTensor.from_flat(random_tensor([l, s, r], np.random.default_rng(1), keep=0.3))
    with pytest.raises(RuntimeError, match="missing"):
        noalloc(src, part)
    # more blocks than the fixed scratch holds: refused before any launch
    big = [make_index(INWARD, "a", LINK, 0, sec(range(0, 300), [1] * 300)),
           make_index(INWARD, "b", LINK, 0, sec(range(0, 300), [1] * 300)),
           make_index(OUTWARD, "c", LINK, 0, sec(range(0, 600), [1] * 600))]
    t = T.DTensor.init_blocks(big)
    assert t.info()[1] > 65536
    with pytest.raises(RuntimeError, match="65536"):
        noalloc(t, t)
    # and the entry point keeps working afterwards
    ref = full.copy()
    T.embed(src, ref)
    got = full.copy()
    noalloc(src, got)
    assert np.array_equal(got.data(), ref.data())
