"""Device pool limit: a tensor larger than the pool's budget (the device's
free memory less a reserve for the HIP runtime) is refused with an error
before any device allocation is attempted, and the library keeps working."""
import numpy as np
import pytest

from ttgpu import lib as T
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor

pytestmark = pytest.mark.gpu


def test_allocation_past_the_budget_is_refused():
    # 16 diagonal blocks of 46340^2 doubles (each below 2^31 elements):
    # 275 GB in all, above the budget of a 288 GB device less its reserve
    secs = [((q, 0), 46340) for q in range(-15, 17, 2)]
    a = make_index(INWARD, "a", LINK, 0, secs)
    b = make_index(OUTWARD, "b", LINK, 0, secs)
    with pytest.raises(RuntimeError, match="refused"):
        T.DTensor.init_blocks([a, b])
    rng = np.random.default_rng(2)
    l = make_index(INWARD, "l", LINK, 0, [((q, 0), d) for q, d in zip([-1, 1], [5, 7])])
    s = make_index(INWARD, "s", SITE, 0, [((-1, 0), 1), ((1, 0), 1)])
    r = make_index(OUTWARD, "r", LINK, 0, [((q, 0), d) for q, d in zip([-2, 0, 2], [3, 6, 4])])
    f = random_tensor([l, s, r], rng)
    x = T.DTensor.from_flat(f)
    assert x.norm() == pytest.approx(np.linalg.norm(f.data), rel=1e-14)
