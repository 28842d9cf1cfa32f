"""Sector SVD (svd_bond, qtensor_op.cpp:1097-1544) on sectors larger than
the single-workgroup Jacobi handles (above 64 rows), against the oracle:
kept dimension and truncation error, the singular values (to the Gram
eigensolver's resolution, about sqrt(eps) s_max), the truncated product U*V
and the isometry of the unabsorbed factor.  cutoff > 0 runs the LR-reduced
eigensolver (pivoted Cholesky, Jacobi on R R^T); cutoff = 0 keeps zero
singular values and runs the block Jacobi on the Gram matrix itself."""
import numpy as np
import pytest

import oracle as O
from ttgpu import host as P
from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor

pytestmark = pytest.mark.gpu


def sec(qs, ds):
    return [((q, 0), d) for q, d in zip(qs, ds)]


def big_two_site(seed):
    rng = np.random.default_rng(seed)
    l = make_index(INWARD, "ID1Link3", LINK, 0, sec([-2, 0, 2], [70, 150, 90]))
    s1 = make_index(INWARD, "Site3", SITE, 0, sec([-1, 1], [1, 1]))
    mo = make_index(OUTWARD, "ID1Link4", LINK, 0, sec([-3, -1, 1, 3], [60, 200, 230, 80]))
    s2 = make_index(INWARD, "Site4", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "ID1Link5", LINK, 0, sec([-2, 0, 2], [110, 160, 75]))
    A = random_tensor([l, s1, mo], rng)
    B = random_tensor([mo.dag(), s2, r], rng)
    # a decaying spectrum: scale B's blocks so that x = A * B has singular
    # values over many orders of magnitude (as a DMRG two-site state does)
    off = 0
    for b in range(B.nblocks):
        n = int(B.block_sizes()[b])
        d = B.block(b)
        k = np.arange(d.shape[0])[:, None, None] if d.ndim == 3 else 0
        B.data[off:off + n] = (d * np.exp(-k / 12.0)).reshape(-1, order="F")
        off += n
    return A, B


@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("cutoff,K", [(1e-12, 0), (1e-9, 300), (0.0, 0), (0.0, 250)])
def test_block_jacobi_svd_matches_oracle(direction, cutoff, K):
    A, B = big_two_site(3)
    oA, oB = O.OTensor.from_flat(A), O.OTensor.from_flat(B)
    pA, pB = P.PTensor.from_flat(A), P.PTensor.from_flat(B)
    ox, px = oA * oB, pA * pB
    otr, okept, oS = O.svd_bond(ox, oA, oB, "ID1Link4", direction, cutoff, K)
    ptr, pkept, pS = P.svd_bond(px, pA, pB, "ID1Link4", direction, cutoff, K)
    assert pkept == okept
    assert ptr == pytest.approx(otr, rel=1e-6, abs=1e-12)
    so = np.sort(oS.to_flat().data)
    sp_ = np.sort(pS.to_flat().data)
    assert np.allclose(sp_, so, rtol=1e-9, atol=3e-8 * so.max())
    rec_p = (pA * pB).to_flat().to_dense()
    rec_o = (oA * oB).to_flat().to_dense()
    # cutoff > 0: the LR eigensolver converges with the relative test, so the
    # kept subspace is exact to ~1e-9.  cutoff = 0 keeps the numerical null
    # space; the Gram-matrix Jacobi leaves couplings below 1e-14 of the largest
    # diagonal unrotated, which bounds the kept subspace at the Gram
    # resolution (~1e-8 relative)
    rtol = 1e-9 if cutoff > 0 else 1e-8
    assert np.linalg.norm(rec_p - rec_o) <= rtol * np.linalg.norm(rec_o)
    iso = pA if direction == 0 else pB
    D = iso.to_flat().to_dense()
    if direction == 0:
        M = D.reshape(-1, D.shape[-1], order="F")
        G = M.T @ M
    else:
        M = D.reshape(D.shape[0], -1, order="F")
        G = M @ M.T
    assert np.allclose(G, np.eye(G.shape[0]), atol=1e-11)
