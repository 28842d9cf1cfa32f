"""Dense numpy references used to pin the oracle (closed forms, no reference code)."""
import numpy as np

from ttgpu.flat import FlatTensor

SZ = np.array([[-0.5, 0.0], [0.0, 0.5]])   # basis order Dn(0), Up(1)
SP = np.array([[0.0, 0.0], [1.0, 0.0]])    # <Up|S+|Dn> = 1 ; rows = bra
SM = SP.T.copy()
ID = np.eye(2)


def kron_chain(ops):
    out = np.ones((1, 1))
    for o in ops:
        out = np.kron(out, o)
    return out


def site_op(N, i, op):
    ops = [ID] * N
    ops[i] = op
    return ops


def heisenberg_dense(N, bonds, J):
    """H = sum_b J_b (Sz Sz + 1/2 (S+S- + S-S+)); site 0 is the most significant factor."""
    H = np.zeros((2 ** N, 2 ** N))
    for (a, b), j in zip(bonds, J):
        for A, B, c in ((SZ, SZ, 1.0), (SP, SM, 0.5), (SM, SP, 0.5)):
            ops = [ID] * N
            ops[a] = A
            ops[b] = B
            H += j * c * kron_chain(ops)
    return H


def mpo_to_dense(Ws):
    """Contract a list of dense MPO tensors W[l, ket, bra, r] into H[bra, ket]
    with site 0 the most significant factor."""
    T = Ws[0][0]            # (ket, bra, r)
    T = T.transpose(1, 0, 2)  # (bra, ket, r)
    for W in Ws[1:]:
        # T (BRA, KET, l) x W (l, ket, bra, r) -> (BRA, bra, KET, ket, r)
        T = np.einsum("BKl,lkbr->BbKkr", T, W)
        s = T.shape
        T = T.reshape(s[0] * s[1], s[2] * s[3], s[4])
    return T[:, :, 0]


def einsum_contract(A: FlatTensor, B: FlatTensor, pairs):
    """Dense contraction of densified tensors over the given (posA, posB) pairs.
    Output index order: A free (A order) then B free (B order)."""
    DA, DB = A.to_dense(), B.to_dense()
    letters = "abcdefghijklmnopqrstuvwxyz"
    la = list(letters[:A.rank])
    lb = list(letters[A.rank:A.rank + B.rank])
    for pa, pb in pairs:
        lb[pb] = la[pa]
    paired_a = {p for p, _ in pairs}
    paired_b = {p for _, p in pairs}
    out = [la[i] for i in range(A.rank) if i not in paired_a] + \
          [lb[i] for i in range(B.rank) if i not in paired_b]
    return np.einsum("".join(la) + "," + "".join(lb) + "->" + "".join(out), DA, DB)


# ---------------------------------------------------------------- truncation
def sector_singular_values(x: FlatTensor, nleft: int):
    """{mid QN: singular values (descending)} of the matrix whose rows are
    x's first nleft indices and whose columns are the rest, one block per
    sector of the left flux (Inward +, Outward -), from numpy's SVD of the
    densified sector matrices.  Only sectors that hold blocks of x count."""
    D = x.to_dense()
    rows = int(np.prod(D.shape[:nleft]))
    M = D.reshape(rows, -1)

    def axis_flux(ix, sign):
        f0 = np.concatenate([[sign * q[0]] * d for q, d in zip(ix.qn, ix.qdim)])
        f1 = np.concatenate([[sign * q[1]] * d for q, d in zip(ix.qn, ix.qdim)])
        return f0, f1

    def flux(idx, signs):
        t0, t1 = np.zeros(1, np.int64), np.zeros(1, np.int64)
        for ix, sg in zip(idx, signs):
            a0, a1 = axis_flux(ix, sg)
            t0 = np.add.outer(t0, a0).ravel()
            t1 = np.add.outer(t1, a1).ravel()
        return t0, t1

    r0, r1 = flux(x.idx[:nleft], [1 if ix.arrow == 0 else -1 for ix in x.idx[:nleft]])
    c0, c1 = flux(x.idx[nleft:], [-1 if ix.arrow == 0 else 1 for ix in x.idx[nleft:]])
    keys = set()
    for b in range(x.nblocks):
        q0 = q1 = 0
        for j in range(nleft):
            ix, s = x.idx[j], int(x.block_qi[b, j])
            sg = 1 if ix.arrow == 0 else -1
            q0 += sg * ix.qn[s][0]
            q1 += sg * ix.qn[s][1]
        keys.add((q0, q1))
    out = {}
    for q in sorted(keys):
        ri = np.nonzero((r0 == q[0]) & (r1 == q[1]))[0]
        ci = np.nonzero((c0 == q[0]) & (c1 == q[1]))[0]
        out[q] = np.linalg.svd(M[np.ix_(ri, ci)], compute_uv=False)
    return out


def check_keep_rule(svals, cutoff, K, kept, trunc, kept_per_sector=None, zero_tol=0.0):
    """Asserts what the reference's truncation rule (qtensor_op.cpp:1289-1301)
    fixes, stated as invariants of the globally sorted singular values s
    (descending, nm = sum s^2, tr(k) = 1 - sum_{i<k} s_i^2 / nm):
      cutoff = 0, K = 0: every value is kept;
      K > 0 binding:     kept = K;
      cutoff > 0:        a zero value is never kept, and if the cutoff ends
                         the loop, tr(kept) < cutoff <= tr(kept - 1);
      the reported truncation error is max(tr(kept), 0) (to rounding);
      per sector, the kept count is the number of that sector's values among
      the kept ones (values are assumed distinct across the cut).
    Values at or below zero_tol count as zero."""
    allv = np.sort(np.concatenate([v for v in svals.values()]))[::-1]
    nm = float(np.sum(allv ** 2))
    tr = lambda k: 1.0 - float(np.sum(allv[:k] ** 2)) / nm   # noqa: E731
    nz = int(np.sum(allv > zero_tol))
    total = len(allv)
    if cutoff == 0 and K == 0:
        assert kept == total, (kept, total)
    else:
        limit = total if cutoff == 0 else nz
        if K > 0:
            limit = min(limit, K)
        assert 1 <= kept <= limit, (kept, limit)
        if kept < limit:   # the cutoff ended the loop
            assert cutoff > 0 and tr(kept) < cutoff + 1e-13 and tr(kept - 1) >= cutoff - 1e-13, \
                (kept, tr(kept), tr(kept - 1), cutoff)
        if cutoff > 0:
            assert allv[kept - 1] > zero_tol
    assert abs(trunc - max(tr(kept), 0.0)) < 1e-12, (trunc, tr(kept))
    if kept_per_sector is not None:
        thr = allv[kept - 1]
        for q, v in svals.items():
            want = int(np.sum(v >= thr))
            assert kept_per_sector.get(q, 0) == want, (q, kept_per_sector.get(q, 0), want)
