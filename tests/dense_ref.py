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
