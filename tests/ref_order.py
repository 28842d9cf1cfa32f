"""Independent numpy statement of the bookkeeping order of
qtensor::operator* (qtensor.cpp:606-722), used to pin the oracle's and the
device's block lists and task lists.

For C = A * B with contracted pairs (pa, pb) in the reference's pairing
order, the free positions keep their tensor order.  A block's sector key is
the flux of its free indices (A: Inward +, Outward -; B: Inward -,
Outward +).  Result blocks run over the keys A holds, ascending, then B's
free-position tuples of that key in lexicographic order, then A's free
tuples in lexicographic order; every such (left, right) pair is a result
block, whether or not a product reaches it.  Inside a result block the
contributions run over the contracted tuples B holds for that key, in
lexicographic order, and a task is emitted when both the A block and the B
block of the tuple exist."""
from collections import defaultdict

import numpy as np

from ttgpu.flat import INWARD, FlatTensor


def _flux(t: FlatTensor, qi, positions, sign_in):
    q0 = q1 = 0
    for p in positions:
        s = sign_in if t.idx[p].arrow == INWARD else -sign_in
        q = t.idx[p].qn[int(qi[p])]
        q0 += s * q[0]
        q1 += s * q[1]
    return (q0, q1)


def contraction_order(A: FlatTensor, B: FlatTensor, pairs):
    """(result block tuples, tasks) where tasks are (c, a, b, M, N, K) rows."""
    pa = [p for p, _ in pairs]
    pb = [q for _, q in pairs]
    fa = [j for j in range(A.rank) if j not in pa]
    fb = [j for j in range(B.rank) if j not in pb]
    akey = {tuple(int(v) for v in qi): b for b, qi in enumerate(A.block_qi)}
    bkey = {tuple(int(v) for v in qi): b for b, qi in enumerate(B.block_qi)}
    left, right, mid = defaultdict(set), defaultdict(set), defaultdict(set)
    for qi in A.block_qi:
        left[_flux(A, qi, fa, 1)].add(tuple(int(qi[j]) for j in fa))
    for qi in B.block_qi:
        q = _flux(B, qi, fb, -1)
        right[q].add(tuple(int(qi[j]) for j in fb))
        mid[q].add(tuple(int(qi[j]) for j in pb))
    blocks, tasks = [], []
    for q in sorted(left):
        for r in sorted(right.get(q, ())):
            for lt in sorted(left[q]):
                c = len(blocks)
                blocks.append(list(lt) + list(r))
                M = int(np.prod([A.idx[j].qdim[v] for j, v in zip(fa, lt)])) if fa else 1
                N = int(np.prod([B.idx[j].qdim[v] for j, v in zip(fb, r)])) if fb else 1
                for m in sorted(mid.get(q, ())):
                    at = [0] * A.rank
                    for j, v in zip(fa, lt):
                        at[j] = v
                    for j, v in zip(pa, m):
                        at[j] = v
                    bt = [0] * B.rank
                    for j, v in zip(pb, m):
                        bt[j] = v
                    for j, v in zip(fb, r):
                        bt[j] = v
                    a, b = akey.get(tuple(at)), bkey.get(tuple(bt))
                    if a is None or b is None:
                        continue
                    K = int(np.prod([B.idx[j].qdim[v] for j, v in zip(pb, m)])) if pb else 1
                    tasks.append((c, a, b, M, N, K))
    return (np.array(blocks, np.int32).reshape(-1, len(fa) + len(fb)),
            np.array(tasks, np.int64).reshape(-1, 6))
