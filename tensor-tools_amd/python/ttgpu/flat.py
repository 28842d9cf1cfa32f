"""Flat, C-ABI-shaped description of a U(1) block-sparse tensor.

This is the host-side data format that crosses the C-ABI (include/ttgpu.h):
per-index sector lists sorted by QN, a block list of sector-position tuples,
and the block data packed in block order, each block column-major (first
index fastest) like the reference's blocks (qtensor/qtensor.h:84-92).

Arrow 0 = Inward, 1 = Outward; type 0 = Link, 1 = Site (util/types_and_headers.h:137-141).
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np

INWARD, OUTWARD = 0, 1
LINK, SITE = 0, 1


@dataclasses.dataclass
class Index:
    arrow: int
    name: str
    type: int
    level: int
    qn: List[tuple]      # [(q0, q1), ...] sorted ascending
    qdim: List[int]

    def copy(self) -> "Index":
        return Index(self.arrow, self.name, self.type, self.level, list(self.qn), list(self.qdim))

    def dag(self) -> "Index":
        c = self.copy()
        c.arrow = 1 - c.arrow
        return c


@dataclasses.dataclass
class FlatTensor:
    idx: List[Index]
    block_qi: np.ndarray            # (nblocks, rank) int32
    data: np.ndarray                # float64, blocks packed in order

    @property
    def rank(self) -> int:
        return len(self.idx)

    @property
    def nblocks(self) -> int:
        return int(self.block_qi.shape[0])

    def block_dims(self, b: int):
        return [self.idx[j].qdim[int(self.block_qi[b, j])] for j in range(self.rank)]

    def block_sizes(self) -> np.ndarray:
        if self.nblocks == 0:
            return np.zeros(0, dtype=np.int64)
        d = np.array([[ix.qdim[int(q)] for ix, q in zip(self.idx, row)] for row in self.block_qi],
                     dtype=np.int64)
        return d.prod(axis=1)

    def block_offsets(self) -> np.ndarray:
        s = self.block_sizes()
        return np.concatenate([[0], np.cumsum(s)]).astype(np.int64)

    def block(self, b: int) -> np.ndarray:
        off = self.block_offsets()
        return self.data[off[b]:off[b + 1]].reshape(self.block_dims(b), order="F")

    def block_qn_key(self, b: int):
        return tuple(self.idx[j].qn[int(self.block_qi[b, j])] for j in range(self.rank))

    # ---- flat arrays for ctypes -------------------------------------
    def c_arrays(self):
        arrows = np.array([ix.arrow for ix in self.idx], dtype=np.int32)
        levels = np.array([ix.level for ix in self.idx], dtype=np.int32)
        types = np.array([ix.type for ix in self.idx], dtype=np.int32)
        names = ("".join(ix.name + "\n" for ix in self.idx)).encode()
        nsec = np.array([len(ix.qn) for ix in self.idx], dtype=np.int32)
        qn = np.array([q for ix in self.idx for q in ix.qn], dtype=np.int32).reshape(-1)
        qdim = np.array([d for ix in self.idx for d in ix.qdim], dtype=np.int32)
        bqi = np.ascontiguousarray(self.block_qi, dtype=np.int32).reshape(-1)
        data = np.ascontiguousarray(self.data, dtype=np.float64)
        return arrows, levels, types, names, nsec, qn, qdim, bqi, data

    # ---- dense view (tests) ------------------------------------------
    def dense_shape(self):
        return [int(sum(ix.qdim)) for ix in self.idx]

    def to_dense(self) -> np.ndarray:
        """Embed blocks into a dense array (sector ranges in sorted order)."""
        starts = [np.concatenate([[0], np.cumsum(ix.qdim)]) for ix in self.idx]
        D = np.zeros(self.dense_shape())
        for b in range(self.nblocks):
            sl = tuple(slice(starts[j][q], starts[j][q] + self.idx[j].qdim[q])
                       for j, q in enumerate(self.block_qi[b]))
            D[sl] = self.block(b)
        return D


def make_index(arrow, name, type_, level, sectors) -> Index:
    """sectors: iterable of ((q0, q1), dim); sorted by QN here."""
    s = sorted(((tuple(q), int(d)) for q, d in sectors), key=lambda t: t[0])
    return Index(arrow, name, type_, level, [q for q, _ in s], [d for _, d in s])


def legal_blocks(idx: List[Index]) -> np.ndarray:
    """All zero-flux sector tuples, index 0 fastest (qtensor.cpp:302-345)."""
    sizes = [len(ix.qn) for ix in idx]
    n = int(np.prod(sizes)) if sizes else 0
    out = []
    for i in range(n):
        pos, tot0, tot1, r = [], 0, 0, i
        for j, ix in enumerate(idx):
            p = r % sizes[j]
            r //= sizes[j]
            pos.append(p)
            q = ix.qn[p]
            sgn = 1 if ix.arrow == INWARD else -1
            tot0 += sgn * q[0]
            tot1 += sgn * q[1]
        if tot0 == 0 and tot1 == 0:
            out.append(pos)
    return np.array(out, dtype=np.int32).reshape(-1, len(idx))


def random_tensor(idx: List[Index], rng: np.random.Generator, block_qi: Optional[np.ndarray] = None,
                  keep: float = 1.0) -> FlatTensor:
    """Random block-sparse tensor over legal blocks (optionally a random subset)."""
    bq = legal_blocks(idx) if block_qi is None else block_qi
    if keep < 1.0 and len(bq):
        mask = rng.random(len(bq)) < keep
        mask[0] = True
        bq = bq[mask]
    t = FlatTensor([ix.copy() for ix in idx], bq, np.zeros(0))
    t.data = rng.standard_normal(int(t.block_sizes().sum()))
    return t
