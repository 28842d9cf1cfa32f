"""Seeded synthetic inputs for the parity tests and the bench.

Sector model (the matvec / sweep workloads): a link of bond dimension m is
split over Sz sectors s in [-S, S] (QN q0 = 2*Sz, q1 = 0) with dims
proportional to exp(-s^2 / (2 sigma^2)), rounded, minimum 1; the rounding
residue is then spread one unit at a time over the largest sectors (ties to
the lower Sz) until the dims sum to m exactly.

Environments L, R and the two-site vector x are filled block by block
(block order = the reference's initBlock order) with N(0,1) FP64 values from
numpy's PCG64 seeded with seed+0 (L), seed+1 (R), seed+2 (x).  W holds the
exact MPO coefficients of the reference's AutoMPO builder (oracle restatement).
"""
from __future__ import annotations

import numpy as np

from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, FlatTensor, Index, legal_blocks, make_index


def sector_dims(m: int, smax: int, sigma: float):
    s = np.arange(-smax, smax + 1)
    w = np.exp(-s.astype(float) ** 2 / (2 * sigma * sigma))
    d = np.maximum(1, np.round(m * w / w.sum())).astype(int)
    order = sorted(range(len(s)), key=lambda i: (-d[i], s[i]))
    k = 0
    while d.sum() != m:
        i = order[k % len(order)]
        if d.sum() < m:
            d[i] += 1
        elif d[i] > 1:
            d[i] -= 1
        k += 1
    return [(2 * int(q), int(dd)) for q, dd in zip(s, d)]


def link_index(arrow, name, level, sectors):
    return make_index(arrow, name, LINK, level, [((q, 0), d) for q, d in sectors])


def site_index(arrow, name, level):
    return make_index(arrow, name, SITE, level, [((-1, 0), 1), ((1, 0), 1)])


def fill(idx, seed):
    rng = np.random.default_rng(seed)
    bq = legal_blocks(idx)
    t = FlatTensor([ix.copy() for ix in idx], bq, np.zeros(0))
    t.data = rng.standard_normal(int(t.block_sizes().sum()))
    return t


def index_of(t: FlatTensor, pos: int) -> Index:
    return t.idx[pos].copy()


def heff_inputs(W1: FlatTensor, W2: FlatTensor, site: int, m: int, smax: int, sigma: float,
                seed: int, psi_id: int = 1):
    """L, R, x for the bond (site, site+1) around MPO tensors W1 = W_site,
    W2 = W_site+1 (index order: link In, site Out lvl0, site In lvl1, link Out).
    Index names follow the reference: ID<id>Link<i>, Site<i>."""
    secs = sector_dims(m, smax, sigma)
    l = link_index(INWARD, f"ID{psi_id}Link{site}", 0, secs)
    r = link_index(OUTWARD, f"ID{psi_id}Link{site + 2}", 0, secs)
    si = site_index(INWARD, f"Site{site}", 0)
    sj = site_index(INWARD, f"Site{site + 1}", 0)
    wl = W1.idx[0].dag()        # TL's MPO leg is Outward (dmrg.cpp:60-65)
    wr = W2.idx[3].dag()        # TR's MPO leg is Inward (dmrg.cpp:69-74)
    lp = l.copy(); lp.level = 1            # TL = (l' In lvl1, w Out, l Out)
    ld = l.dag()
    rp = r.copy(); rp.level = 1            # TR = (r' Out lvl1, w In, r In)
    rd = r.dag()
    L = fill([lp, wl, ld], seed)
    R = fill([rp, wr, rd], seed + 1)
    x = fill([l, si, sj, r], seed + 2)
    return L, R, x


def mpo_bond(kind: str, site: int):
    """MPO tensors (W_site, W_site+1) of a workload, from the oracle's
    AutoMPO restatement.  kind: 'j1j2_chain' (N=100, J2/J1=0.5),
    'cyl6x12' (NN Heisenberg 6x12, y-periodic), 'j1j2_cyl8x16'."""
    import oracle as O
    if kind == "j1j2_chain":
        N, terms = 100, O.heisenberg_terms(1, 100, 1, False, 1.0, 0.5)
    elif kind == "cyl6x12":
        N, terms = 72, O.heisenberg_terms(0, 12, 6, True, 1.0, 0.0)
    elif kind == "j1j2_cyl8x16":
        N, terms = 128, O.heisenberg_terms(1, 16, 8, True, 1.0, 0.5)
    elif kind == "nn_chain":
        N, terms = 20, O.heisenberg_terms(0, 20, 1, False, 1.0, 0.0)
    else:
        raise ValueError(kind)
    H = O.mpo_autompo(N, terms)
    return H.tensor(site).to_flat(), H.tensor(site + 1).to_flat()


WORKLOADS = {
    # name: (mpo kind, bond site, m, smax, sigma)
    "cfg1_j1j2_chain_m512": ("j1j2_chain", 49, 512, 20, 6.0),
    "cfg2_cyl6x12_m2048": ("cyl6x12", 35, 2048, 30, 9.0),
    "cfg4_j1j2_cyl8x16_m8192": ("j1j2_cyl8x16", 63, 8192, 40, 12.0),
}
