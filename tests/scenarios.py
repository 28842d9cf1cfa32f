"""Test scenarios: the package's seeded workloads (ttgpu.workloads) with the
MPO tensors taken from the oracle's AutoMPO restatement, plus small chains."""
from __future__ import annotations

from ttgpu.workloads import (fill, heff_inputs, link_index, sector_dims,  # noqa: F401
                             site_index)
from ttgpu import workloads as _w

# name: (mpo kind, bond site, m, smax, sigma) for the parity tests
WORKLOADS = {
    "cfg1_j1j2_chain_m512": ("j1j2_chain", 49, 512, 20, 6.0),
    "cfg2_cyl6x12_m2048": ("cyl6x12", 35, 2048, 30, 9.0),
    "cfg4_j1j2_cyl8x16_m8192": ("j1j2_cyl8x16", 63, 8192, 40, 12.0),
}


def mpo_bond(kind: str, site: int):
    """(W_site, W_site+1) from the oracle's AutoMPO restatement."""
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


def two_site_case(seed, zero_sector=None):
    """A two-site bond: site tensors A = (l, s1, mid Out), B = (mid In, s2, r)
    as svd_bond's templates, and a random two-site tensor x = (l, s1, s2, r)
    over all legal blocks.  zero_sector: a left flux (q0 of l + s1) whose
    blocks are set to zero, giving exactly zero singular values."""
    import numpy as np
    from ttgpu.flat import INWARD, LINK, OUTWARD, SITE, make_index, random_tensor

    def sec(qs, ds):
        return [((q, 0), d) for q, d in zip(qs, ds)]

    rng = np.random.default_rng(seed)
    l = make_index(INWARD, "ID1Link3", LINK, 0, sec([-4, -2, 0, 2, 4], [2, 4, 6, 4, 2]))
    s1 = make_index(INWARD, "Site3", SITE, 0, sec([-1, 1], [1, 1]))
    mo = make_index(OUTWARD, "ID1Link4", LINK, 0, sec([-5, -3, -1, 1, 3, 5], [1, 3, 6, 6, 3, 1]))
    s2 = make_index(INWARD, "Site4", SITE, 0, sec([-1, 1], [1, 1]))
    r = make_index(OUTWARD, "ID1Link5", LINK, 0, sec([-4, -2, 0, 2, 4], [1, 3, 5, 3, 1]))
    A = random_tensor([l, s1, mo], rng)
    B = random_tensor([mo.dag(), s2, r], rng)
    x = random_tensor([l, s1, s2, r], rng)
    if zero_sector is not None:
        off = x.block_offsets()
        for b in range(x.nblocks):
            q = l.qn[x.block_qi[b, 0]][0] + s1.qn[x.block_qi[b, 1]][0]
            if q == zero_sector:
                x.data[off[b]:off[b + 1]] = 0.0
    return A, B, x
