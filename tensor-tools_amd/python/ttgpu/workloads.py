"""Seeded synthetic effective-Hamiltonian workloads (the matvec benchmarks).

Sector model: a link of bond dimension m is split over Sz sectors s in
[-S, S] (QN q0 = 2*Sz, q1 = 0) with dims proportional to
exp(-s^2 / (2 sigma^2)), rounded, minimum 1; the rounding residue is spread
one unit at a time over the largest sectors (ties to the lower Sz) until the
dims sum to m exactly.

L, R and the two-site vector x are filled block by block (block order = the
reference's initBlock order) with N(0,1) FP64 values from numpy's PCG64
seeded with seed+0 (L), seed+1 (R), seed+2 (x).  W holds the exact MPO
coefficients from the AutoMPO builder of the host layer.
"""
from __future__ import annotations

import numpy as np

from .flat import INWARD, LINK, OUTWARD, SITE, FlatTensor, legal_blocks, make_index

# name: (lattice, Nx, Ny, y-periodic, J2, bond site, m, S, sigma)
WORKLOADS = {
    "j1j2_chain_N100_m512": ("j1j2", 100, 1, False, 0.5, 49, 512, 20, 6.0),
    "heisenberg_cyl6x12_m2048": ("nn", 12, 6, True, 0.0, 35, 2048, 30, 9.0),
    "j1j2_cyl8x16_m8192": ("j1j2", 16, 8, True, 0.5, 63, 8192, 40, 12.0),
}


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


def heff_inputs(W1: FlatTensor, W2: FlatTensor, site: int, m: int, smax: int, sigma: float,
                seed: int, psi_id: int = 1):
    """L, R, x for the bond (site, site+1) around MPO tensors W1 = W_site and
    W2 = W_site+1 (index order: link In, site Out lvl0, site In lvl1, link
    Out).  Index names follow the reference: ID<id>Link<i>, Site<i>; the
    environments have the reference's layouts TL = (l' In lvl1, w Out, l Out)
    and TR = (r' Out lvl1, w In, r In) (dmrg.cpp:60-74)."""
    secs = sector_dims(m, smax, sigma)
    l = link_index(INWARD, f"ID{psi_id}Link{site}", 0, secs)
    r = link_index(OUTWARD, f"ID{psi_id}Link{site + 2}", 0, secs)
    si = site_index(INWARD, f"Site{site}", 0)
    sj = site_index(INWARD, f"Site{site + 1}", 0)
    wl = W1.idx[0].dag()
    wr = W2.idx[3].dag()
    lp = l.copy()
    lp.level = 1
    rp = r.copy()
    rp.level = 1
    L = fill([lp, wl, l.dag()], seed)
    R = fill([rp, wr, r.dag()], seed + 1)
    x = fill([l, si, sj, r], seed + 2)
    return L, R, x


def mps_link_sectors(N: int, b: int, m: int, sigma: float):
    """Sectors ((q0, 0), dim) of link b of an N-site spin-1/2 MPS with total
    Sz = 0 (q0 = 2 Sz of the b sites left of the link, the product-state
    convention of qtt.cpp:213-240): q0 runs over the charges both sides can
    reach; dims follow exp(-Sz^2 / (2 sigma^2)) scaled to m (minimum 1) and
    are capped by the number of left and right basis states of that charge."""
    from math import comb
    qs = [q for q in range(-b, b + 1, 2) if abs(q) <= N - b]
    w = np.exp(-(np.array(qs, dtype=float) / 2.0) ** 2 / (2 * sigma * sigma))
    d = np.maximum(1, np.round(m * w / w.sum())).astype(int)
    out = []
    for q, dd in zip(qs, d):
        cap = min(comb(b, (b + q) // 2), comb(N - b, (N - b + q) // 2))
        out.append((int(q), int(min(int(dd), cap))))
    return out


def random_mps(N: int, m: int, sigma: float, seed: int, mps_id: int = 1):
    """Site tensors psi_i = (ID<id>Link<i> In, Site<i> In, ID<id>Link<i+1> Out)
    (qtt.cpp:246-253) over all legal blocks, N(0,1) from PCG64(seed + i)
    times c_i = 1 / sqrt(2 max_q D_q) over the sectors of link i.

    The scale keeps the canonicalisation finite.  rc() (qtt.cpp:821-850)
    carries psi_i * U_{i+1} leftwards unnormalised, and with N(0,1) entries
    the carried squared norm of sector q grows by about
    D_q(link i) * sum_s n_{q+s} per site: 1e330 at N = 128, m = 8192, past
    the range of the FP64 Gram matrices.  With c_i it stays within
    [1e-3, 1] at every size.  The normalised state is the same, since a
    product of site scales is one global factor."""
    out = []
    for i in range(N):
        lsec = mps_link_sectors(N, i, m, sigma)
        li = make_index(INWARD, f"ID{mps_id}Link{i}", LINK, 0, [((q, 0), d) for q, d in lsec])
        si = site_index(INWARD, f"Site{i}", 0)
        lo = make_index(OUTWARD, f"ID{mps_id}Link{i + 1}", LINK, 0,
                        [((q, 0), d) for q, d in mps_link_sectors(N, i + 1, m, sigma)])
        t = fill([li, si, lo], seed + i)
        t.data *= 1.0 / np.sqrt(2.0 * max(d for _, d in lsec))
        out.append(t)
    return out


def lattice_terms(builder, lattice: str, Nx: int, Ny: int, yper: bool, J2: float):
    kind = 1 if lattice == "j1j2" else 0
    return builder.heisenberg_terms(kind, Nx, Ny, yper, 1.0, J2)


def mpo_bond(builder, name: str):
    """(W_site, W_site+1) as flat tensors from `builder` (the host layer's
    Python binding, or the oracle's in tests)."""
    lat, Nx, Ny, yper, J2, site, *_ = WORKLOADS[name]
    terms = lattice_terms(builder, lat, Nx, Ny, yper, J2)
    H = builder.mpo_autompo(Nx * Ny, terms)
    return H.tensor(site).to_flat(), H.tensor(site + 1).to_flat()


def workload_inputs(builder, name: str, seed: int = 7):
    lat, Nx, Ny, yper, J2, site, m, smax, sigma = WORKLOADS[name]
    W1, W2 = mpo_bond(builder, name)
    L, R, x = heff_inputs(W1, W2, site, m, smax, sigma, seed)
    return dict(W1=W1, W2=W2, L=L, R=R, x=x, site=site, m=m)


def env_live_model(env_elems, N: int, stride: int, scale: float = 1.0, elem_bytes: int = 8):
    """Largest number of environment bytes live at any bond of a sweep under
    the checkpoint policy of the DMRG driver, from the per-bond sizes of the
    environments a sweep produced (env_elems: 2 (N-1) entries per sweep in
    sweep order, the first half moving from the left, so entry i of a left
    half is TL[i+1] and entry i of a right half is TR[N-2-i]), scaled by
    `scale` (e.g. 4 for twice the bond dimension).

    At bond `site` of a half-sweep moving from the left, live are: the kept
    right environments of the previous half-sweep (sites that are multiples
    of `stride`, and the end), the window rebuilt from the nearest kept one
    (at most stride - 1), and the left environments kept so far in this
    half-sweep plus the one just produced; mirrored for the other direction.
    stride = 1 keeps one full set.  Returns (peak bytes, the bond and
    half-sweep of the peak)."""
    env = np.asarray(env_elems, dtype=float) * scale * elem_bytes
    nb = N - 1
    nh = len(env) // nb
    size_L = {}   # site -> bytes of TL[site] / TR[site] most recently produced
    size_R = {}
    best = (0.0, None)
    for h in range(nh):
        e = env[h * nb:(h + 1) * nb]
        left = h % 2 == 0
        if left:
            for i in range(nb):
                size_L[i + 1] = e[i]
        else:
            for i in range(nb):
                size_R[nb - 1 - i] = e[i]
    # sizes of every environment (the last value produced at each site)
    for h in range(nh):
        left = h % 2 == 0
        for i in range(nb):
            site = i if left else nb - i
            if left:
                far = [size_R.get(j, 0.0) for j in range(site + 1, N) if j % stride == 0 or j == N - 1]
                win = [size_R.get(j, 0.0) for j in range(site + 1, min(N, site + stride + 1))]
                near = [size_L.get(j, 0.0) for j in range(0, site + 2) if j % stride == 0 or j == site + 1]
            else:
                far = [size_L.get(j, 0.0) for j in range(0, site) if j % stride == 0 or j == 0]
                win = [size_L.get(j, 0.0) for j in range(max(0, site - stride), site)]
                near = [size_R.get(j, 0.0) for j in range(site - 1, N) if j % stride == 0 or j == site - 1]
            tot = sum(far) + sum(win) + sum(near)
            if tot > best[0]:
                best = (tot, (h, site))
    return best
