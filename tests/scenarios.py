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
