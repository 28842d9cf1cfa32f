"""Exact ground-state energies for the DMRG checks (scipy Lanczos on the
Sz = 0 sector of the open spin-1/2 Heisenberg chain H = J sum S_i.S_{i+1}).
Writes tests/golden/ed_energies.json.  Pure numpy/scipy; no reference code."""
import itertools
import json
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def sector_basis(N):
    states = [s for s in range(1 << N) if bin(s).count("1") == N // 2]
    return np.array(states, dtype=np.int64)


def heisenberg_sector(N, bonds, J=1.0):
    basis = sector_basis(N)
    index = {s: i for i, s in enumerate(basis)}
    rows, cols, vals = [], [], []
    for i, s in enumerate(basis):
        diag = 0.0
        for a, b in bonds:
            sa = (s >> a) & 1
            sb = (s >> b) & 1
            diag += J * (0.25 if sa == sb else -0.25)
            if sa != sb:
                t = s ^ ((1 << a) | (1 << b))
                rows.append(index[t]); cols.append(i); vals.append(0.5 * J)
        rows.append(i); cols.append(i); vals.append(diag)
    n = len(basis)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def main():
    out = {}
    for N in (10, 12, 20):
        bonds = [(i, i + 1) for i in range(N - 1)]
        H = heisenberg_sector(N, bonds)
        e = sla.eigsh(H, k=1, which="SA", tol=1e-14)[0][0]
        out[f"heisenberg_obc_N{N}"] = float(e)
        print(N, H.shape[0], repr(e))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ed_energies.json")
    json.dump(out, open(path, "w"), indent=1)


if __name__ == "__main__":
    main()
