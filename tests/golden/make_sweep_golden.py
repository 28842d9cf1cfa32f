"""Writes sweep_cyl6x12_m2048.json: the oracle's two-site DMRG sweep on the
6x12 Heisenberg cylinder (y-periodic, AutoMPO) at m = 2048, the energy
fixture the GPU sweep is compared with.

The run follows dmrg() (dmrg.cpp:702-755) in the oracle: one sweep (two
half-sweeps of 71 bonds), max bond dimension 2048, SVD cutoff 1e-9,
Davidson max_restart 2 and tol 1e-12, from the seeded random MPS of
ttgpu.workloads.random_mps (N = 72, m = 2048, sigma = 9, seed 11) after the
canonicalisation dmrg() itself runs.  Recorded: the final and half-sweep
energies, the initial energy, and per bond update the Davidson energy, the
truncation error and the kept bond dimension.

Usage: python tests/golden/make_sweep_golden.py [--m M] [--out PATH]
(OMP_NUM_THREADS sets the oracle's threads; at m = 2048 the run takes tens of
minutes on 8 cores)."""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tensor-tools_amd", "python"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from ttgpu.workloads import random_mps  # noqa: E402

NX, NY, SIGMA, SEED = 12, 6, 9.0, 11


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--m", type=int, default=2048)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    N = NX * NY
    H = O.mpo_autompo(N, O.heisenberg_terms(0, NX, NY, True, 1.0, 0.0))
    psi = O.mps_product([i % 2 for i in range(N)])
    for i, t in enumerate(random_mps(N, args.m, SIGMA, seed=SEED)):
        psi.set_tensor(i, O.OTensor.from_flat(t))
    psi.set_center(-1)
    t0 = time.time()
    r = O.dmrg_log(psi, H, 1, [args.m], [1e-9], [2], tol=1e-12)
    out = dict(system="Heisenberg 6x12 cylinder (y-periodic), AutoMPO", sites=N, m=args.m,
               sigma=SIGMA, seed=SEED, sweeps=1, cutoff=1e-9, max_restart=2, tol=1e-12,
               E=r["E"], E_per_site=r["E"] / N, E0=r["E0"],
               halfsweep_E=[float(v) for v in r["halfsweep_E"]],
               bond_E=[float(v) for v in r["bond_E"]],
               trunc_err=[float(v) for v in r["trunc_err"]],
               kept=[int(v) for v in r["kept"]],
               bond_dims=[int(v) for v in psi.bond_dims()],
               oracle_threads=O.threads(), seconds=round(time.time() - t0, 1))
    path = args.out or os.path.join(HERE, f"sweep_cyl6x12_m{args.m}.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"E/N = {out['E_per_site']:.15f} in {out['seconds']} s on {out['oracle_threads']} threads -> {path}")


if __name__ == "__main__":
    main()
