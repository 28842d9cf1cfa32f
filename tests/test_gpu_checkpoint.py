"""Checkpoint and resume (qTensorTrain::save / load, qtt.cpp:678-771;
qtensor::save / load, qtensor.cpp:1379-1415, 1477-1528; the resume form of
dmrg, dmrg.cpp:653-697).

* save -> load gives the same indices (names, arrows, types, levels,
  sectors), block lists and bit-identical block data, bond dimensions and
  centre, for an MPS and an MPO;
* the metadata file holds the reference's dataset names and the data file
  holds every block at the offsets the metadata states (read back with
  numpy from the flat file, not through the loader);
* a run stopped after two half-sweeps, saved, loaded and resumed at
  start_sweep = 2 gives bit-identical energies to the same run continued in
  memory;
* the ranged dmrg (dmrg.cpp:760-813) over the full range equals the
  schedule form, and a proper sub-range lowers the energy."""
import os
import subprocess

import numpy as np
import pytest

from ttgpu import host as P

pytestmark = pytest.mark.gpu

N = 12


def same_tensor(a, b):
    fa, fb = a.to_flat(), b.to_flat()
    assert [(i.arrow, i.name, i.type, i.level, i.qn, i.qdim) for i in fa.idx] == \
           [(i.arrow, i.name, i.type, i.level, i.qn, i.qdim) for i in fb.idx]
    assert np.array_equal(fa.block_qi, fb.block_qi)
    assert np.array_equal(fa.data.view(np.uint64), fb.data.view(np.uint64))


def ground_state(nsweeps=2):
    H = P.mpo_heisenberg_nn(N)
    psi = P.mps_product([i % 2 for i in range(N)])
    P.dmrg(psi, H, nsweeps, [8, 16], [1e-10], [2], 1e-10)
    return psi, H


def test_mps_and_mpo_round_trip(tmp_path):
    psi, H = ground_state()
    fn = str(tmp_path / "psi")
    psi.save(fn)
    assert os.path.exists(fn + ".h5") and os.path.exists(fn + "__T.bin")
    back = P.PMPS.load(fn + ".h5")
    assert len(back) == N and back.center() == psi.center()
    assert np.array_equal(back.bond_dims(), psi.bond_dims())
    for i in range(N):
        same_tensor(psi.tensor(i), back.tensor(i))
    # data file: every tensor's blocks back to back, dense column-major
    raw = np.fromfile(fn + "__T.bin", dtype=np.float64)
    assert np.array_equal(raw, np.concatenate([psi.tensor(i).to_flat().data for i in range(N)]))
    hn = str(tmp_path / "H")
    H.save(hn, str(tmp_path / "H_blocks.bin"))
    H2 = P.PMPO.load(hn + ".h5")
    for i in range(N):
        same_tensor(H.tensor(i), H2.tensor(i))
    assert P.psiHphi(psi, H2, psi) == P.psiHphi(psi, H, psi)


def test_metadata_names_and_offsets(tmp_path):
    psi, _ = ground_state(1)
    fn = str(tmp_path / "m")
    psi.save(fn)
    out = subprocess.run(["/opt/conda/bin/h5dump", "-n", fn + ".h5"], capture_output=True, text=True,
                         check=True).stdout
    names = {ln.split()[-1] for ln in out.splitlines() if ln.strip().startswith(("dataset", "group"))}
    for top in ("/wfn", "/length", "/phy_dim", "/bond_dims", "/center", "/totalQ", "/phy_qn", "/id"):
        assert top in names, top
    t3 = psi.tensor(3).to_flat()
    for k in ("T", "num_blocks", "rank", "idx_arrows", "idx_types", "idx_levels", "idx_qn_0",
              "idx_qdim_2", "idx_name_1", "offset_0", "block_0_qi"):
        assert f"/Tensor3/{k}" in names, k
    assert f"/Tensor3/offset_{t3.nblocks - 1}" in names
    assert f"/Tensor3/offset_{t3.nblocks}" not in names
    # offsets: cumulative bytes over the train, read with h5dump
    off = subprocess.run(["/opt/conda/bin/h5dump", "-d", "/Tensor3/offset_0", "-b", "LE", "-o",
                          str(tmp_path / "o.bin"), fn + ".h5"], capture_output=True, check=True)
    assert off.returncode == 0
    o = int(np.fromfile(str(tmp_path / "o.bin"), dtype=np.int64)[0])
    assert o == 8 * sum(psi.tensor(i).to_flat().data.size for i in range(3))


def test_resume_after_save_is_bit_identical(tmp_path):
    H = P.mpo_heisenberg_nn(N)
    # reference run: half-sweeps 0-1, then 2-5 continued in memory
    a = P.mps_product([i % 2 for i in range(N)])
    r1 = P.dmrg_resume(a, H, 2, 16, 1e-10, 2, 0, 1e-10)
    ra = P.dmrg_resume(a, H, 6, 16, 1e-10, 2, 2, 1e-10)
    # the same first part, saved, loaded into a fresh object and resumed
    b = P.mps_product([i % 2 for i in range(N)])
    r2 = P.dmrg_resume(b, H, 2, 16, 1e-10, 2, 0, 1e-10)
    assert r2["E"] == r1["E"]
    fn = str(tmp_path / "ck")
    b.save(fn)
    del b
    c = P.PMPS.load(fn + ".h5")
    rc = P.dmrg_resume(c, H, 6, 16, 1e-10, 2, 2, 1e-10)
    assert rc["E"] == ra["E"]
    assert np.array_equal(rc["bond_E"], ra["bond_E"])
    assert len(rc["halfsweep_E"]) == 4
    # an odd start resumes moving from the right (position(L-1) first):
    # half-sweeps 1 and 2
    d = P.PMPS.load(fn + ".h5")
    rd = P.dmrg_resume(d, H, 3, 16, 1e-10, 2, 1, 1e-10)
    assert len(rd["halfsweep_E"]) == 2 and rd["E"] < r1["E"] + 1e-9


def test_ranged_dmrg():
    H = P.mpo_heisenberg_nn(N)
    a = P.mps_product([i % 2 for i in range(N)])
    b = P.mps_product([i % 2 for i in range(N)])
    ra = P.dmrg(a, H, 2, [8, 16], [1e-10], [2], 1e-10)
    rb = P.dmrg_range(b, H, 0, N - 1, 2, [8, 16], [1e-10], [2], 1e-10)
    assert rb["E"] == ra["E"]
    c = P.mps_product([i % 2 for i in range(N)])
    rc = P.dmrg_range(c, H, 3, 8, 2, [16], [1e-10], [2], 1e-10)
    assert len(rc["bond_E"]) == 2 * 2 * 5
    assert rc["E"] < rc["E0"] - 0.1
    # only the links inside the range grew.  Right of it (sites 9..) nothing
    # moved: the bond dimensions are those the canonicalisation alone gives
    # the product state (its link sector sets, one state each).  Left of it
    # the final position(0) re-splits links 1-3 at cutoff 0, where a sector
    # that lost its partners inside the range drops out: never larger.
    bd = c.bond_dims()
    q = P.mps_product([i % 2 for i in range(N)])
    q.position(3)
    q.position(0)
    bq = q.bond_dims()
    assert np.array_equal(bd[9:], bq[9:]) and np.all(bd[:4] <= bq[:4]), (bd, bq)
    assert max(bd[4:9]) > max(bq[4:9]), (bd, bq)


def test_env_checkpoint_stride_gives_identical_results():
    # environments kept at every third site and rebuilt when needed: the same
    # bits as keeping every environment
    H = P.mpo_heisenberg_nn(20)
    runs = []
    for k in (1, 3):
        P.set_env_checkpoint_stride(k)
        try:
            psi = P.mps_product([i % 2 for i in range(20)])
            runs.append(P.dmrg(psi, H, 4, [16, 32], [1e-10], [2], 1e-10))
        finally:
            P.set_env_checkpoint_stride(1)
    assert runs[0]["E"] == runs[1]["E"]
    assert np.array_equal(runs[0]["bond_E"], runs[1]["bond_E"])
    assert np.array_equal(runs[0]["env_elems"], runs[1]["env_elems"])
