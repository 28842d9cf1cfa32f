#!/usr/bin/env python3
"""Headline benchmark: effective-Hamiltonian matvec TFLOP/s at m=2048.

Workload (the reference's batching configuration): spin-1/2 Heisenberg on a
6x12 cylinder (periodic in the width-6 direction), AutoMPO qMPO (bond
dimension at the middle bond reported as w), two-site effective Hamiltonian
L * (W35 W36) * R around the middle bond; MPS bond dimension m=2048 over
Sz sectors -30..30 with Gaussian (sigma=9) sector dims; L, R, x N(0,1) from a
fixed seed.  One step = one H_eff matvec y = ((L*x)*W12)*R (three grouped
MFMA GEMM launches replayed from a HIP graph) with all operands resident in
HBM.  FP64 is the reported `value`; FP32 and split-BF16 (3 bf16 MFMA
products, FP32 accumulation) are measured on the same inputs with their error
against the FP64 result.

cpu_baseline: the oracle (FP64 CPU restatement of the reference's
qtensor::operator* / big_qtensor::product, OpenMP over output blocks) timed
on this host for a bounded number of products of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tensor-tools_amd", "python"))

import numpy as np  # noqa: E402

MFMA_PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}
HBM_PEAK_GBS = 8000.0
WORKLOAD = "heisenberg_cyl6x12_m2048"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stage_traffic(workload, stage, cfg):
    """HBM bytes per launch of one FP64 H_eff stage kernel, from the committed
    rocprofv3 --pmc summary of that workload (FETCH_SIZE and WRITE_SIZE in
    KiB per dispatch, each scaled by what the counter reads for a device copy
    of known size), or None when no summary for this workload, stage and work
    shape exists."""
    path = os.path.join(ROOT, "profiles", "heff_cyl6x12_m2048_traffic.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    if d.get("workload") != workload:
        return None
    rec = d.get("stages", {}).get(str(stage + 1), {}).get(str(cfg))
    if not rec:
        return None
    return int(round((rec["fetch_kib"] * d["fetch_scale"] + rec["write_kib"] * d["write_scale"]) * 1024))


def sweep_leg(P, T, args):
    """--sweeps full two-site DMRG sweeps (both half-sweeps, N-1 bonds each)
    on the workload's lattice, max bond dimension --sweep-m, SVD cutoff
    --sweep-cutoff (1e-9),
    Davidson max_restart 2, tol 1e-12, from the same seeded random MPS
    (Gaussian Sz sectors of width sigma, m per link where the Hilbert space
    allows) in every precision form.  Reduced-precision forms end each
    Davidson call with the FP64 residual-correction step (two FP64 products
    and a 2x2 FP64 Rayleigh-Ritz problem), so after one sweep from a random
    state they are one Krylov step ahead of the FP64 form; the converged
    comparison is converged_precision()."""
    from ttgpu.workloads import WORKLOADS, lattice_terms, random_mps
    lat, Nx, Ny, yper, J2, site, m_w, smax, sigma = WORKLOADS[args.workload]
    N = Nx * Ny
    H = P.mpo_autompo(N, lattice_terms(P, lat, Nx, Ny, yper, J2))
    t0 = time.time()
    tensors = random_mps(N, args.sweep_m, sigma, seed=11)
    log(f"[bench] sweep: random MPS N={N} m={args.sweep_m} built in {time.time() - t0:.1f}s")
    forms = {"fp64": T.FP64, "fp32": T.FP32, "bf16x3": T.BF16X3}
    res = {}
    P.set_dmrg_progress(12)
    P.set_env_checkpoint_stride(args.env_stride)
    for name in args.sweep_forms.split(","):
        P.set_precision(forms[name])
        psi = P.mps_product([i % 2 for i in range(N)])
        for i, t in enumerate(tensors):
            psi.set_tensor(i, P.PTensor.from_flat(t))
        psi.set_center(-1)
        r = P.dmrg(psi, H, args.sweeps, [args.sweep_m], [args.sweep_cutoff], [2], tol=1e-12)
        hs = np.asarray(r["halfsweep_s"], dtype=float)
        sweep_s = float(np.sum(hs)) / args.sweeps
        warm = None
        if args.sweep_warm and args.sweep_m <= 4096:
            # the same sweeps again from the same state: the same work, with the
            # tile-shape choices the first run timed already cached (a DMRG run's
            # later sweeps find most of them there)
            del psi
            psi = P.mps_product([i % 2 for i in range(N)])
            for i, t in enumerate(tensors):
                psi.set_tensor(i, P.PTensor.from_flat(t))
            psi.set_center(-1)
            r2 = P.dmrg(psi, H, args.sweeps, [args.sweep_m], [args.sweep_cutoff], [2], tol=1e-12)
            warm = round(float(np.sum(r2["halfsweep_s"])) / args.sweeps, 4)
            if r2["E"] != r["E"]:
                log(f"[bench] note: {name} repeat sweep energy differs by {abs(r2['E'] - r['E']):.2e} "
                    f"(tile shapes chosen by timing differ between the runs)")
        res[name] = dict(sweep_s=round(sweep_s, 4), sweep_s_warm=warm,
                         sweep_s_each=[round(float(v), 4) for v in hs.reshape(-1, 2).sum(axis=1)],
                         halfsweep_s=[round(float(v), 4) for v in hs],
                         total_s=round(r["seconds"], 4), E=r["E"], E_per_site=r["E"] / N,
                         n_matvec=r["n_matvec"], t_davidson=round(r["t_davidson"], 4),
                         t_svd=round(r["t_svd"], 4), t_env=round(r["t_env"], 4),
                         trunc_err_max=r["trunc_err_max"],
                         trunc_err_last_halfsweep=r["trunc_err_last_halfsweep"],
                         max_bond=int(np.max(psi.bond_dims())),
                         env_gb_max=round(8e-9 * float(np.max(r["env_elems"])), 3),
                         env_gb_last_halfsweep=round(8e-9 * float(np.sum(r["env_elems"][-(N - 1):])), 3),
                         pool_live_gb_max=round(r["pool_live_max"] / 1e9, 3))
        if args.env_log:
            with open(args.env_log, "a") as f:
                f.write(json.dumps({"form": name, "m": args.sweep_m, "workload": args.workload,
                                    "env_elems": [int(v) for v in r["env_elems"]]}) + "\n")
        log(f"[bench] sweep {name}: {sweep_s:.3f} s, repeated {warm} s (halves {res[name]['halfsweep_s']}, "
            f"davidson {r['t_davidson']:.2f} s, svd {r['t_svd']:.2f} s, env {r['t_env']:.2f} s, "
            f"{r['n_matvec']} matvecs), E/N = {r['E'] / N:.12f}, max bond {res[name]['max_bond']}")
        del psi
    P.set_precision(T.FP64)
    P.set_env_checkpoint_stride(1)
    if "fp64" in res:
        for name, v in res.items():
            v["dE_per_site_vs_fp64"] = abs(v["E_per_site"] - res["fp64"]["E_per_site"])
            v["dtrunc_last_halfsweep_vs_fp64"] = abs(v["trunc_err_last_halfsweep"]
                                                    - res["fp64"]["trunc_err_last_halfsweep"])
    out = dict(m=args.sweep_m, sites=N, sweeps=args.sweeps, cutoff=args.sweep_cutoff, max_restart=2,
               env_stride=args.env_stride, forms=res)
    if args.cpu_seconds > 0 and args.sweep_m == m_w and args.sweeps == 1 and args.sweep_cutoff == 1e-9:
        out["cpu_baseline"] = cpu_bond_sample(args, N)
    return out


def converged_precision(P, T, args):
    """The precision forms at convergence on the sweep leg's system: FP64
    sweeps from the seeded random MPS until the energy per site changes by
    less than --converge-tol between sweeps (at most --converge-sweeps), then
    every form runs two more sweeps from that same converged state; reported:
    each form's final energy per site and its difference from the FP64 run's,
    beside the per-sweep FP64 energies that show the convergence."""
    from ttgpu.workloads import WORKLOADS, lattice_terms, random_mps
    lat, Nx, Ny, yper, J2, site, m_w, smax, sigma = WORKLOADS[args.workload]
    N = Nx * Ny
    m = args.sweep_m
    H = P.mpo_autompo(N, lattice_terms(P, lat, Nx, Ny, yper, J2))
    P.set_precision(T.FP64)
    psi = P.mps_product([i % 2 for i in range(N)])
    for i, t in enumerate(random_mps(N, m, sigma, seed=11)):
        psi.set_tensor(i, P.PTensor.from_flat(t))
    psi.set_center(-1)
    hist = []
    t0 = time.time()
    for s in range(args.converge_sweeps):
        r = P.dmrg(psi, H, 1, [m], [args.sweep_cutoff], [2], tol=1e-12)
        hist.append(r["E"] / N)
        log(f"[bench] converge: FP64 sweep {s + 1}: E/N = {hist[-1]:.13f}")
        if len(hist) > 1 and abs(hist[-1] - hist[-2]) < args.converge_tol:
            break
    t_conv = time.time() - t0
    start = [psi.tensor(i).to_flat() for i in range(N)]
    out = {}
    for name, prec in (("fp64", T.FP64), ("fp32", T.FP32), ("bf16x3", T.BF16X3)):
        P.set_precision(prec)
        q = P.mps_product([i % 2 for i in range(N)])
        for i, t in enumerate(start):
            q.set_tensor(i, P.PTensor.from_flat(t))
        q.set_center(0)
        r = P.dmrg(q, H, 2, [m], [args.sweep_cutoff], [2], tol=1e-12)
        out[name] = {"E_per_site": r["E"] / N, "halfsweep_E_per_site": [float(v) / N for v in r["halfsweep_E"]],
                     "trunc_err_last_halfsweep": r["trunc_err_last_halfsweep"],
                     "sweep_s": round(float(np.sum(r["halfsweep_s"])) / 2, 4)}
        del q
    P.set_precision(T.FP64)
    for v in out.values():
        v["dE_per_site_vs_fp64"] = abs(v["E_per_site"] - out["fp64"]["E_per_site"])
    conv = abs(hist[-1] - hist[-2]) if len(hist) > 1 else None
    log("[bench] converged precision (6x12, m=%d): " % m + ", ".join(
        f"{k} E/N {v['E_per_site']:.13f} |dE|/site vs fp64 {v['dE_per_site_vs_fp64']:.2e}" for k, v in out.items()))
    return {"system": f"{args.workload} lattice, m={m}, cutoff {args.sweep_cutoff}", "fp64_sweeps": len(hist),
            "fp64_E_per_site_by_sweep": hist, "last_sweep_change_per_site": conv,
            "converge_tol": args.converge_tol, "fp64_converge_s": round(t_conv, 2),
            "forms_two_more_sweeps": out}


def precision_check(P, T):
    """Converged comparison of the precision forms: the N=20 open Heisenberg
    chain, two-site DMRG from the Neel state, 10 sweeps, bond dimension
    16, 32, 48, then 64 for the last seven sweeps, cutoff 1e-10, Davidson
    max_restart 2, tol 1e-10; the final energy per site of each form against the FP64 run and against exact
    diagonalisation (tests/golden/ed_energies.json)."""
    N = 20
    ed = json.load(open(os.path.join(ROOT, "tests", "golden", "ed_energies.json")))["heisenberg_obc_N20"]
    H = P.mpo_heisenberg_nn(N)
    out = {}
    for name, prec in (("fp64", T.FP64), ("fp32", T.FP32), ("bf16x3", T.BF16X3)):
        P.set_precision(prec)
        psi = P.mps_product([i % 2 for i in range(N)])
        r = P.dmrg(psi, H, 10, [16, 32, 48, 64, 64, 64, 64, 64, 64, 64], [1e-10], [2], 1e-10)
        out[name] = {"E_per_site": r["E"] / N, "dE_per_site_vs_ed": abs(r["E"] - ed) / N}
    P.set_precision(T.FP64)
    for v in out.values():
        v["dE_per_site_vs_fp64"] = abs(v["E_per_site"] - out["fp64"]["E_per_site"])
    log("[bench] precision check (N=20 chain, 10 sweeps): " + ", ".join(
        f"{k} |dE|/site vs fp64 {v['dE_per_site_vs_fp64']:.2e} vs ED {v['dE_per_site_vs_ed']:.2e}"
        for k, v in out.items()))
    return {"system": "Heisenberg chain N=20 OBC, 10 sweeps, m 16,32,48,64x7, cutoff 1e-10", "forms": out}


def cpu_bond_sample(args, N):
    """The oracle's cost of one bulk bond update at the sweep's m (Davidson
    with max_restart 2 on the middle-bond H_eff of the matvec workload, the
    sector SVD with cutoff 1e-9 / K = m, one environment step), times the
    2(N-1) bond updates of a sweep.  Bulk bonds are the largest, so this
    over-estimates the edges of the chain."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from ttgpu import workloads as W
    from ttgpu.flat import LINK, OUTWARD, make_index
    W1, W2 = W.mpo_bond(O, args.workload)
    lat, Nx, Ny, yper, J2, site, m, smax, sigma = W.WORKLOADS[args.workload]
    L, R, x = W.heff_inputs(W1, W2, site, m, smax, sigma, 7)
    oL, oR, oW1, oW2, ox = (O.OTensor.from_flat(t) for t in (L, R, W1, W2, x))
    ob = O.OBig(oL, oW1, oW2, oR)
    t0 = time.time()
    ob.davidson(ox, 2, 1e-12)
    t1 = time.time()
    l, si, sj, r = x.idx
    mid = make_index(OUTWARD, f"ID1Link{site + 1}", LINK, 0, list(zip(l.qn, l.qdim)))
    Al, Ar = O.init_block([l, si, mid]), O.init_block([mid.dag(), sj, r])
    O.svd_bond(ox, Al, Ar, mid.name, 0, 1e-9, m)
    t2 = time.time()
    O.env_step(Al, oL, oW1)
    t3 = time.time()
    per_bond = t3 - t0
    log(f"[bench] cpu bond sample: davidson {t1 - t0:.2f} s, svd {t2 - t1:.2f} s, env {t3 - t2:.2f} s "
        f"on {O.threads()} threads")
    return {"sweep_s_estimate": round(per_bond * 2 * (N - 1), 2), "unit": "s", "cores": O.threads(),
            "kind": "port", "blas": O.blas_name(), "per_bond_s": round(per_bond, 3),
            "sample": "one bulk bond update (Davidson, sector SVD, environment step) of the oracle "
                      "at the sweep's m, times 2(N-1) bonds"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default=WORKLOAD)
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate CPU time budget of the cpu_baseline leg (0 disables)")
    ap.add_argument("--forms", default="fp64,fp32,bf16x3")
    ap.add_argument("--sweep-m", type=int, default=2048,
                    help="bond dimension of the DMRG sweep leg (0 skips the sweep)")
    ap.add_argument("--sweep-forms", default="fp64,bf16x3,fp32")
    ap.add_argument("--sweeps", type=int, default=1, help="sweeps of the sweep leg (sweep_s is per sweep)")
    ap.add_argument("--sweep-warm", type=int, default=1,
                    help="1: run the sweep leg a second time from the same state and report its time "
                         "(sweep_s_warm: tile-shape choices already timed); m <= 4096 only")
    ap.add_argument("--sweep-cutoff", type=float, default=1e-9, help="SVD cutoff of the sweep leg")
    ap.add_argument("--env-stride", type=int, default=0,
                    help="keep every k-th environment between half-sweeps (0: 1 up to m=4096, else 4)")
    ap.add_argument("--env-log", default="",
                    help="append each sweep form's per-bond environment sizes to this JSON-lines file")
    ap.add_argument("--converge-sweeps", type=int, default=16,
                    help="most FP64 sweeps of the converged precision comparison (0 skips it)")
    ap.add_argument("--converge-tol", type=float, default=1e-10,
                    help="energy-per-site change between sweeps that counts as converged")
    ap.add_argument("--precision-check", type=int, default=1,
                    help="1: after the sweep leg, the N=20 chain comparison of the precision forms")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a dispatch of "
                         "known traffic that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library has no multi-GPU path (asked for --gpus "
              f"{args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.env_stride == 0:
        args.env_stride = 1 if args.sweep_m <= 4096 else 4
    if args.sweep_m > 4096 and args.env_stride < 2:
        print("bench.py: a sweep leg above m=4096 needs --env-stride >= 2.  One full environment set "
              "of the 8x16 cylinder at m=8192 is projected above the device's 288 GB (m=4096: 79 GB "
              "after two sweeps, environments growing as m^2); with stride k only every k-th "
              "environment is kept between half-sweeps.", file=sys.stderr)
        sys.exit(2)

    from ttgpu import host as P
    from ttgpu import lib as T
    from ttgpu.workloads import WORKLOADS, workload_inputs

    T.init(0)
    if args.copy_calib_mib > 0:
        from ttgpu.flat import INWARD, LINK, FlatTensor, make_index
        n = args.copy_calib_mib * (1 << 20) // 8
        ix = make_index(INWARD, "calib", LINK, 0, [((0, 0), n)])
        src = T.DTensor.from_flat(FlatTensor([ix], np.zeros((1, 1), np.int32), np.ones(n)))
        dst = src.copy()
        T.sync()
        log(f"[bench] copy calibration: {n * 8} bytes read, {n * 8} bytes written")
        del src, dst
    t0 = time.time()
    inp = workload_inputs(P, args.workload)
    W1, W2, L, R, x = (inp[k] for k in ("W1", "W2", "L", "R", "x"))
    w_mid = int(sum(W1.idx[3].qdim))
    log(f"[bench] inputs built in {time.time() - t0:.1f}s: |L|={L.data.size} |R|={R.data.size} "
        f"|x|={x.data.size} w={w_mid}")
    pL, pR, pW1, pW2, px = (P.PTensor.from_flat(t) for t in (L, R, W1, W2, x))

    forms = {"fp64": T.FP64, "fp32": T.FP32, "bf16x3": T.BF16X3}
    results = {}
    y_ref = None
    for name in args.forms.split(","):
        prec = forms[name]
        P.set_precision(prec)
        big = P.PBig(pL, pW1, pW2, pR)
        info = big.info(px)
        y = big.product(px).to_flat().data
        if name == "fp64":
            y_ref = y
        big.time(px, max(1, args.warmup))            # warmup replays
        ms, ms3 = big.time(px, args.steps)             # timed replays (HIP events)
        relerr = (float(np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref))
                  if y_ref is not None else None)
        results[name] = dict(ms=ms, stage_ms=ms3.tolist(), flops=info["flops"],
                             stage_flops=info["stage_flops"].tolist(),
                             ntasks=info["ntasks"].tolist(), tile_cfg=info["tile_cfg"].tolist(),
                             ntiles=info["ntiles"].tolist(),
                             tflops=info["flops"] / (ms * 1e-3) / 1e12, relerr_vs_fp64=relerr)
        log(f"[bench] {name}: {ms:.3f} ms/matvec, {results[name]['tflops']:.2f} TFLOP/s, "
            f"stages {np.round(ms3, 4).tolist()} ms, rel err vs fp64 {relerr}")
        del big
    P.set_precision(T.FP64)

    main_form = results["fp64"]
    flops = main_form["flops"]
    # dominant kernel: the slowest of the three grouped-GEMM stages
    k = int(np.argmax(main_form["stage_ms"]))
    achieved = main_form["stage_flops"][k] / (main_form["stage_ms"][k] * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": MFMA_PEAK_TFLOPS["f64"],
                "unit": "TFLOP/s", "frac": round(achieved / MFMA_PEAK_TFLOPS["f64"], 4),
                "traffic": stage_traffic(args.workload, k, main_form["tile_cfg"][k]),
                "kernel": f"stage {k + 1} grouped GEMM (FP64, work shape {main_form['tile_cfg'][k]})"}

    cpu = None
    if args.cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        oL, oR, oW1, oW2, ox = (O.OTensor.from_flat(t) for t in (L, R, W1, W2, x))
        ob = O.OBig(oL, oW1, oW2, oR)
        t1 = ob.time_products(ox, 1)
        n = int(max(1, min(50, args.cpu_seconds / max(t1, 1e-6))))
        tn = ob.time_products(ox, n)
        yo = ob.product(ox).to_flat()
        # FP64 GPU result vs the oracle on the same inputs (dense comparison
        # via block keys: the oracle's block order differs from the device's)
        yg = P.PBig(pL, pW1, pW2, pR).product(px).to_flat()
        go = {yg.block_qn_key(b): yg.block(b) for b in range(yg.nblocks)}
        num = den = 0.0
        for b in range(yo.nblocks):
            ob_ = yo.block(b)
            gb = go.get(yo.block_qn_key(b))
            num += float(np.sum((ob_ - (gb if gb is not None else 0.0)) ** 2))
            den += float(np.sum(ob_ ** 2))
        cpu = {"value": round(flops / (tn / n) / 1e12, 5), "unit": "TFLOP/s", "cores": O.threads(),
               "kind": "port", "blas": O.blas_name(),
               "sample": f"{n} H_eff matvecs of the same workload ({tn / n * 1e3:.1f} ms each, "
                         f"FP64 oracle, OpenMP over output blocks, block products on {O.blas_name()})",
               "fp64_gpu_vs_oracle_relerr": float(np.sqrt(num / den))}
        log(f"[bench] cpu baseline: {tn / n * 1e3:.1f} ms/matvec on {O.threads()} threads; "
            f"GPU FP64 vs oracle rel err {cpu['fp64_gpu_vs_oracle_relerr']:.2e}")

    sweep = sweep_leg(P, T, args) if args.sweep_m > 0 else None
    if sweep is not None:
        if args.precision_check:
            sweep["precision_check"] = precision_check(P, T)
        if args.converge_sweeps > 0:
            sweep["precision_converged"] = converged_precision(P, T, args)

    lat, Nx, Ny, yper, J2, site, m, smax, sigma = WORKLOADS[args.workload]
    out = {
        "metric": "H_eff matvec TFLOP/s at m=2048 (FP64; split-BF16 and FP32 forms reported beside)",
        "value": round(main_form["tflops"], 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_form["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded N(0,1) environments and two-site vector; exact AutoMPO W)",
        "config": {"workload": args.workload, "lattice": f"{Ny}x{Nx} cylinder" if Ny > 1 else
                   f"chain N={Nx}", "bond": [site, site + 1], "m": m, "sz_sectors": [-smax, smax],
                   "sigma": sigma, "w": w_mid, "flops_per_matvec": flops,
                   "x_elems": int(x.data.size), "L_elems": int(L.data.size)},
        "forms": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in results.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "sweep": sweep,
    }
    if "bf16x3" in results:
        out["bf16x3_speedup_vs_fp64"] = round(results["bf16x3"]["tflops"] / main_form["tflops"], 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
