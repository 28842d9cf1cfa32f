"""Python binding of the C++ host layer (qtensor / big_qtensor /
tensor_davidsonIT / svd_bond / qMPS / dmrg over device blocks), through the
tth_* entry points of libttgpu.so.  Method names follow the reference API."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import lib as T
from .flat import FlatTensor, Index

_i, _d, _l, _p = C.c_int, C.c_double, C.c_long, C.c_void_p
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

_sig_done = False


def _lib():
    global _sig_done
    L = T.load()
    if not _sig_done:
        sig = {
            "tth_init": ([_i], _i), "tth_set_precision": ([_i], None),
            "tth_set_env_precision": ([_i], None),
            "tth_last_error": ([], C.c_char_p),
            "tth_tensor_import": ([_i, _ip, _ip, _ip, C.c_char_p, _ip, _ip, _ip, _i, _ip, _dp], _p),
            "tth_init_block": ([_i, _ip, _ip, _ip, C.c_char_p, _ip, _ip, _ip], _p),
            "tth_tensor_info": ([_p, _ip, C.POINTER(_l)], _i),
            "tth_tensor_export": ([_p, _ip, _ip, _ip, C.c_char_p, _ip, _ip, _ip, _ip, _lp, _dp], _i),
            "tth_tensor_free": ([_p], None), "tth_tensor_copy": ([_p], _p),
            "tth_contract": ([_p, _p], _p), "tth_inner": ([_p, _p], _d), "tth_norm": ([_p], _d),
            "tth_add": ([_p, _p, _d], _i), "tth_scale": ([_p, _d], _i),
            "tth_prime": ([_p, _i], None), "tth_dag": ([_p], _i),
            "tth_big_new": ([_p, _p, _p, _p], _p), "tth_big_free": ([_p], None),
            "tth_big_product": ([_p, _p], _p), "tth_big_expec": ([_p, _p], _d),
            "tth_big_size": ([_p], _l), "tth_big_heff": ([_p, _p], _p),
            "tth_big_time": ([_p, _p, _i, C.POINTER(_d), _dp], _i),
            "tth_big_info": ([_p, _p, C.POINTER(_d), _dp, _lp, _ip, _ip], _i),
            "tth_davidson": ([_p, _p, _i, _d, C.POINTER(_i), C.POINTER(_d)], _d),
            "tth_svd_bond": ([_p, _p, _p, C.c_char_p, _i, _d, _l, C.POINTER(_i), C.POINTER(_p)], _d),
            "tth_mpo_heisenberg_nn": ([_i, _i], _p),
            "tth_mpo_autompo": ([_i, _i, _i, _dp, _ip, _ip, _ip, _ip], _p),
            "tth_heisenberg_terms": ([_i, _i, _i, _i, _d, _d, _i, _dp, _ip, _ip, _ip, _ip], _i),
            "tth_mps_product": ([_i, _i, _ip], _p),
            "tth_mpo_tensor": ([_p, _i], _p), "tth_mpo_length": ([_p], _i),
            "tth_qtt_tensor": ([_p, _i], _p), "tth_qtt_set_tensor": ([_p, _i, _p], _i),
            "tth_qtt_length": ([_p], _i), "tth_qtt_center": ([_p], _i),
            "tth_qtt_set_center": ([_p, _i], None),
            "tth_psiHphi": ([_p, _p, _p], _d), "tth_psiphi": ([_p, _p], _d),
            "tth_set_dmrg_progress": ([_i], None),
            "tth_mps_free": ([_p], None), "tth_mpo_free": ([_p], None),
            "tth_mps_position": ([_p, _i], _d), "tth_mps_normalize": ([_p], _i),
            "tth_qtt_bond_dims": ([_p, _ip], None),
            "tth_dmrg": ([_p, _p, _i, _i, _ip, _i, _dp, _i, _ip, _d, _dp, _dp, C.POINTER(_d), _dp,
                          _dp, _dp, _ip, _dp], _d),
            "tth_dmrg_range": ([_p, _p, _i, _i, _i, _i, _ip, _i, _dp, _i, _ip, _d, _dp, _dp,
                                C.POINTER(_d), _dp, _dp, _dp, _ip, _dp], _d),
            "tth_dmrg_resume": ([_p, _p, _i, _i, _d, _i, _i, _d, _dp, _dp, C.POINTER(_d), _dp, _dp,
                                 _dp, _ip, _dp], _d),
            "tth_set_env_checkpoint_stride": ([_i], None),
            "tth_mps_new": ([], _p), "tth_mpo_new": ([], _p),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _sig_done = True
    return L


def set_precision(prec: int):
    _lib().tth_set_precision(prec)


def set_env_precision(prec: int):
    """Arithmetic of the environment updates' contractions (FP64 default)."""
    _lib().tth_set_env_precision(prec)


def _err(what):
    return RuntimeError(f"ttgpu host: {what}: {_lib().tth_last_error().decode()}")


def _check(h, what):
    if not h:
        raise _err(what)
    return h


def _ok(rc, what):
    if rc is None:
        return rc
    if isinstance(rc, float):
        if rc != rc:   # NaN marks an error return
            raise _err(what)
    elif rc < 0:
        raise _err(what)
    return rc


class PTensor:
    def __init__(self, h):
        self.h = _check(h, "tensor")

    def __del__(self):
        if getattr(self, "h", None) and T._lib is not None:
            T._lib.tth_tensor_free(self.h)
            self.h = None

    @classmethod
    def from_flat(cls, t: FlatTensor) -> "PTensor":
        a, lv, ty, nm, ns, qn, qd, bqi, data = t.c_arrays()
        if len(data) == 0:
            data = np.zeros(1)
        if len(bqi) == 0:
            bqi = np.zeros(1, np.int32)
        return cls(_lib().tth_tensor_import(t.rank, a, lv, ty, nm, ns, qn, qd, t.nblocks, bqi, data))

    def to_flat(self) -> FlatTensor:
        info = np.zeros(4, np.int32)
        tot = _l(0)
        _ok(_lib().tth_tensor_info(self.h, info, C.byref(tot)), "tensor info")
        rank, nst, nb, nl = (int(v) for v in info)
        arrows, levels, types, nsec = (np.zeros(max(rank, 1), np.int32) for _ in range(4))
        names = C.create_string_buffer(nl + 1)
        qn = np.zeros(max(2 * nst, 1), np.int32)
        qd = np.zeros(max(nst, 1), np.int32)
        bqi = np.zeros(max(nb * rank, 1), np.int32)
        boff = np.zeros(nb + 1, np.int64)
        data = np.zeros(max(tot.value, 1))
        _ok(_lib().tth_tensor_export(self.h, arrows, levels, types, names, nsec, qn, qd, bqi, boff, data),
            "tensor export")
        nml = names.value.decode().split("\n")
        idx, off = [], 0
        for i in range(rank):
            k = int(nsec[i])
            idx.append(Index(int(arrows[i]), nml[i], int(types[i]), int(levels[i]),
                             [(int(qn[2 * (off + s)]), int(qn[2 * (off + s) + 1])) for s in range(k)],
                             [int(qd[off + s]) for s in range(k)]))
            off += k
        return FlatTensor(idx, bqi[:nb * rank].reshape(nb, rank).copy(), data[:tot.value].copy())

    def copy(self):
        return PTensor(_lib().tth_tensor_copy(self.h))

    def __mul__(self, other):
        return PTensor(_lib().tth_contract(self.h, other.h))

    def inner(self, other):
        return _ok(_lib().tth_inner(self.h, other.h), "inner")

    def norm(self):
        return _ok(_lib().tth_norm(self.h), "norm")

    def add(self, other, c=1.0):
        _ok(_lib().tth_add(self.h, other.h, c), "add")

    def scale(self, c):
        _ok(_lib().tth_scale(self.h, c), "scale")

    def prime(self, inc=1):
        _lib().tth_prime(self.h, inc)

    def dag(self):
        _ok(_lib().tth_dag(self.h), "dag")


def init_block(idx):
    t = FlatTensor(idx, np.zeros((0, len(idx)), np.int32), np.zeros(0))
    a, lv, ty, nm, ns, qn, qd, _, _ = t.c_arrays()
    return PTensor(_lib().tth_init_block(len(idx), a, lv, ty, nm, ns, qn, qd))


class PBig:
    def __init__(self, L: PTensor, W1: PTensor, W2, R: PTensor):
        self._keep = (L, W1, W2, R)
        self.h = _check(_lib().tth_big_new(L.h, W1.h if W1 else None, W2.h if W2 else None, R.h),
                        "big_qtensor")

    def __del__(self):
        if getattr(self, "h", None) and T._lib is not None:
            T._lib.tth_big_free(self.h)

    def product(self, x: PTensor) -> PTensor:
        return PTensor(_lib().tth_big_product(self.h, x.h))

    def expec(self, x):
        return _ok(_lib().tth_big_expec(self.h, x.h), "expec")

    def size(self):
        return _ok(_lib().tth_big_size(self.h), "big_qtensor size")

    def heff(self, x: PTensor):
        """The device operator (ttg_heff) planned for x's indices."""
        return _check(_lib().tth_big_heff(self.h, x.h), "heff")

    def time(self, x: PTensor, n: int):
        """(ms per product over n HIP-graph replays, ms of each stage kernel)"""
        ms = _d(0)
        ms3 = np.zeros(3)
        _ok(_lib().tth_big_time(self.h, x.h, n, C.byref(ms), ms3), "time")
        return ms.value, ms3

    def info(self, x: PTensor):
        fl = _d(0)
        sf = np.zeros(3)
        nt = np.zeros(3, np.int64)
        cfg = np.zeros(3, np.int32)
        ntl = np.zeros(3, np.int32)
        _ok(_lib().tth_big_info(self.h, x.h, C.byref(fl), sf, nt, cfg, ntl), "info")
        return dict(flops=fl.value, stage_flops=sf, ntasks=nt, tile_cfg=cfg, ntiles=ntl)

    def davidson(self, x: PTensor, max_restart=2, tol=1e-12):
        n = _i(0)
        r = _d(0)
        e = _ok(_lib().tth_davidson(self.h, x.h, max_restart, tol, C.byref(n), C.byref(r)), "davidson")
        return e, n.value


def svd_bond(combined, A_left, A_right, mid_name, direction, cutoff, K):
    kept = _i(0)
    S = _p()
    tr = _lib().tth_svd_bond(combined.h, A_left.h, A_right.h, mid_name.encode(), direction, cutoff,
                             K, C.byref(kept), C.byref(S))
    _ok(tr, "svd_bond")
    return tr, kept.value, PTensor(S.value)


_io = None


def _io_lib():
    """libttgpu_io.so: qTensorTrain save / load (HDF5 metadata + flat block
    data, qtt.cpp:678-771)."""
    global _io
    if _io is None:
        _lib()
        path = os.path.join(os.path.dirname(T.LIB_PATH), "libttgpu_io.so")
        if not os.path.exists(path):
            raise RuntimeError(f"libttgpu_io.so not built ({path}); run __graft_entry__.build()")
        L = C.CDLL(path)
        for name in ("tth_mps_save", "tth_mpo_save", "tth_mps_load", "tth_mpo_load"):
            f = getattr(L, name)
            f.argtypes = [_p, C.c_char_p, C.c_char_p]
            f.restype = _i
        L.tth_io_last_error.argtypes = []
        L.tth_io_last_error.restype = C.c_char_p
        _io = L
    return _io


def _io_ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"ttgpu io: {what}: {_io_lib().tth_io_last_error().decode()}")


class _Train:
    """save / load of qMPS and qMPO (qtt.cpp:678-771)"""
    _kind = ""

    def save(self, fn: str, wfn: str = ""):
        """metadata to fn + '.h5', block data to wfn (default fn + '__T.bin')"""
        _io_ok(getattr(_io_lib(), f"tth_{self._kind}_save")(self.h, fn.encode(), wfn.encode()), "save")

    @classmethod
    def load(cls, fn: str, data_prefix: str = ""):
        """a train read from fn (the .h5 file); its data file name, as stored,
        is prefixed with data_prefix"""
        t = cls(getattr(_lib(), f"tth_{cls._kind}_new")())
        _io_ok(getattr(_io_lib(), f"tth_{cls._kind}_load")(t.h, fn.encode(), data_prefix.encode()), "load")
        return t


class PMPO(_Train):
    _kind = "mpo"

    def __init__(self, h):
        self.h = _check(h, "mpo")

    def __del__(self):
        if getattr(self, "h", None) and T._lib is not None:
            T._lib.tth_mpo_free(self.h)

    def __len__(self):
        return _lib().tth_mpo_length(self.h)

    def tensor(self, i) -> PTensor:
        return PTensor(_lib().tth_mpo_tensor(self.h, i))


class PMPS(_Train):
    _kind = "mps"

    def __init__(self, h):
        self.h = _check(h, "mps")

    def __del__(self):
        if getattr(self, "h", None) and T._lib is not None:
            T._lib.tth_mps_free(self.h)

    def __len__(self):
        return _lib().tth_qtt_length(self.h)

    def tensor(self, i) -> PTensor:
        return PTensor(_lib().tth_qtt_tensor(self.h, i))

    def set_tensor(self, i, t: PTensor):
        _ok(_lib().tth_qtt_set_tensor(self.h, i, t.h), "set_tensor")

    def position(self, site):
        return _ok(_lib().tth_mps_position(self.h, site), "position")

    def normalize(self):
        _ok(_lib().tth_mps_normalize(self.h), "normalize")

    def center(self):
        return _lib().tth_qtt_center(self.h)

    def set_center(self, c: int):
        """-1: not canonical (the next position() right-canonicalises first)"""
        _lib().tth_qtt_set_center(self.h, c)

    def bond_dims(self):
        out = np.zeros(len(self) + 1, np.int32)
        _lib().tth_qtt_bond_dims(self.h, out)
        return out


def heisenberg_terms(kind: int, Nx: int, Ny: int, yper: bool, J1=1.0, J2=0.0):
    cap = 6 * 4 * Nx * Ny + 16
    coef = np.zeros(cap)
    o1, i1, o2, i2 = (np.zeros(cap, np.int32) for _ in range(4))
    n = _lib().tth_heisenberg_terms(kind, Nx, Ny, int(yper), J1, J2, cap, coef, o1, i1, o2, i2)
    return coef[:n].copy(), o1[:n].copy(), i1[:n].copy(), o2[:n].copy(), i2[:n].copy()


def mpo_autompo(N: int, terms, mpo_id: int = 2) -> PMPO:
    coef, o1, i1, o2, i2 = terms
    return PMPO(_lib().tth_mpo_autompo(N, mpo_id, len(coef), np.ascontiguousarray(coef), o1, i1, o2, i2))


def mpo_heisenberg_nn(N: int, mpo_id: int = 2) -> PMPO:
    return PMPO(_lib().tth_mpo_heisenberg_nn(N, mpo_id))


def mps_product(state, mps_id: int = 1) -> PMPS:
    st = np.ascontiguousarray(np.array(state, dtype=np.int32))
    return PMPS(_lib().tth_mps_product(len(st), mps_id, st))


def set_dmrg_progress(every: int):
    """dmrg() prints a progress line to stderr every `every` bonds (0: silent)"""
    _lib().tth_set_dmrg_progress(every)


def psiHphi(psi: PMPS, H: PMPO, phi: PMPS) -> float:
    """<psi|H|phi> (mps/observables.cpp:64-105)"""
    return _ok(_lib().tth_psiHphi(psi.h, H.h, phi.h), "psiHphi")


def psiphi(psi: PMPS, phi: PMPS) -> float:
    """<psi|phi> (mps/observables.cpp:415-450)"""
    return _ok(_lib().tth_psiphi(psi.h, phi.h), "psiphi")


def _dmrg_out(E, hs, hsec, e0, stats, bE, tr, kp, env, nb):
    return dict(E=E, halfsweep_E=hs, halfsweep_s=hsec, E0=e0.value, t_davidson=stats[0],
                t_svd=stats[1], t_env=stats[2], n_matvec=int(stats[3]), seconds=stats[4],
                trunc_err_max=stats[5], trunc_err_last_halfsweep=stats[6], pool_live_max=stats[7],
                bond_E=bE[:nb], trunc_err=tr[:nb], kept=kp[:nb], env_elems=env[:nb])


def _dmrg_bufs(nhalf, nb):
    return (np.zeros(max(nhalf, 1)), np.zeros(max(nhalf, 1)), _d(0), np.zeros(8), np.zeros(max(nb, 1)),
            np.zeros(max(nb, 1)), np.zeros(max(nb, 1), np.int32), np.zeros(max(nb, 1)))


def set_env_checkpoint_stride(k: int):
    """Keep the environments of every k-th site between half-sweeps and
    rebuild the others when a bond needs them (1: one full set; results do
    not depend on k)."""
    _lib().tth_set_env_checkpoint_stride(k)


def _schedules(maxbd, cutoff, max_restart):
    return (np.ascontiguousarray(np.array(maxbd, np.int32)), np.ascontiguousarray(np.array(cutoff, np.float64)),
            np.ascontiguousarray(np.array(max_restart, np.int32)))


def dmrg(psi: PMPS, H: PMPO, nsweeps: int, maxbd, cutoff, max_restart, tol=1e-12):
    """dmrg(qMPS, qMPO, num_sweeps, max_bd, cutoff, max_restart) (dmrg.cpp:702-755)"""
    mb, ct, mr = _schedules(maxbd, cutoff, max_restart)
    nb = 2 * nsweeps * (len(H) - 1)
    hs, hsec, e0, stats, bE, tr, kp, env = _dmrg_bufs(2 * nsweeps, nb)
    E = _ok(_lib().tth_dmrg(psi.h, H.h, nsweeps, len(mb), mb, len(ct), ct, len(mr), mr, tol, hs,
                            hsec, C.byref(e0), stats, bE, tr, kp, env), "dmrg")
    return _dmrg_out(E, hs, hsec, e0, stats, bE, tr, kp, env, nb)


def dmrg_range(psi: PMPS, H: PMPO, start: int, stop: int, nsweeps: int, maxbd, cutoff, max_restart,
               tol=1e-12):
    """dmrg(qMPS, qMPO, start, stop, ...) (dmrg.cpp:760-813): bonds of sites
    [start, stop] only"""
    mb, ct, mr = _schedules(maxbd, cutoff, max_restart)
    nb = 2 * nsweeps * (stop - start)
    hs, hsec, e0, stats, bE, tr, kp, env = _dmrg_bufs(2 * nsweeps, nb)
    E = _ok(_lib().tth_dmrg_range(psi.h, H.h, start, stop, nsweeps, len(mb), mb, len(ct), ct, len(mr), mr,
                                  tol, hs, hsec, C.byref(e0), stats, bE, tr, kp, env), "dmrg")
    return _dmrg_out(E, hs, hsec, e0, stats, bE, tr, kp, env, nb)


def dmrg_resume(psi: PMPS, H: PMPO, num_half: int, max_bd: int, cutoff: float, max_restart: int,
                start_sweep: int, tol=1e-12):
    """dmrg(qMPS, qMPO, num_sweeps, max_bd, cutoff, mode, search_space_size,
    max_restart, start_sweep) (dmrg.cpp:653-697): half-sweeps start_sweep ..
    num_half - 1, even ones moving from the left"""
    nh = max(num_half - start_sweep, 0)
    nb = nh * (len(H) - 1)
    hs, hsec, e0, stats, bE, tr, kp, env = _dmrg_bufs(nh, nb)
    E = _ok(_lib().tth_dmrg_resume(psi.h, H.h, num_half, max_bd, cutoff, max_restart, start_sweep, tol, hs,
                                   hsec, C.byref(e0), stats, bE, tr, kp, env), "dmrg")
    return _dmrg_out(E, hs, hsec, e0, stats, bE, tr, kp, env, nb)
