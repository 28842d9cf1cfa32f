"""ctypes binding of libttgpu.so (C-ABI declared in include/ttgpu.h).

The library must be built (tensor-tools_amd/Makefile); there is no fallback:
every call goes to the HIP implementation, and a missing library or GPU is an
error.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

from .flat import FlatTensor, Index

PKG = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
ROOT = os.path.dirname(PKG)
LIB_PATH = os.path.join(PKG, "lib", "libttgpu.so")
HEADER = os.path.join(ROOT, "include", "ttgpu.h")

FP64, FP32, BF16X3 = 0, 1, 2
MOVE_FROM_LEFT, MOVE_FROM_RIGHT = 0, 1

_i = C.c_int
_d = C.c_double
_p = C.c_void_p
_sz = C.c_size_t
_i64 = C.c_int64
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")

_lib = None


def header_symbols():
    """Every function name declared in include/ttgpu.h."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(ttg_\w+)\s*\(", txt)) |
                  set(re.findall(r"\bconst char\*\s*(ttg_\w+)\s*\(", txt)))


def load(path: str = LIB_PATH):
    """Load the library without touching the GPU."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libttgpu.so not built ({path}); run __graft_entry__.build()")
    L = C.CDLL(path)
    sig = {
        "ttg_init": ([_i], _i), "ttg_finalize": ([], _i), "ttg_last_error": ([], C.c_char_p),
        "ttg_sync": ([], _i), "ttg_mem_info": ([C.POINTER(_sz)] * 3, _i),
        "ttg_tensor_create": ([_i, _ip, _ip, _ip, _ip, _i, _ip, C.POINTER(_p)], _i),
        "ttg_tensor_init_blocks": ([_i, _ip, _ip, _ip, _ip, C.POINTER(_p)], _i),
        "ttg_tensor_upload": ([_p, _dp, _sz], _i), "ttg_tensor_download": ([_p, _dp, _sz], _i),
        "ttg_tensor_info": ([_p, C.POINTER(_i), C.POINTER(_i), C.POINTER(_sz)], _i),
        "ttg_tensor_meta_download": ([_p, _ip, _lp], _i),
        "ttg_tensor_copy": ([_p, C.POINTER(_p)], _i), "ttg_tensor_free": ([_p], _i),
        "ttg_tensor_fill": ([_p, _d], _i), "ttg_tensor_set_random": ([_p, C.c_uint64], _i),
        "ttg_dot": ([_p, _p, C.POINTER(_d)], _i), "ttg_nrm2": ([_p, C.POINTER(_d)], _i),
        "ttg_dot_matched": ([_p, _p, C.POINTER(_d)], _i),
        "ttg_scale": ([_p, _d], _i), "ttg_axpy": ([_d, _p, _p], _i),
        "ttg_contract_plan": ([_p, _p, _i, _ip, _ip, _i, C.POINTER(_p)], _i),
        "ttg_plan_info": ([_p, C.POINTER(_i), C.POINTER(_i64), C.POINTER(_d)], _i),
        "ttg_plan_tasks_download": ([_p, _ip], _i),
        "ttg_plan_new_output": ([_p, C.POINTER(_p)], _i),
        "ttg_contract_exec": ([_p, _p, _p, _p], _i), "ttg_plan_free": ([_p], _i),
        "ttg_contract": ([_p, _p, _i, _ip, _ip, _i, C.POINTER(_p)], _i),
        "ttg_heff_create": ([_p, _p, _p, _p, _ip, _i, C.POINTER(_p)], _i),
        "ttg_heff_create_twin": ([_p, _i, C.POINTER(_p)], _i),
        "ttg_heff_new_output": ([_p, C.POINTER(_p)], _i),
        "ttg_heff_apply": ([_p, _p, _p], _i),
        "ttg_heff_info": ([_p, C.POINTER(_d), _lp, _dp], _i),
        "ttg_heff_time": ([_p, _p, _p, _i, C.POINTER(_d)], _i),
        "ttg_heff_free": ([_p], _i),
        "ttg_heff_stage_time": ([_p, _p, _i, _dp], _i),
        "ttg_heff_tiles": ([_p, _ip, _ip], _i),
        "ttg_heff_stage_plan": ([_p, _i, C.POINTER(_p)], _i),
        "ttg_env_update": ([_p, _p, _p, _i, _i, C.POINTER(_p)], _i),
        "ttg_env_pairs": ([_i, _ip], _i),
        "ttg_same_blocks": ([_p, _p, C.POINTER(_i)], _i),
        "ttg_svd": ([_p, _i, _i, _d, _i64, C.POINTER(_p), C.POINTER(_p), C.POINTER(_p),
                     C.POINTER(_d), C.POINTER(_i64)], _i),
        "ttg_tensor_index_download": ([_p, _i, C.POINTER(_i), _ip, _ip], _i),
        "ttg_small_eigh": ([_i, _dp, _dp, _dp], _i),
        "ttg_lincomb": ([_i, C.POINTER(_p), _dp, _p], _i),
        "ttg_embed": ([_p, _p], _i),
        "ttg_embed_noalloc": ([_p, _p], _i),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def lib():
    return load()


def check(rc: int):
    if rc != 0:
        raise RuntimeError("ttgpu: " + lib().ttg_last_error().decode())


_inited = False


def init(device: int = 0):
    global _inited
    if not _inited:
        check(lib().ttg_init(device))
        _inited = True


class DTensor:
    """Owning handle of a device tensor; keeps the host index descriptors."""

    def __init__(self, h, idx):
        self.h = h
        self.idx = [ix.copy() for ix in idx]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ttg_tensor_free(self.h)
            self.h = None

    # ---- construction ---------------------------------------------------
    @classmethod
    def from_flat(cls, t: FlatTensor) -> "DTensor":
        init()
        a, _, _, _, ns, qn, qd, bqi, data = t.c_arrays()
        h = _p()
        if len(bqi) == 0:
            bqi = np.zeros(1, np.int32)
        check(lib().ttg_tensor_create(t.rank, a, ns, qn, qd, t.nblocks, bqi, C.byref(h)))
        d = cls(h, t.idx)
        if len(data):
            check(lib().ttg_tensor_upload(h, np.ascontiguousarray(data), len(data)))
        return d

    @classmethod
    def init_blocks(cls, idx) -> "DTensor":
        init()
        t = FlatTensor(idx, np.zeros((0, len(idx)), np.int32), np.zeros(0))
        a, _, _, _, ns, qn, qd, _, _ = t.c_arrays()
        h = _p()
        check(lib().ttg_tensor_init_blocks(len(idx), a, ns, qn, qd, C.byref(h)))
        return cls(h, idx)

    # ---- queries ----------------------------------------------------------
    def info(self):
        r, nb, tot = _i(0), _i(0), _sz(0)
        check(lib().ttg_tensor_info(self.h, C.byref(r), C.byref(nb), C.byref(tot)))
        return r.value, nb.value, tot.value

    def meta(self):
        r, nb, _ = self.info()
        qi = np.zeros(max(nb * r, 1), np.int32)
        off = np.zeros(nb + 1, np.int64)
        check(lib().ttg_tensor_meta_download(self.h, qi, off))
        return qi[:nb * r].reshape(nb, r), off

    def data(self) -> np.ndarray:
        _, _, tot = self.info()
        out = np.zeros(max(tot, 1))
        check(lib().ttg_tensor_download(self.h, out, tot))
        return out[:tot]

    def upload(self, data: np.ndarray):
        check(lib().ttg_tensor_upload(self.h, np.ascontiguousarray(data, dtype=np.float64), len(data)))

    def to_flat(self) -> FlatTensor:
        qi, _ = self.meta()
        return FlatTensor([ix.copy() for ix in self.idx], qi, self.data())

    # ---- vector ops -------------------------------------------------------
    def dot(self, other) -> float:
        out = _d(0)
        check(lib().ttg_dot(self.h, other.h, C.byref(out)))
        return out.value

    def dot_matched(self, other) -> float:
        """qtensor::contract over the blocks both tensors hold"""
        out = _d(0)
        check(lib().ttg_dot_matched(self.h, other.h, C.byref(out)))
        return out.value

    def norm(self) -> float:
        out = _d(0)
        check(lib().ttg_nrm2(self.h, C.byref(out)))
        return out.value

    def scale(self, a: float):
        check(lib().ttg_scale(self.h, a))

    def axpy(self, a: float, x: "DTensor"):
        check(lib().ttg_axpy(a, x.h, self.h))

    def copy(self) -> "DTensor":
        h = _p()
        check(lib().ttg_tensor_copy(self.h, C.byref(h)))
        return DTensor(h, self.idx)

    def fill(self, v: float):
        check(lib().ttg_tensor_fill(self.h, v))


def _pairs_arr(pairs):
    pa = np.array([p[0] for p in pairs] or [0], np.int32)
    pb = np.array([p[1] for p in pairs] or [0], np.int32)
    return pa, pb


def result_indices(A_idx, B_idx, pairs):
    pa = {p for p, _ in pairs}
    pb = {p for _, p in pairs}
    return [A_idx[i] for i in range(len(A_idx)) if i not in pa] + \
           [B_idx[i] for i in range(len(B_idx)) if i not in pb]


class Plan:
    def __init__(self, A: DTensor, B: DTensor, pairs, prec=FP64):
        self.h = _p()
        pa, pb = _pairs_arr(pairs)
        check(lib().ttg_contract_plan(A.h, B.h, len(pairs), pa, pb, prec, C.byref(self.h)))
        self.cidx = result_indices(A.idx, B.idx, pairs)

    @classmethod
    def borrowed(cls, h, cidx, owner) -> "Plan":
        p = cls.__new__(cls)
        p.h, p.cidx, p._owner = h, cidx, owner
        return p

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None and getattr(self, "_owner", None) is None:
            _lib.ttg_plan_free(self.h)

    def info(self):
        nb, nt, fl = _i(0), _i64(0), _d(0)
        check(lib().ttg_plan_info(self.h, C.byref(nb), C.byref(nt), C.byref(fl)))
        return nb.value, nt.value, fl.value

    def tasks(self) -> np.ndarray:
        _, nt, _ = self.info()
        out = np.zeros(6 * max(nt, 1), np.int32)
        check(lib().ttg_plan_tasks_download(self.h, out))
        return out[:6 * nt].reshape(nt, 6)

    def new_output(self) -> DTensor:
        h = _p()
        check(lib().ttg_plan_new_output(self.h, C.byref(h)))
        return DTensor(h, self.cidx)

    def execute(self, A: DTensor, B: DTensor, Cout: DTensor):
        check(lib().ttg_contract_exec(self.h, A.h, B.h, Cout.h))


def contract(A: DTensor, B: DTensor, pairs, prec=FP64) -> DTensor:
    h = _p()
    pa, pb = _pairs_arr(pairs)
    check(lib().ttg_contract(A.h, B.h, len(pairs), pa, pb, prec, C.byref(h)))
    return DTensor(h, result_indices(A.idx, B.idx, pairs))


class Heff:
    """y = ((L*x)*W12)*R, three planned stages replayed from a HIP graph."""

    def __init__(self, L: DTensor, W12: DTensor, R: DTensor, x_like: DTensor, pairs3, prec=FP64):
        enc = []
        for pr in pairs3:
            enc.append(len(pr))
            for a, b in pr:
                enc += [a, b]
        self._keep = (L, W12, R)
        self.h = _p()
        check(lib().ttg_heff_create(L.h, W12.h, R.h, x_like.h, np.array(enc, np.int32), prec,
                                    C.byref(self.h)))
        self.xidx = x_like.idx

    @classmethod
    def twin(cls, src: "Heff", prec) -> "Heff":
        """The same operator at precision `prec` on src's plans and tile shapes."""
        t = cls.__new__(cls)
        t._keep = src._keep + (src,)
        t.h = _p()
        check(lib().ttg_heff_create_twin(src.h, prec, C.byref(t.h)))
        t.xidx = src.xidx
        return t

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ttg_heff_free(self.h)

    def new_output(self) -> DTensor:
        h = _p()
        check(lib().ttg_heff_new_output(self.h, C.byref(h)))
        return DTensor(h, self.xidx)

    def apply(self, x: DTensor, y: DTensor):
        check(lib().ttg_heff_apply(self.h, x.h, y.h))

    def info(self):
        fl = _d(0)
        nt = np.zeros(3, np.int64)
        sf = np.zeros(3)
        check(lib().ttg_heff_info(self.h, C.byref(fl), nt, sf))
        return fl.value, nt, sf

    def time(self, x: DTensor, y: DTensor, n: int) -> float:
        ms = _d(0)
        check(lib().ttg_heff_time(self.h, x.h, y.h, n, C.byref(ms)))
        return ms.value

    def stage_time(self, x: DTensor, n: int):
        out = np.zeros(3)
        check(lib().ttg_heff_stage_time(self.h, x.h, n, out))
        return out

    def stage_plan(self, stage: int, A_idx, B_idx, pairs) -> "Plan":
        """Borrowed view of stage `stage`'s plan (freed with the operator)."""
        h = _p()
        check(lib().ttg_heff_stage_plan(self.h, stage, C.byref(h)))
        return Plan.borrowed(h, result_indices(A_idx, B_idx, pairs), self)

    def tiles(self):
        cfg = np.zeros(3, np.int32)
        nt = np.zeros(3, np.int32)
        check(lib().ttg_heff_tiles(self.h, cfg, nt))
        return cfg, nt


def env_pairs(direction: int):
    """The three stages' index pairs of ttg_env_update for a direction."""
    out = np.zeros(16, np.int32)
    check(lib().ttg_env_pairs(direction, out))
    res, k = [], 0
    for _ in range(3):
        n = int(out[k]); k += 1
        res.append([(int(out[k + 2 * i]), int(out[k + 2 * i + 1])) for i in range(n)])
        k += 2 * n
    return res


def env_update(psi: DTensor, env: DTensor, W: DTensor, direction: int, prec=FP64) -> DTensor:
    """env' = ((psi^dag' * env) * W) * psi (updateEnv, dmrg.cpp:491-509)."""
    h = _p()
    check(lib().ttg_env_update(psi.h, env.h, W.h, direction, prec, C.byref(h)))
    p = env_pairs(direction)
    psid = [ix.dag() for ix in psi.idx]
    for ix in psid:
        ix.level += 1
    i1 = result_indices(psid, env.idx, p[0])
    i2 = result_indices(i1, W.idx, p[1])
    return DTensor(h, result_indices(i2, psi.idx, p[2]))


def embed(src: DTensor, dst: DTensor):
    check(lib().ttg_embed(src.h, dst.h))


def embed_noalloc(src: DTensor, dst: DTensor):
    """ttg_embed on the library's fixed scratch (no device allocation)."""
    check(lib().ttg_embed_noalloc(src.h, dst.h))


def small_eigh(M: np.ndarray):
    n = M.shape[0]
    ev = np.zeros(n)
    V = np.zeros(n * n)
    check(lib().ttg_small_eigh(n, np.ascontiguousarray(M.reshape(-1, order="F")), ev, V))
    return ev, V.reshape((n, n), order="F")


def sync():
    check(lib().ttg_sync())
