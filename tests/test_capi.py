"""CPU checks of the drop-in boundary: the C-ABI library loads and exports
every entry point include/ttgpu.h declares (no GPU calls)."""
import ctypes

from ttgpu import lib as T


def test_library_exports_every_header_symbol():
    L = T.load()
    syms = T.header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert missing == []


def test_binding_covers_header():
    # every declared function has a ctypes signature in the Python binding
    L = T.load()
    for s in T.header_symbols():
        f = getattr(L, s)
        assert f.restype is not None, s


def test_last_error_without_gpu_is_a_string():
    L = T.load()
    assert isinstance(L.ttg_last_error(), bytes)


def test_io_library_loads_and_exports_save_load():
    from ttgpu import host as P
    L = P._io_lib()
    for s in ("tth_mps_save", "tth_mps_load", "tth_mpo_save", "tth_mpo_load", "tth_io_last_error"):
        assert hasattr(L, s), s
