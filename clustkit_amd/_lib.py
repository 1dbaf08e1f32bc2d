"""ctypes binding of libclustkit_hip.so (the C-ABI in include/clustkit_hip.h).

The shared library is built in-tree by ``__graft_entry__.build()`` /
``make -C clustkit_amd/csrc``.  There is no CPU fallback: if the library is
missing or no GPU is visible, the entry points raise.
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libclustkit_hip.so"

CK_SW_LENGTH_RATIO = 0x1
CK_SW_MASK_IDENTITY = 0x2
CK_SW_PLAIN = 0x4
CK_SW_NO_GRAPH = 0x8
CK_SW_TIMING = 0x10
CK_SW_LATENCY = 0x20
CK_NW_PLAIN = 0x4
CK_RANK_IDENTITY = 0
CK_RANK_EVALUE = 1
CK_RANK_SCORE = 2
CK_NW_RAW = 0x8
CK_NW_STRIPE = 0x10

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F32 = ctypes.c_float
_U32 = ctypes.c_uint32

# name -> (restype, argtypes); every symbol include/clustkit_hip.h declares.
SIGNATURES = {
    "ck_last_error": (ctypes.c_char_p, []),
    "ck_sw_last_paths": (ctypes.c_int, [_P, _P]),
    "ck_sw_last_status": (ctypes.c_int, []),
    "ck_sw_set_config": (ctypes.c_int, [_I32, _I32, _I32]),
    "ck_sw_get_config": (ctypes.c_int, [_P, _P, _P]),
    "ck_kmer_index_build_device": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _I64, _P, _P,
                                                  _P, _P, _P, _P]),
    "ck_kmer_tile_offsets_device": (ctypes.c_int, [_P, _P, _I64, _I32, _P, _P]),
    "ck_version": (ctypes.c_char_p, []),
    "ck_device_count": (ctypes.c_int, []),
    "ck_init": (ctypes.c_int, [ctypes.c_int]),
    "ck_device_props": (ctypes.c_int, [ctypes.c_int, _P, _P, _P, ctypes.c_char_p, _I32]),
    "ck_batch_sw_align": (ctypes.c_int, [_P, _P, _P, _P, _I32, _I32, _P, _F32, _P, _P, _P, _P]),
    "ck_sw_align_host": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _P, _I64, _I32, _P, _P, _F32,
                                        _U32, _P, _P, _P, _P]),
    "ck_graph_csr_host": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _P, _P, _P]),
    "ck_graph_csr_device": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _P, _P, _P, _P]),
    "ck_nw_align_host": (ctypes.c_int, [_P, _I64, _P, _I64, _P, _P, _I64, _I32, _P, _F32, _U32,
                                        _P, _P, _P]),
    "ck_nw_align_device": (ctypes.c_int, [_P, _I64, _P, _P, _P, _I32, _I32, _P, _F32, _U32,
                                          _P, _P, _P, _P]),
    "ck_sw_align_device": (ctypes.c_int, [_P, _I64, _P, _P, _P, _I32, _I32, _P, _P, _F32,
                                          _U32, _P, _P, _P, _P, _P]),
    "ck_sw_align_device_noalloc": (ctypes.c_int, [_P, _I64, _P, _P, _P, _I32, _I32, _P, _P,
                                                  _F32, _U32, _P, _P, _P, _P, _P]),
    "ck_sw_last_stats": (ctypes.c_int, [_P, _P, _P]),
    "ck_sw_last_timing": (ctypes.c_int, [_P, _P, _P]),
    "ck_batch_score_queries": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _P, _P, _P, _P]
                               + [_I32] * 7 + [_P, _P, _P, _P]),
    "ck_kmer_score_device": (ctypes.c_int, [_P, _I64, _P, _P, _I32, _I32, _I32, _I32, _P, _P,
                                            _P, _P, _P, _P, _P, _I64, _P, _P, _P, _I32, _I32,
                                            _I32, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _P,
                                            _P]),
    "ck_kmer_db_codes_device": (ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _P, _P, _P]),
    "ck_kmer_score_similar_device": (ctypes.c_int, [_P, _I64, _P, _P, _I32, _I32, _P, _P, _P,
                                                    _P, _I32, _I32, _P, _I32, _I32, _I32,
                                                    _I32, _P, _P, _P, _P, _P]),
    "ck_kmer_tile_offsets_host": (ctypes.c_int, [_P, _P, _I64, _I32, _P]),
    "ck_search_pairs_device": (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _P, _P]),
    "ck_kmer_set_workspace_budget": (ctypes.c_int, [_I64]),
    "ck_kmer_query_bytes": (_I64, [_I32, _I32, _I32, _I32, _I32, _I32, _I32]),
    "ck_jaccard_device": (ctypes.c_int, [_P, _I64, _P, _I32, _P, _P]),
    "ck_jaccard_host": (ctypes.c_int, [_P, _I64, _P, _I64, _I32, _P]),
    "ck_sketch_device": (ctypes.c_int, [_P, _P, _P, _I64, _I32, _I32, _I32, ctypes.c_uint64,
                                        _P, _P]),
    "ck_lsh_candidates_device": (ctypes.c_int, [_P, _I64, _I32, _P, _P, _I32, _I32, _I32, _P,
                                                _P]),
    "ck_lsh_fetch_pairs_device": (ctypes.c_int, [_P, _I64, _P]),
    "ck_search_topk_device": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _I64, _I32, _I32,
                                             _I32, _I32, _P, _P, _P, _P, _P, _P]),
}

_lock = threading.Lock()
_lib = None


class HipLibraryError(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Load the library (raises when it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise HipLibraryError(
                    f"{LIB_PATH} is missing: build it with `make -C clustkit_amd/csrc` "
                    "(or __graft_entry__.build()); there is no CPU fallback")
            # One HIP runtime per process.  torch ships its own libamdhip64 /
            # libhsa-runtime64; importing it first makes the dynamic loader
            # bind this library's libamdhip64 dependency to that already
            # loaded copy, so device pointers and streams are shared.  Loaded
            # the other way round, the process holds two runtimes and the
            # second one finds no GPU.
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().ck_last_error().decode(errors="replace")
        raise HipLibraryError(f"clustkit_hip error {rc}: {msg}")


def device_props(device: int = 0) -> dict:
    """CU count, peak engine clock, HBM bytes and architecture name of a GPU
    (ck_device_props)."""
    lib = require_gpu()
    cu, clk, mem = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
    name = ctypes.create_string_buffer(64)
    check(lib.ck_device_props(int(device), ctypes.byref(cu), ctypes.byref(clk),
                              ctypes.byref(mem), name, 64))
    return {"num_cu": cu.value, "clock_mhz": clk.value / 1000.0, "hbm_bytes": mem.value,
            "arch": name.value.decode(errors="replace")}


def require_gpu() -> ctypes.CDLL:
    lib = load()
    if lib.ck_device_count() <= 0:
        raise HipLibraryError("no GPU visible: the alignment stage runs on the MI355X only")
    return lib
