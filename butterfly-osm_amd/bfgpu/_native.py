"""ctypes bindings of the in-tree native libraries.

libbfgpu.so (HIP engine, C ABI in include/bfgpu.h) and libbfcch.so (CCH
builder).  Loading fails loudly when a library has not been built: there is no
CPU fallback for the engine.
"""
import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class GraphDesc(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32),
                ("up_offsets", u64p), ("up_targets", u32p), ("up_weights", u32p),
                ("down_offsets", u64p), ("down_targets", u32p), ("down_weights", u32p)]


class CtxInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_uint32), ("m_up", C.c_uint64), ("m_down", C.c_uint64),
                ("etree_depth", C.c_uint32), ("n_top", C.c_uint32), ("n_segments", C.c_uint32),
                ("n_segment_levels", C.c_uint32), ("n_bottom_levels", C.c_uint32),
                ("lanes", C.c_uint32), ("lane_capacity", C.c_uint32),
                ("device_bytes", C.c_uint64)]


class Timings(C.Structure):
    _fields_ = [("up_ms", C.c_float), ("sweep_ms", C.c_float), ("output_ms", C.c_float),
                ("total_ms", C.c_float), ("batches", C.c_uint32), ("sweep_launches", C.c_uint32),
                ("lanes_run", C.c_uint64), ("nodes_swept", C.c_uint64), ("arcs_swept", C.c_uint64),
                ("dom_ms", C.c_float), ("dom_launches", C.c_uint32), ("dom_bytes", C.c_uint64),
                ("gate_seg_units", C.c_uint64), ("gate_seg_open", C.c_uint64),
                ("gate_bt_units", C.c_uint64), ("gate_bt_open", C.c_uint64),
                ("join_ms", C.c_float), ("join_visits", C.c_uint64),
                ("join_label_bits", C.c_uint32), ("join_tiles_u32", C.c_uint64)]


class RelaxStats(C.Structure):
    _fields_ = [("triangles", C.c_uint64), ("improvements", C.c_uint64), ("levels", C.c_uint32),
                ("launches", C.c_uint32), ("passes", C.c_uint32), ("kernel_ms", C.c_float),
                ("total_ms", C.c_float)]


class BfcParams(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("seed", C.c_uint64),
                ("spacing_m", C.c_double), ("jitter_frac", C.c_double), ("delete_frac", C.c_double),
                ("oneway_frac", C.c_double), ("chord_frac", C.c_double), ("ban_frac", C.c_double),
                ("leaf_threshold", C.c_int32), ("n_threads", C.c_int32), ("verbose", C.c_int32)]


class BfcCch(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m_up", C.c_uint64), ("m_down", C.c_uint64),
                ("up_off", u64p), ("up_tgt", u32p), ("up_w", u32p),
                ("dn_off", u64p), ("dn_tgt", u32p), ("dn_w", u32p),
                ("x", C.POINTER(C.c_float)), ("y", C.POINTER(C.c_float)),
                ("rank_to_filtered", u32p),
                ("m_ebg", C.c_uint64), ("ebg_off", u64p), ("ebg_head", u32p), ("ebg_w", u32p),
                ("nbg_nodes", C.c_uint32), ("nbg_edges", C.c_uint32), ("n_components", C.c_uint32),
                ("n_shortcuts", C.c_uint64),
                ("t_graph_s", C.c_double), ("t_order_s", C.c_double), ("t_contract_s", C.c_double)]


# entry points of include/bfgpu.h: name -> (argtypes)
BFGPU_SYMBOLS = {
    "bfg_ctx_create": [C.POINTER(GraphDesc), C.c_int, C.POINTER(C.c_void_p)],
    "bfg_ctx_destroy": [C.c_void_p],
    "bfg_last_error": [],
    "bfg_ctx_get_info": [C.c_void_p, C.POINTER(CtxInfo)],
    "bfg_ctx_last_timings": [C.c_void_p, C.POINTER(Timings)],
    "bfg_ctx_configure": [C.c_void_p, C.c_uint32, C.c_uint32],
    "bfg_ctx_tune": [C.c_void_p, C.c_uint32, C.c_uint32],
    "bfg_ctx_set_concurrency": [C.c_void_p, C.c_uint32],
    "bfg_ctx_set_label_width": [C.c_void_p, C.c_uint32],
    "bfg_validate_graph": [C.POINTER(GraphDesc), C.POINTER(CtxInfo)],
    "bfg_matrix": [C.c_void_p, u32p, C.c_uint32, u32p, C.c_uint32, u32p],
    "bfg_matrix_bounded": [C.c_void_p, u32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32, u32p],
    "bfg_matrix_device": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p],
    "bfg_matrix_device_noalloc": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32,
                                  C.c_void_p],
    "bfg_matrix_multi": [C.c_void_p, C.c_uint32, u32p, u32p, u32p, u32p, u32p, u32p],
    "bfg_query_batch": [C.c_void_p, u32p, C.c_uint32, C.c_uint32, u32p],
    "bfg_isochrone": [C.c_void_p, u32p, u32p, C.c_uint32, u64p, u32p],
    "bfg_phast_seeded": [C.c_void_p, u32p, u32p, u32p, C.c_uint32, u32p, u32p],
    "bfg_isochrone_lists": [C.c_void_p, u32p, u32p, u32p, C.c_uint32, u32p, u64p, u64p, u32p, u32p, C.c_uint64,
                            u64p],
    "bfg_isochrone_lists_device": [C.c_void_p, u32p, u32p, u32p, C.c_uint32, u32p, C.c_void_p, u64p, u32p,
                                   C.c_void_p, C.c_uint64, u64p],
    "bfg_table_seeded": [C.c_void_p, u32p, u32p, u32p, C.c_uint32, u32p, u32p, u32p, C.c_uint32, C.c_uint32,
                         u32p, C.POINTER(C.c_uint8)],
    "bfg_ctx_attach_length": [C.c_void_p, C.POINTER(GraphDesc), u32p, u32p],
    "bfg_ctx_pair_reverse": [C.c_void_p, C.c_void_p],
    "bfg_device_allocations": [u64p],
    "bfg_ctx_create_reverse": [C.POINTER(GraphDesc), C.c_int, C.POINTER(C.c_void_p)],
    "bfg_table_bucket": [C.c_void_p, u32p, C.c_uint32, u32p, C.c_uint32, C.c_uint32, u32p],
    "bfg_table_bucket_device": [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32,
                                C.c_void_p],
    "bfg_phast_seeded_2ch": [C.c_void_p, u32p, u32p, u32p, u32p, C.c_uint32, u32p, u32p, u32p],
    "bfg_table_seeded_2ch": [C.c_void_p, u32p, u32p, u32p, u32p, C.c_uint32, u32p, u32p, u32p, u32p, C.c_uint32,
                             C.c_uint32, u32p, u32p, C.POINTER(C.c_uint8)],
    "bfg_triangle_relax": [C.POINTER(GraphDesc), C.c_int, u32p, u32p, u32p, u32p, u32p, u32p,
                           C.POINTER(RelaxStats)],
    "bfg_triangle_relax_lex": [C.POINTER(GraphDesc), C.c_int, u32p, u32p, u64p, u64p, u32p, u32p, u32p, u32p,
                               u32p, u32p, C.POINTER(RelaxStats)],
}

_bfgpu = None
_bfcch = None


def lib_path(name):
    return os.path.join(LIB_DIR, name)


def bfgpu():
    """The HIP engine library; raises if it has not been built."""
    global _bfgpu
    if _bfgpu is None:
        path = lib_path("libbfgpu.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build() (no CPU fallback exists)")
        lib = C.CDLL(path)
        for name, args in BFGPU_SYMBOLS.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name == "bfg_last_error" else (None if name == "bfg_ctx_destroy" else C.c_int)
        _bfgpu = lib
    return _bfgpu


def bfcch():
    global _bfcch
    if _bfcch is None:
        path = lib_path("libbfcch.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run __graft_entry__.build()")
        lib = C.CDLL(path)
        lib.bfc_build_grid.argtypes = [C.POINTER(BfcParams)]
        lib.bfc_build_grid.restype = C.POINTER(BfcCch)
        lib.bfc_free.argtypes = [C.POINTER(BfcCch)]
        lib.bfc_free.restype = None
        lib.bfc_last_error.argtypes = []
        lib.bfc_last_error.restype = C.c_char_p
        _bfcch = lib
    return _bfcch


def ptr(a, typ=u32p):
    return a.ctypes.data_as(typ)
