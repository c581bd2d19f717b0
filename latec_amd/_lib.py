"""ctypes binding of liblatec.so (include/latec.h).

The library is built in-tree by ``latec_amd.csrc.build`` (``__graft_entry__``
``build()``).  There is no fallback: if the shared library is missing or a
call fails, this module raises.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "liblatec.so")

LATEC_OK = 0
LATEC_E_ARG = -1
LATEC_E_HIP = -2
LATEC_E_CAPACITY = -3
LATEC_E_FORMAT = -4
LATEC_E_COUNT = -5


class InterpOpts(C.Structure):
    _fields_ = [
        ("cubic", C.c_int32),
        ("center", C.c_int32),
        ("end_mode", C.c_int32),
        ("round_mode", C.c_int32),
        ("span", C.c_double),
        ("offset", C.c_double),
        ("radius", C.c_uint32),
        ("code_bytes", C.c_int32),
    ]


class Fit(C.Structure):
    _fields_ = [
        ("kexp", C.c_int32),
        ("level", C.c_int32),
        ("shape", C.c_int32),
        ("owx", C.c_int32),
        ("bits", C.c_float),
        ("n_codes", C.c_uint32),
        ("n_out", C.c_uint32),
        ("pad", C.c_uint32),
    ]


class GnnFrame(C.Structure):
    _fields_ = [
        ("pad", C.c_int32),
        ("has_halo", C.c_int32),
        ("origin", C.c_int64 * 8),
        ("tshape", C.c_int64 * 8),
        ("edges", C.c_int64 * 8),
        ("grid", C.c_int64 * 8),
        ("d_coded", C.c_void_p),
    ]


class GnnStageIO(C.Structure):
    _fields_ = [
        ("M", C.c_int64), ("n", C.c_int64),
        ("d_keep", C.c_void_p), ("d_pred", C.c_void_p), ("d_scale", C.c_void_p),
        ("d_gidx", C.c_void_p), ("d_x", C.c_void_p), ("d_recon", C.c_void_p),
        ("d_block", C.c_void_p), ("opts", C.c_void_p), ("cand_opts", C.c_void_p),
        ("gate", C.c_int32), ("interp_kind", C.c_int32), ("reuse_choice", C.c_int32),
        ("pad0", C.c_int32),
        ("choice", C.c_void_p), ("gate_word", C.c_void_p),
        ("eb", C.c_double), ("radius", C.c_uint32), ("pad1", C.c_uint32),
        ("gate_thr", C.c_void_p), ("gate_grid", C.c_void_p), ("level_thr", C.c_void_p),
        ("d_hist", C.c_void_p), ("codes", C.c_void_p), ("levels", C.c_void_p),
        ("outliers", C.c_void_p), ("n_out", C.c_void_p),
    ]


class GnnDecodeStageIO(C.Structure):
    _fields_ = [
        ("M", C.c_int64), ("n", C.c_int64),
        ("d_keep", C.c_void_p), ("d_pred", C.c_void_p), ("d_scale", C.c_void_p),
        ("d_gidx", C.c_void_p), ("d_block", C.c_void_p), ("cand_opts", C.c_void_p),
        ("interp_kind", C.c_int32), ("gate_word", C.c_uint32),
        ("eb", C.c_double), ("level_thr", C.c_void_p), ("levels", C.c_void_p),
    ]


p = C.c_void_p
i64, i32, u32, f64, f32 = C.c_int64, C.c_int, C.c_uint32, C.c_double, C.c_float

# name -> (restype, argtypes); every symbol declared in include/latec.h
SIGNATURES = {
    "latec_last_error": (C.c_char_p, []),
    "latec_abi_version": (i32, []),
    "latec_tune_set": (i32, [C.c_char_p, i64]),
    "latec_tune_get": (i64, [C.c_char_p]),
    "latec_plan_create": (i32, [i32, p, i32, i32, i32, i32, C.POINTER(p)]),
    "latec_plan_destroy": (None, [p]),
    "latec_plan_n_stages": (i32, [p]),
    "latec_plan_stage_info": (i32, [p, i32, p, p, p, p]),
    "latec_plan_class_info": (i32, [p, i32, i32, p]),
    "latec_plan_stage_offset": (i64, [p, i32]),
    "latec_interp_encode": (i32, [p, i32, i32, p, p, p, p, p, p, p, p, p, p]),
    "latec_interp_decode": (i32, [p, i32, i32, p, p, p, p, p, p, p, p]),
    "latec_interp_encode_graph": (i32, [p, p, p, p, p, p, p, p, p, p, p]),
    "latec_interp_predict": (i32, [p, i32, p, p, p, p]),
    "latec_code_hist": (i32, [p, i64, i32, u32, p, p]),
    "latec_hist_words": (i64, [u32]),
    "latec_quantize": (i32, [p, p, i64, f64, p, p, p, p, p, p, p]),
    "latec_dequantize": (i32, [p, p, i64, p, p, p, f64, p, p, p]),
    "latec_level_weights": (i32, [i32, p, p]),
    "latec_model_weights": (i32, [i32, i32, i32, p, p]),
    "latec_neg_log2_weights": (i32, [i32, i32, i32, i32, p, p]),
    "latec_build_cdfs": (i32, [i32, i32, i32, i32, i32, i32, p, p]),
    "latec_fit_const": (i32, [p, i32, i32, p, p, u32, i32, i32, p, p]),
    "latec_fit_codes": (i32, [p, p, i64, u32, i32, i32, p, p]),
    "latec_fit_stages": (i32, [i32, p, p, p, p, u32, i32, i32, p, p]),
    "latec_scale_to_level": (i32, [p, i64, p, p, p]),
    "latec_rans_encode": (i64, [p, i64, p, i64, p, i64]),
    "latec_rans_encode_levels": (i64, [p, p, i64, p, i64, p, i64]),
    "latec_rans_encode_codes_levels": (i64, [p, i32, p, i64, u32, i32, p, i64, p, i64]),
    "latec_code_kexp": (i32, [p, i32, i64, u32]),
    "latec_rans_decode": (i32, [p, i64, p, i64, p, i64]),
    "latec_rans_decode_levels": (i32, [p, i64, p, p, i64, p, i64]),
    "latec_rans_decode_codes_levels": (i32, [p, i64, p, p, i64, u32, i32, p, i32, i64]),
    "latec_pack_stage": (i64, [p, i32, i64, p, i64, u32, i32, p, p, p, i64]),
    "latec_stage_header": (i32, [p, i64, p, p, p, p, p, p, p]),
    "latec_unpack_stage": (i64, [p, i64, u32, p, i64, p, i32, p, p, i64, p]),
    "latec_pack_stages": (i64, [i32, p, i32, p, p, p, p, p, u32, i32, p, p, p, p, i64, p, i32]),
    "latec_pack_stages_framed": (i64, [i32, p, i32, p, p, p, p, p, u32, i32, p, p, p, p, i64, p,
                                       i32, i32, p, i64, p]),
    "latec_zstd_bound": (i64, [i64]),
    "latec_zstd_compress": (i64, [p, i64, i32, p, i64]),
    "latec_zstd_content_size": (i64, [p, i64]),
    "latec_zstd_decompress": (i64, [p, i64, p, i64]),
    "latec_gnn_param_count": (i64, [i32, i32]),
    "latec_gnn_create": (i32, [i32, i32, i32, i32, p, i64, C.POINTER(p)]),
    "latec_gnn_destroy": (None, [p]),
    "latec_gnn_n_lines": (i32, [p]),
    "latec_gnn_geometry": (i32, [p, p, i32, p, p, p, p, p]),
    "latec_gnn_embed": (i32, [p, p, p, p, p, i64, p, p]),
    "latec_gnn_head": (i32, [p, p, i64, f64, f32, f32, p, p, p, p, p]),
    "latec_gnn_finalize": (i32, [p, p, p, p, p, p, i64, f32, f32, p]),
    "latec_gnn_halo": (i32, [p, p, p, p, p, i64, f32, f32, p]),
    "latec_gnn_halo_frame": (i32, [p, p, p, p, p, f32, f32, p]),
    "latec_gnn_gate_costs": (i32, [p, p, p, p, p, i32, i64, f64, u32, p, p, p, p]),
    "latec_gnn_gate_apply": (i32, [p, p, p, i64, f64, f32, i32, p]),
    "latec_gnn_code_stage": (i32, [p, p, i32, p, p]),
    "latec_gnn_decode_stage_begin": (i32, [p, p, i32, p, p]),
    "latec_gnn_decode_stage_end": (i32, [p, p, p, i64, f64, p, p, p]),
    "latec_gather_f32": (i32, [p, p, i64, p, p]),
    "latec_scatter_f32": (i32, [p, p, i64, p, p]),
    "latec_train_gather": (i32, [p, i64, p, i64, p, p]),
    "latec_train_scatter_add": (i32, [p, i64, p, i64, p, i64, p, p]),
    "latec_train_write_rows": (i32, [p, i64, p, i64, p, p]),
    "latec_train_gemm": (i32, [i64, i64, i64, p, i64, i64, p, i64, i64, p, i64, p, i32, p, i64,
                               p]),
    "latec_train_gemm_bf16": (i32, [i64, i64, i64, p, i64, i64, p, i64, i64, p, i64, p, i32, p, i64,
                               p]),
    "latec_train_colsum": (i32, [i64, i64, p, p, i32, p, i64, p]),
    "latec_train_gelu": (i32, [p, p, i64, p]),
    "latec_train_gelu_bwd": (i32, [p, p, p, i64, p]),
    "latec_train_attn": (i32, [i64, p, p, p, i32, f32, p, p, p, p, p]),
    "latec_train_attn_bwd": (i32, [i64, p, p, i32, p, p, p, p, p, p, p, p, i64, p]),
    "latec_train_nll": (i32, [i64, p, p, p, f32, p, p, p, p]),
    "latec_mlrans_lanes": (i64, [i64]),
    "latec_mlrans_scratch_words": (i64, [i32, p]),
    "latec_mlrans_encode": (i32, [i32, p, i32, p, p, p, u32, p, p, p, i64, p, i64, p, p, p]),
    "latec_mlrans_decode": (i32, [i32, p, i64, p, p, p, p, u32, p, p, p, i32, p, p, i64, p]),
    "latec_mlrans_noalloc_words": (i64, []),
    "latec_interp_stats_words": (i64, [p]),
    "latec_minmax_words": (i64, [i32]),
    "latec_minmax": (i32, [i32, p, p, p, p, p, i64, p]),
    "latec_order_outliers": (i32, [i32, p, i32, p, p, p, p, p, p, p, i64, p]),
    "latec_zero_ranks_words": (i64, [i32, p]),
    "latec_interp_stats": (i32, [p, p, p, p, p, p, i64, p]),
    "latec_mlrans_encode_noalloc": (i32, [i32, p, i32, p, p, p, u32, p, p, p, i64, p, p, p]),
    "latec_zero_ranks": (i32, [i32, p, i32, p, p, p, p, p, i64, p]),
    "latec_batch_create": (i32, [i32, p, p, p, p, p, p, p, p, i32, p, p, p, p, p,
                                 C.POINTER(p)]),
    "latec_batch_n_steps": (i32, [p]),
    "latec_batch_run": (i32, [p, i32, p]),
    "latec_batch_destroy": (None, [p]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m latec_amd.csrc.build` "
                "(the HIP path has no CPU fallback)"
            )
        # torch's wheel ships its own HIP runtime, which its libraries load
        # as "libamdhip64.so"; loaded first, this library would pull in
        # /opt/rocm's copy under "libamdhip64.so.7" and torch a second one,
        # and torch's device pointers would mean nothing to this library's
        # runtime.  Importing torch first makes both use torch's copy.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().latec_last_error().decode("utf-8", "replace")


def check(rc: int, what: str = "") -> int:
    """Raise ValueError (argument/format errors) or RuntimeError (HIP errors)."""
    if rc < 0:
        msg = last_error() or what
        if rc == LATEC_E_HIP:
            raise RuntimeError(msg)
        raise ValueError(msg)
    return rc


def ptr(a) -> C.c_void_p:
    """Data pointer of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return C.c_void_p(0)
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def stream_ptr(stream=None) -> C.c_void_p:
    """hipStream_t of a torch stream (None -> torch's current stream on the
    current device, read without building a Stream object: the per-stage
    library calls make thousands of these)."""
    import torch

    if stream is not None:
        return C.c_void_p(stream.cuda_stream)
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None and torch.cuda.is_initialized():
        return C.c_void_p(raw(torch._C._cuda_getDevice()))
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)
