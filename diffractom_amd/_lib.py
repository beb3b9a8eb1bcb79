"""ctypes binding of libdiffractom_hip.so (the C ABI declared in include/diffractom_hip.h).

The library is built in-tree (diffractom_amd/build.py). There is no fallback: without the library, or without a
GPU, every operator call raises. Status codes map to the exception types the reference raises for the same
conditions (ValueError, NotImplementedError, MemoryError) and RuntimeError for launch failures.
"""
from __future__ import annotations

import ctypes
from ctypes import c_float, c_int32, c_int64, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libdiffractom_hip.so"

P = c_void_p
I32 = c_int32
I64 = c_int64
F32 = c_float

# name -> argument types (the last argument of every kernel entry point is the stream)
SIGNATURES = {
    "dfm_gather_k": [P, P, I64, I64, I32, I32, P],
    "dfm_scatter_k": [P, P, I64, I64, I32, I32, P],
    "dfm_radon_forward": [P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, F32, I32, P],
    "dfm_radon_backward": [P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, I32, P],
    "dfm_radon_backward_spdhg": [P, P, I32, I32, I32, I32, I32, I32, I32, I32, I32, P, P, P, F32, F32, P],
    "dfm_radon_forward_table": [P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, P],
    "dfm_radon_backward_table": [P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, P, P, P, F32, F32, I32, I32,
                                 P],
    "dfm_pf_eval_dense": [P, P, P, P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, I32, F32, I32, P],
    "dfm_pf_count_fwd": [P, P, I32, I32, I32, I32, P],
    "dfm_pf_count_adj": [P, P, I32, I32, I32, I32, P],
    "dfm_pf_fill_fwd": [P, P, P, P, I32, I32, I32, I32, P],
    "dfm_pf_fill_adj": [P, P, P, P, I32, I32, I32, I32, P],
    "dfm_spmm_forward": [P, P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, P],
    "dfm_spmm_forward_split": [P, P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, P, I32, I32, P],
    "dfm_smooth_tv": [P, P, I32, I32, I32, F32, P, P, P],
    "dfm_radon_forward_sweep": [P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, I32, I32, P],
    "dfm_radon_sweep_lines": [I32, I32],
    "dfm_radon_forward_pairs_sweep": [P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, I32, I32, P],
    "dfm_radon_forward_band": [P, P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, P],
    "dfm_spmm_t": [P, P, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, F32, I32, I32, P],
    "dfm_spmm_t_rows_per_block": [I32],
    "dfm_spmm_t_direct": [P, I64, I32, P, P, P, P, I64, I32, P, I32, I32, I32, I32, I32, I32, F32, I32, I32, P],
    "dfm_transpose_sum": [P, P, P, I32, I32, I32, I32, I64, I64, I64, I64, I32, I64, I32, P],
    "dfm_spmm_adjoint": [P, P, P, P, P, P, I32, I32, I32, I32, I32, F32, I32, P],
    "dfm_zero_rows": [P, P, I32, I64, P],
    "dfm_dual_update_sub": [P, P, P, P, P, I64, P, I64, P, I64, F32, P, I32, I64, P],
    "dfm_spdhg_update": [P, P, P, P, F32, F32, I64, P],
    "dfm_tinit": [P, I64, I32, P],
    "dfm_tmin": [P, P, P, I64, I32, I32, F32, P],
    "dfm_tfinal": [P, I64, I32, I32, P],
    "dfm_dual_steps": [P, P, P, I64, F32, I64, P],
    "dfm_set_mask_ones": [P, P, I64, I64, P],
    "dfm_support_ones_k": [P, P, I64, I32, I32, P],
    "dfm_project_support_k": [P, P, I64, I32, I32, P],
    "dfm_deviance_partial": [P, P, P, I64, P, I64, P, I64, F32, I64, P, P, P],
    "dfm_mean_counted_partial": [P, P, I64, I64, P, P, P],
    "dfm_deviance_noalloc": [P, P, P, I64, P, I64, P, I64, F32, I64, P, P],
    "dfm_poisson_em_ratio": [P, P, P, I64, P, I64, P, I64, F32, I64, P, P],
    "dfm_poisson_weight_over_gain": [P, I64, P, I64, I64, P, P],
    "dfm_em_update": [P, P, P, P, I64, F32, I64, P],
    "dfm_pdhg_primal_update": [P, P, P, P, F32, I64, P],
    "dfm_average": [P, P, I64, P],
    "dfm_fista_value_grad_partial": [P, P, P, I64, P, I64, P, I64, F32, I32, I32, F32, I64, P, P, P],
    "dfm_fista_curvature_partial": [P, P, P, P, I64, P, I64, P, I64, F32, I32, I64, P, P, P],
    "dfm_fista_sums_partial": [P, P, I64, P, P, I64, I32, I64, P, P, P],
    "dfm_fista_elementwise": [I32, P, P, P, P, P, I64, F32, F32, I64, P],
    "dfm_fista_step_sums_partial": [P, P, P, P, I64, P, P, P],
    "dfm_dot_partial": [P, P, I64, P, P, P],
    "dfm_scale": [P, F32, I64, P],
    "dfm_tv_dual_k": [P, P, P, P, P, P, I32, I32, I32, F32, F32, P],
    "dfm_tv_adjoint_update_k": [P, P, P, P, P, I32, I32, I32, F32, F32, P],
    "dfm_tv_value_partial": [P, P, I32, I32, I32, P, P, P],
    "dfm_mfma_pack_fwd": [P, P, P, I32, I32, I32, I32, I32, I32, P],
    "dfm_mfma_pack_adj": [P, P, P, I32, I32, I32, I32, I32, I32, P],
    "dfm_mfma_forward": [P, I32, I32, P, P, P, I32, I32, I32, I32, I32, P, P, I32, I32, I32, P],
    "dfm_mfma_grouped": [I32, P, P, I64, I32, I32, P],
    "dfm_mfma_adjoint": [P, P, P, P, I32, I32, I32, I32, I32, P, I32, F32, I32, P],
    "dfm_pfblk_count": [P, P, I32, I32, I32, I32, P],
    "dfm_pfblk_fill": [P, P, P, P, I32, P, I32, I32, I32, I32, I32, P],
    "dfm_pfblk_forward": [P, I32, I32, P, P, P, I32, P, I32, P, I32, I32, I32, P, P, I32, I32, P],
    "dfm_pfblk_adjoint": [P, P, P, P, I32, P, I32, P, I32, I32, I32, I32, P, I32, F32, P],
    "dfm_pfblk_grouped": [I32, P, P, I64, I32, P],
    "dfm_pfblk_desc_size": [],
    "dfm_pfgen": [I32, P, P, P, P, P, P, P, I32, P, P, P, P, I32, I32, I32, I32, I32, I32, I32, I32, F32,
                  ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, P, P, P, P, I32, P, P],
    "dfm_pf_grad": [P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, I32, P, P, P, I32, I32, I32, I32, I32, I32, I32,
                    I32, I32, F32, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, P, P, P],
    "dfm_pf_grad_reduce": [P, P, I32, I32, I32, I64, I32, P, P, P],
    "dfm_angle_dots": [P, P, I32, I32, I32, I32, P, P],
    "dfm_nufft_pad": [P, I32, I32, I32, P, P, I64, P, P],
    "dfm_nufft_interp": [P, P, P, P, P, P, I32, I32, I32, P, P],
    "dfm_nufft_spread": [P, P, P, P, P, P, I32, I32, I64, P, P],
    "dfm_nufft_crop": [P, P, P, I64, I32, F32, P, I32, I32, I32, P],
    "dfm_spdhg_update_2d": [P, P, P, P, I64, I32, I32, I32, F32, F32, P],
    "dfm_dense_f32": [I32, P, I64, I32, I32, P, I64, I32, P, I64, I32, I32, P, I32, I32, I32, I32, F32, I32, I32, P,
                      P],
    "dfm_local_strips": [P, P, P, P, P, P, I32, P, P, I32, I32, I32, P, P, I32, F32, P, P],
    "dfm_local_assemble": [P, I32, P, P, P, I32, P, I32, P, I32, I32, I32, P, P],
    "dfm_local_extract": [P, I32, I32, I32, P, P, I32, P, P, P, I32, P, P],
    "dfm_local_backproject": [P, I32, P, P, P, I32, I32, P, P, I32, P, P, P, P, I64, I64, P, I32, I32, I32, P, P, P,
                              P, F32, F32, P],
    "dfm_tv_dual_nbr": [P, P, P, P, P, P, P, P, F32, F32, I64, P],
    "dfm_tv_adjoint_update_nbr": [P, P, P, P, P, P, P, F32, F32, I64, P],
    "dfm_tv_value_nbr_partial": [P, P, P, P, I64, P, P, P],
}

DFM_ERR_ARG, DFM_ERR_UNSUPPORTED, DFM_ERR_LAUNCH, DFM_ERR_MEMORY = -1, -2, -3, -4

_lib = None


def load():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH.name} is not built: run `python -m diffractom_amd.build` "
                               "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.argtypes = args
            fn.restype = c_int32
        lib.dfm_last_error.restype = ctypes.c_char_p
        lib.dfm_last_error.argtypes = []
        lib.dfm_version.restype = c_int32
        _lib = lib
    return _lib


def exported_symbols():
    """Names of SIGNATURES that the library exports."""
    lib = load()
    return [n for n in SIGNATURES if hasattr(lib, n)]


def call(name, *args):
    """Call an entry point; a non-zero status raises the matching exception with the library's message."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != 0:
        msg = lib.dfm_last_error().decode(errors="replace")
        if st == DFM_ERR_ARG:
            raise ValueError(msg)
        if st == DFM_ERR_UNSUPPORTED:
            raise NotImplementedError(msg)
        if st == DFM_ERR_MEMORY:
            raise MemoryError(msg)
        raise RuntimeError(msg)
    return st


MAX_PARTIALS = 2048   # DFM_MAX_PARTIALS of the header: partial sums a reduction entry point writes at most


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    return None if t is None else c_void_p(t.data_ptr())


def stream_handle(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return c_void_p(s.cuda_stream)
