"""ctypes binding of libhipfem.so (the C ABI declared in include/hipfem.h).

The library is loaded after ``import torch`` so that it binds to the HIP
runtime torch already loaded (same SONAME), sharing streams and allocations.
There is no fallback: a missing library or a failed call raises.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhipfem.so")

c_i32, c_i64, c_f64, c_sz = ctypes.c_int32, ctypes.c_int64, ctypes.c_double, ctypes.c_size_t
vp = ctypes.c_void_p

HF_MAT_POISSON, HF_MAT_LINEAR_ELASTIC, HF_MAT_NEO_HOOKEAN = 0, 1, 2
HF_PARAM_NONE, HF_PARAM_SCALE_E, HF_PARAM_SIMP, HF_PARAM_MU, HF_PARAM_SCALE_K = 0, 1, 2, 3, 4
HF_KRYLOV_PCG, HF_KRYLOV_BICGSTAB = 0, 1
HF_MG_JACOBI, HF_MG_CHEBYSHEV = 0, 1
HF_MG_FINE_FULL, HF_MG_FINE_SELL, HF_MG_FINE_F16 = 1, 2, 4
HF_MG_SKIP_FINE_POST = 1
HF_MG_SOLVE_EAGER = 4
HF_MG_SOLVE_SPLIT_SCALARS = 8


class HFMaterial(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("param_mode", c_i32), ("E", c_f64), ("nu", c_f64),
                ("E0", c_f64), ("Emin", c_f64), ("penal", c_f64), ("kappa", c_f64), ("k", c_f64),
                ("quad1d", c_i32), ("pad_", c_i32)]


class HFKrylovInfo(ctypes.Structure):
    _fields_ = [("iterations", c_i32), ("status", c_i32), ("res_norm", c_f64), ("b_norm", c_f64)]


class HFMmaInfo(ctypes.Structure):
    _fields_ = [("newton_steps", c_i32), ("trials", c_i32), ("res_max", c_f64), ("res_norm", c_f64)]


class HipFemError(RuntimeError):
    pass


_SIGS = {
    "hf_last_error": (c_i32, [ctypes.c_char_p, c_sz]),
    "hf_version": (ctypes.c_char_p, []),
    "hf_pattern_build": (c_i32, [vp, c_i64, c_i32, c_i64, c_i32, vp, ctypes.POINTER(vp)]),
    "hf_pattern_sizes": (c_i32, [vp, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)]),
    "hf_pattern_arrays": (c_i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]),
    "hf_pattern_csr": (c_i32, [vp, vp, vp, vp]),
    "hf_pattern_destroy": (c_i32, [vp]),
    "hf_cell_residual_f64": (c_i32, [c_i32, ctypes.POINTER(HFMaterial), vp, vp, c_i64, vp, vp, vp, vp]),
    "hf_hex8_box_geometry_bytes": (c_sz, []),
    "hf_hex8_box_geometry": (c_i32, [vp, vp, c_i64, vp, vp]),
    "hf_cell_residual_box_f64": (c_i32, [ctypes.POINTER(HFMaterial), vp, vp, c_i64, vp, vp, vp, vp]),
    "hf_cell_residual_f32": (c_i32, [c_i32, ctypes.POINTER(HFMaterial), vp, vp, c_i64, vp, vp, vp, vp]),
    "hf_cell_jacobian_f64": (c_i32, [c_i32, ctypes.POINTER(HFMaterial), vp, vp, c_i64, vp, vp, vp, vp]),
    "hf_cell_jacobian_f32": (c_i32, [c_i32, ctypes.POINTER(HFMaterial), vp, vp, c_i64, vp, vp, vp, vp]),
    "hf_cell_param_vjp_f64": (c_i32, [c_i32, ctypes.POINTER(HFMaterial), vp, vp, c_i64, vp, vp, vp, vp, vp]),
    "hf_generic_ugrad_f64": (c_i32, [vp, c_i64, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, vp, vp, vp]),
    "hf_generic_residual_f64": (c_i32, [c_i64, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, vp, vp, vp, vp]),
    "hf_generic_jacobian_f64": (c_i32, [c_i64, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, vp, vp, vp, vp]),
    "hf_generic_jacobian_f32": (c_i32, [c_i64, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, vp, vp, vp, vp]),
    "hf_gather_residual_f64": (c_i32, [vp, vp, vp, vp, vp]),
    "hf_gather_residual_box_f64": (c_i32, [vp, c_i32, c_i32, c_i32, vp, vp, vp, vp, vp]),
    "hf_matrix_values": (c_i64, [vp]),
    "hf_assemble_f64": (c_i32, [vp, vp, vp, vp]),
    "hf_pattern_max_row_blocks": (c_i32, [vp]),
    "hf_assemble_hex8_f64": (c_i32, [vp, vp, vp, vp]),
    "hf_assemble_hex8_f32": (c_i32, [vp, vp, vp, vp]),
    "hf_assemble_f32": (c_i32, [vp, vp, vp, vp]),
    "hf_matrix_to_csr_f64": (c_i32, [vp, vp, vp, vp]),
    "hf_bc_rows_f64": (c_i32, [vp, vp, c_i64, vp, vp]),
    "hf_bc_rows_f32": (c_i32, [vp, vp, c_i64, vp, vp]),
    "hf_bc_residual_f64": (c_i32, [vp, vp, c_i64, c_f64, vp, vp, vp]),
    "hf_bc_initial_guess_f64": (c_i32, [vp, vp, c_i64, vp, vp, c_i64, vp]),
    "hf_spmv_f64": (c_i32, [vp, vp, vp, vp, vp]),
    "hf_spmv_f32": (c_i32, [vp, vp, vp, vp, vp]),
    "hf_spmv_t_f64": (c_i32, [vp, vp, vp, vp, vp]),
    "hf_spmv_t_f32": (c_i32, [vp, vp, vp, vp, vp]),
    "hf_jacobi_f64": (c_i32, [vp, vp, vp, vp]),
    "hf_jacobi_f32": (c_i32, [vp, vp, vp, vp]),
    "hf_krylov_workspace_size": (c_sz, [c_i32, c_i64, c_i32]),
    "hf_krylov_solve_f64": (c_i32, [c_i32, vp, vp, vp, vp, vp, c_i32, c_i32, c_f64, c_f64, c_i32, vp,
                                    ctypes.POINTER(HFKrylovInfo), vp]),
    "hf_krylov_solve_f32": (c_i32, [c_i32, vp, vp, vp, vp, vp, c_i32, c_i32, c_f64, c_f64, c_i32, vp,
                                    ctypes.POINTER(HFKrylovInfo), vp]),
    "hf_init": (c_i32, []),
    "hf_krylov_solve_noalloc_f64": (c_i32, [c_i32, vp, vp, vp, vp, vp, c_i32, c_i32, c_f64, c_f64, c_i32,
                                            ctypes.POINTER(HFKrylovInfo), vp]),
    "hf_profile_enable": (c_i32, [c_i32]),
    "hf_profile_read": (c_i32, [ctypes.POINTER(c_i64), ctypes.POINTER(c_f64)]),
    "hf_cnn_conv3d_fwd_f32": (c_i32, [vp, c_i32, vp, vp, c_i32, c_i32, c_i32, c_i32, c_i32, c_i32, vp, vp]),
    "hf_cnn_conv3d_bwd_data_f32": (c_i32, [vp, c_i32, vp, c_i32, c_i32, c_i32, c_i32, c_i32, vp, vp]),
    "hf_cnn_bwd_weight_workspace": (c_sz, [c_i32, c_i32, c_i32, c_i32, c_i32, c_i32]),
    "hf_cnn_conv3d_bwd_weight_f32": (c_i32, [vp, c_i32, vp, c_i32, c_i32, c_i32, c_i32, c_i32, vp, vp, vp, vp]),
    "hf_cnn_tanh_bwd_f32": (c_i32, [vp, vp, c_i64, vp, vp]),
    "hf_lpc_features_f32": (c_i32, [vp, vp, vp, c_i64, c_i32, vp, vp, vp]),
    "hf_lpc_output_f32": (c_i32, [vp, vp, vp, vp, c_i64, c_i32, vp, vp, vp]),
    "hf_lpc_apply_workspace_size": (c_sz, [c_i64]),
    "hf_lpc_apply_f32": (c_i32, [vp, vp, vp, vp, vp, c_i32, c_i32, c_i32, vp, vp, vp]),
    "hf_lpc_workspace_size": (c_sz, [c_i64]),
    "hf_lpc_solve_f32": (c_i32, [vp, vp, vp, vp, vp, vp, vp, vp, c_i32, c_i32, c_i32, c_i32, c_f64, c_f64, c_i32,
                                 c_i32, c_i32, c_i32, vp, ctypes.POINTER(HFKrylovInfo), vp]),
    "hf_mg_create": (c_i32, [c_i32, c_i32, c_i32, c_i32, ctypes.POINTER(vp)]),
    "hf_mg_destroy": (c_i32, [vp]),
    "hf_mg_levels": (c_i32, [vp, ctypes.POINTER(c_i32), vp, c_i32]),
    "hf_mg_set_smoother": (c_i32, [vp, c_i32, c_i32, ctypes.c_float]),
    "hf_mg_set_post_degree": (c_i32, [vp, c_i32]),
    "hf_mg_set_cell_operator": (c_i32, [vp, vp, vp]),
    "hf_mg_assemble_cellop_f32": (c_i32, [vp, vp, vp]),
    "hf_mg_set_learned": (c_i32, [vp, c_i32, vp, vp, vp]),
    "hf_mg_bounds": (c_i32, [vp, vp, c_i32]),
    "hf_mg_setup_f32": (c_i32, [vp, vp, vp, vp, ctypes.POINTER(c_i32), vp]),
    "hf_mg_apply_f32": (c_i32, [vp, vp, vp, c_i32, vp]),
    "hf_mg_copy_stencil": (c_i32, [vp, c_i32, vp, vp]),
    "hf_mg_workspace_size": (c_sz, [vp]),
    "hf_mg_solve_f32": (c_i32, [vp, vp, vp, vp, vp, c_i32, c_i32, c_f64, c_f64, c_i32, vp,
                                ctypes.POINTER(HFKrylovInfo), vp]),
    "hf_mg_stencil_size": (c_sz, [vp]),
    "hf_mg_assemble_hex8_f32": (c_i32, [vp, vp, vp, c_i64, vp, vp, vp, vp]),
    "hf_hex8_jacobian_upper_f32": (c_i32, [vp, vp, vp, c_i64, vp, vp, vp, vp]),
    "hf_mg_setup_stencil_f32": (c_i32, [vp, vp, vp, vp, ctypes.POINTER(c_i32), vp]),
    "hf_mg_solve_stencil_f32": (c_i32, [vp, vp, vp, vp, c_i32, c_i32, c_f64, c_f64, c_i32, vp,
                                        ctypes.POINTER(HFKrylovInfo), vp]),
    "hf_mg_stencil_matvec_f32": (c_i32, [vp, vp, vp, vp, vp]),
    "hf_mma_sensitivity_filter_f64": (c_i32, [vp, vp, vp, c_i64, c_i32, vp, vp, c_i32, vp, vp]),
    "hf_mma_density_filter_f64": (c_i32, [vp, vp, vp, c_i64, c_i32, vp, vp, vp]),
    "hf_mma_create": (c_i32, [c_i64, c_i32, ctypes.POINTER(vp)]),
    "hf_mma_destroy": (c_i32, [vp]),
    "hf_mma_update_f64": (c_i32, [vp, c_i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, c_f64,
                                  ctypes.POINTER(c_f64), ctypes.POINTER(c_f64), ctypes.POINTER(c_f64), c_f64, vp,
                                  ctypes.POINTER(HFMmaInfo), vp]),
    "hf_dot_workspace_size": (c_sz, []),
    "hf_dot_f64": (c_i32, [vp, vp, c_i64, vp, ctypes.POINTER(c_f64), vp]),
    "hf_dot_f32": (c_i32, [vp, vp, c_i64, vp, ctypes.POINTER(c_f64), vp]),
}

_lib = None


def declared_symbols():
    """Entry points declared in include/hipfem.h that this binding uses."""
    return sorted(_SIGS)


def load(path=LIB_PATH):
    """Load libhipfem.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HipFemError("libhipfem.so not found at %s; build it with "
                          "`python jax-fem_amd/build.py`" % path)
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().hf_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def call(name, *args):
    rc = getattr(load(), name)(*args)
    if rc != 0:
        raise HipFemError("%s failed (%d): %s" % (name, rc, last_error()))
    return rc


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "expected a contiguous device tensor"
    return ctypes.c_void_p(t.data_ptr())


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream(dev=None):
    """The current HIP stream of `dev` (default: the current device) as a handle."""
    if _raw_stream is not None and dev is None:
        # the raw handle of torch's current stream, without building a Stream object
        return ctypes.c_void_p(_raw_stream(torch._C._cuda_getDevice()))
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
