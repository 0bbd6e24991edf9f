"""ctypes binding of ``libcasadi_hip.so`` (the C ABI declared in include/casadi_hip.h).

The shared library is built in-tree by ``casadi_amd/csrc/Makefile``.  There is
no fallback: importing this module fails loudly when the library is missing.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcasadi_hip.so")
# Code objects of the specialised kernels are cached by hiprtc (comgr) on disk;
# keep that cache beside the package unless the caller chose a location.
os.environ.setdefault("AMD_COMGR_CACHE_DIR", os.path.join(os.path.dirname(_HERE), "_build", "comgr"))


class CasadiHipError(RuntimeError):
    """Error reported by the native library."""


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C casadi_amd/csrc` "
            "(or __graft_entry__.build()) before importing casadi_amd")
    return C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)


lib = _load()

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
U64 = C.c_uint64
D = C.c_double
PI64 = C.POINTER(C.c_int64)
PU64 = C.POINTER(C.c_uint64)
PD = C.POINTER(C.c_double)
PP = C.POINTER(C.c_void_p)
CSTR = C.c_char_p


def _fn(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


# --- errors ------------------------------------------------------------------
cxh_last_error = _fn("cxh_last_error", CSTR)


def check(rc, what=""):
    if rc < 0:
        msg = cxh_last_error().decode(errors="replace")
        raise CasadiHipError(f"{what}: {msg}" if what else msg)
    return rc


# --- GPU boundary --------------------------------------------------------------
cxh_tape_create = _fn("cxh_tape_create", I32, P, I64, I64, I32, PI64, I32, PI64, PP)
cxh_tape_destroy = _fn("cxh_tape_destroy", I32, P)
cxh_tape_config = _fn("cxh_tape_config", I32, P, I32, I32, I32)
cxh_tape_ipt = _fn("cxh_tape_ipt", I32, P, I32)
cxh_tape_jit_async = _fn("cxh_tape_jit_async", I32, P, I32)
cxh_tape_jit_pending = _fn("cxh_tape_jit_pending", I32, P)
cxh_tape_jit_fallback = _fn("cxh_tape_jit_fallback", I32, P, C.c_char_p, I64)
cxh_tape_tune = _fn("cxh_tape_tune", I32, P, I64, PP, PP, I32, I64, I32, P, C.POINTER(I32), C.POINTER(C.c_double))
cxh_map_eval = _fn("cxh_map_eval", I32, P, I64, PP, PP, I32, I64, P)
cxh_map_eval_host = _fn("cxh_map_eval_host", I32, P, I64, PP, PP)
cxh_map_fwd_adj = _fn("cxh_map_fwd_adj", I32, P, I64, PP, PP, PP, PP, I32, I64, I64, P)
cxh_tape_regularity_check = _fn("cxh_tape_regularity_check", I32, P, I32, P, I64)
cxh_tape_io = _fn("cxh_tape_io", I32, P, C.POINTER(I32), C.POINTER(I32), PI64, PI64)
cxh_tape_deserialize = _fn("cxh_tape_deserialize", I32, C.c_char_p, I64, C.POINTER(P))
cxh_map_eval_sum = _fn("cxh_map_eval_sum", I32, P, I64, PP, PP, C.POINTER(I32), C.POINTER(I32), I64, P)
cxh_map_eval_sum_noalloc = _fn("cxh_map_eval_sum_noalloc", I32, P, I64, PP, PP, C.POINTER(I32), C.POINTER(I32), I64,
                               P)
cxh_sp_forward = _fn("cxh_sp_forward", I32, P, I64, PP, PP, I64, P)
cxh_tape_create_nested = _fn("cxh_tape_create_nested", I32, P, I64, I64, I32, PI64, I32, PI64, I32, PP, PP, PP,
                             PP)
cxh_tape_instructions = _fn("cxh_tape_instructions", I32, P, P, I64, PI64, PI64)
cxh_sp_reverse = _fn("cxh_sp_reverse", I32, P, I64, PP, PP, I64, P)
cxh_map_fwd_adj_sum = _fn("cxh_map_fwd_adj_sum", I32, P, I64, PP, PP, PP, PP, C.POINTER(I32), C.POINTER(I32),
                          I64, I64, P)
cxh_ckpt_size = _fn("cxh_ckpt_size", I32, P, I64, I64, I32, PI64)
cxh_map_fwd_ckpt = _fn("cxh_map_fwd_ckpt", I32, P, I64, PP, PP, I32, I64, I64, P, P)
cxh_map_rev_ckpt = _fn("cxh_map_rev_ckpt", I32, P, I64, PP, PP, PP, I32, I64, I64, P, P)
cxh_jac_create = _fn("cxh_jac_create", I32, P, I32, I64, PI64, PI64, I64, PI64, PI64, PP)
cxh_jac_destroy = _fn("cxh_jac_destroy", I32, P)
cxh_map_jac = _fn("cxh_map_jac", I32, P, I64, PP, PP, P, I32, I64, P)
cxh_device_info = _fn("cxh_device_info", I32, C.POINTER(I32), C.POINTER(I32), C.c_char_p, I32)
cxh_tape_engine = _fn("cxh_tape_engine", I32, P, I32)
cxh_tape_jit_prepare = _fn("cxh_tape_jit_prepare", I32, P, I64, I32, I32)
cxh_tape_jit_source = _fn("cxh_tape_jit_source", I32, P, I64, I32, I32, C.c_char_p, I64, PI64,
                          C.POINTER(I32))
cxh_tape_jit_plan = _fn("cxh_tape_jit_plan", I32, P, I64, PI64, PI64)
cxh_tape_roll_plan = _fn("cxh_tape_roll_plan", I32, P, I64, PI64)
cxh_tape_roll_source = _fn("cxh_tape_roll_source", I32, P, I64, I32, C.c_char_p, I64, PI64)
cxh_jit_compile_check = _fn("cxh_jit_compile_check", I32, CSTR, CSTR)
cxh_jac_jit_source = _fn("cxh_jac_jit_source", I32, P, I32, C.c_char_p, I64, PI64, C.POINTER(I32),
                         C.POINTER(I32))

# optional groups (linear solvers, heterogeneous batches)
def _opt(name, res, *args):
    try:
        return _fn(name, res, *args)
    except AttributeError:
        return None


cxh_ldl_create = _opt("cxh_ldl_create", I32, PI64, PI64, PI64, PP)
cxh_ldl_destroy = _opt("cxh_ldl_destroy", I32, P)
cxh_ldl_factor_solve = _opt("cxh_ldl_factor_solve", I32, P, I64, P, P, I32, P, I32, P)
cxh_ldl_engine = _opt("cxh_ldl_engine", I32, P, I32)
cxh_ldl_jit_source = _opt("cxh_ldl_jit_source", I32, P, I32, C.c_char_p, I64, PI64)
cxh_qr_engine = _opt("cxh_qr_engine", I32, P, I32)
cxh_qr_jit_source = _opt("cxh_qr_jit_source", I32, P, I32, I32, I32, C.c_char_p, I64, PI64, C.POINTER(I32))
cxh_qr_create = _opt("cxh_qr_create", I32, PI64, PI64, PI64, PI64, PI64, D, PP)
cxh_qr_destroy = _opt("cxh_qr_destroy", I32, P)
cxh_qr_factor_solve = _opt("cxh_qr_factor_solve", I32, P, I64, P, P, I32, I32, P, I32, P)
cxh_ldl_nfact = _opt("cxh_ldl_nfact", I32, P, I64, P, P, I32, P)
cxh_ldl_solve = _opt("cxh_ldl_solve", I32, P, I64, P, I32, I32, I32, P)
cxh_ldl_neig = _opt("cxh_ldl_neig", I32, P, P, P)
cxh_ldl_rank = _opt("cxh_ldl_rank", I32, P, P, P)
cxh_qr_nfact = _opt("cxh_qr_nfact", I32, P, I64, P, P, I32, P)
cxh_qr_solve = _opt("cxh_qr_solve", I32, P, I64, P, I32, I32, I32, P)
cxh_qr_det = _opt("cxh_qr_det", I32, P, P, P)
cxh_qr_cache = _opt("cxh_qr_cache", I32, P, I32)
cxh_batch_create = _opt("cxh_batch_create", I32, I32, PP, PI64, PP)
cxh_batch_bind = _opt("cxh_batch_bind", I32, P, I32, PP, PP, PP, PP, I32, I64)
cxh_batch_config = _opt("cxh_batch_config", I32, P, I32, I32, I32)
cxh_batch_tune = _opt("cxh_batch_tune", I32, P, P, I32, C.POINTER(I32), C.POINTER(D))
cxh_batch_launch = _opt("cxh_batch_launch", I32, P, P)
cxh_batch_launch_best = _opt("cxh_batch_launch_best", I32, P, P)
cxh_batch_prefers_graph = _opt("cxh_batch_prefers_graph", I32, P)
cxh_batch_launch_direct = _opt("cxh_batch_launch_direct", I32, P, P)
cxh_batch_info = _opt("cxh_batch_info", I32, P, C.POINTER(I32), PI64)
cxh_batch_destroy = _opt("cxh_batch_destroy", I32, P)
cxh_mapaccum_create = _fn("cxh_mapaccum_create", I32, P, I64, I32, PP)
cxh_mapaccum_destroy = _fn("cxh_mapaccum_destroy", I32, P)
cxh_mapaccum_eval = _fn("cxh_mapaccum_eval", I32, P, I64, PP, PP, I32, I64, P)
cxh_mapaccum_fwd_adj = _fn("cxh_mapaccum_fwd_adj", I32, P, I64, PP, PP, PP, PP, I32, I64, I64, P)
cxh_mapaccum_source = _fn("cxh_mapaccum_source", I32, P, I32, I64, C.c_char_p, I64, PI64)
cxh_mapaccum_ckpt_size = _fn("cxh_mapaccum_ckpt_size", I32, P, I64, I64, PI64)
cxh_mapaccum_fwd_ckpt = _fn("cxh_mapaccum_fwd_ckpt", I32, P, I64, PP, PP, I32, I64, I64, P, P)
cxh_mapaccum_rev_ckpt = _fn("cxh_mapaccum_rev_ckpt", I32, P, I64, PP, PP, PP, I32, I64, I64, P, P)
cxh_batch_source = _opt("cxh_batch_source", I32, I32, PP, C.c_char_p, I64, PI64)

# --- host symbolic layer ------------------------------------------------------
cxs_sym = _fn("cxs_sym", U64, CSTR)
cxs_const = _fn("cxs_const", U64, D)
cxs_unary = _fn("cxs_unary", U64, I32, U64)
cxs_binary = _fn("cxs_binary", U64, I32, U64, U64)
cxs_neg = _fn("cxs_neg", U64, U64)
cxs_if_else = _fn("cxs_if_else", U64, U64, U64, U64)
cxs_unary_n = _fn("cxs_unary_n", I32, I32, I64, PU64, PU64)
cxs_binary_n = _fn("cxs_binary_n", I32, I32, I64, PU64, I64, PU64, I64, PU64)
cxs_node_info = _fn("cxs_node_info", I32, U64, C.POINTER(I32), PD, PU64, PU64)
cxs_sym_name = _fn("cxs_sym_name", CSTR, U64)
cxs_is_equal = _fn("cxs_is_equal", I32, U64, U64, I32)
cxs_node_count = _fn("cxs_node_count", I64)

cxs_function_new = _fn("cxs_function_new", I32, CSTR, I32, PP, PP, PP, I32, PP, PP, PP, I32, PP)
cxs_function_free = _fn("cxs_function_free", I32, P)
cxs_function_info = _fn("cxs_function_info", I32, P, C.POINTER(I32), C.POINTER(I32), PI64, PI64, PI64)
cxs_function_name = _fn("cxs_function_name", CSTR, P)
cxs_function_name_in = _fn("cxs_function_name_in", CSTR, P, I32)
cxs_function_name_out = _fn("cxs_function_name_out", CSTR, P, I32)
cxs_function_sparsity_in = _fn("cxs_function_sparsity_in", I32, P, I32, PP)
cxs_function_sparsity_out = _fn("cxs_function_sparsity_out", I32, P, I32, PP)
cxs_function_nz_out = _fn("cxs_function_nz_out", I32, P, I32, PP)
cxs_function_nz_in = _fn("cxs_function_nz_in", I32, P, I32, PP)
cxs_function_tape = _fn("cxs_function_tape", I32, P, P)
cxs_function_eval = _fn("cxs_function_eval", I32, P, PP, PP)
cxs_function_map_eval = _fn("cxs_function_map_eval", I32, P, I64, PP, PP, I32)
cxs_function_call = _fn("cxs_function_call", I32, P, PP, PP)
cxs_function_jacobian = _fn("cxs_function_jacobian", I32, P, PP)
cxs_function_forward = _fn("cxs_function_forward", I32, P, I64, PP)
cxs_function_reverse = _fn("cxs_function_reverse", I32, P, I64, PP)
cxs_function_jac_plan = _fn("cxs_function_jac_plan", I32, P, I32, I32, C.POINTER(I32), PP, PP, PP, PP, PP)
cxs_function_jac_sparsity = _fn("cxs_function_jac_sparsity", I32, P, I32, I32, I32, I32, PP)
cxs_jacobian = _fn("cxs_jacobian", I32, PI64, PU64, PI64, PU64, I32, PP, PP)
cxs_gradient = _fn("cxs_gradient", I32, PI64, PU64, PI64, PU64, PP, PP)
cxs_hessian = _fn("cxs_hessian", I32, PI64, PU64, PI64, PU64, PP, PP)
cxs_sp_uni_coloring = _fn("cxs_sp_uni_coloring", I32, PI64, I64, PP)
cxs_sp_star_coloring = _fn("cxs_sp_star_coloring", I32, PI64, I64, PP)
cxs_sp_largest_first = _fn("cxs_sp_largest_first", I32, PI64, PP)
cxs_sp_amd = _fn("cxs_sp_amd", I32, PI64, PP)
cxs_sp_etree = _fn("cxs_sp_etree", I32, PI64, I32, PP)
cxs_sp_ldl = _fn("cxs_sp_ldl", I32, PI64, I32, PP, PP)
cxs_sp_qr = _fn("cxs_sp_qr", I32, PI64, I32, PP, PP, PP, PP)
cxs_sp_mtimes = _fn("cxs_sp_mtimes", I32, PI64, PI64, PP)
cxs_buf_size = _fn("cxs_buf_size", I64, P)
cxs_buf_copy = _fn("cxs_buf_copy", I32, P, P)
cxs_buf_free = _fn("cxs_buf_free", I32, P)


def take_buf(h, dtype="int64"):
    """Copy a cxs_buf out as a numpy array and free it."""
    import numpy as np
    n = cxs_buf_size(h)
    out = np.empty(n, dtype=np.int64)
    if n:
        cxs_buf_copy(h, out.ctypes.data)
    cxs_buf_free(h)
    return out.view(np.uint64) if dtype == "uint64" else out


def ptr_array(ptrs):
    """C array of void* from a list of ints / None."""
    arr = (C.c_void_p * max(len(ptrs), 1))()
    for i, p in enumerate(ptrs):
        arr[i] = p if p else None
    return arr
