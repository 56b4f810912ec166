"""ctypes binding of libpolyeri (include/polyeri.h).

The library is the product: every compute call below runs HIP kernels on the
MI355X.  There is no CPU fallback -- if the shared library is missing or a
device call fails, a PolyeriError is raised.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpolyeri.so")

# every entry point include/polyeri.h declares
EXPORTS = [
    "polyeri_basis_read", "polyeri_ctx_create", "polyeri_ctx_destroy", "polyeri_jk_build",
    "polyeri_jk_build_device", "polyeri_jk_build_device_noalloc", "polyeri_sync", "polyeri_one_electron", "polyeri_eri_tensor",
    "polyeri_scf", "polyeri_mp2_energy", "polyeri_get_stats", "polyeri_set_profiling",
    "polyeri_last_error", "polyeri_device_alloc", "polyeri_device_free", "polyeri_memcpy",
    "polyeri_fock_build_device", "polyeri_device_copy", "polyeri_jk_build_screened",
    "polyeri_autotune", "polyeri_set_graph", "polyeri_set_general_contraction",
    "polyeri_mlp_create", "polyeri_mlp_destroy", "polyeri_mlp_nparam", "polyeri_mlp_params",
    "polyeri_mlp_grads", "polyeri_mlp_forward", "polyeri_mlp_train_step", "polyeri_gemm_bf16",
    "polyeri_mp2_energy_ex", "polyeri_gemm_f64", "polyeri_ctx_create_periodic", "polyeri_jk_build_cells",
    "polyeri_scf_periodic", "polyeri_mp2_energy_periodic", "polyeri_mp2_energy_qp", "polyeri_store_fourindex_eri", "polyeri_restore_fourindex_eri",
    "polyeri_restore_fourindex_eri_cells", "polyeri_eri_tensor_cells", "polyeri_reduced_cel1",
    "polyeri_load_fourindex_eri", "polyeri_set_direct", "polyeri_get_launch_config", "polyeri_set_launch_config",
    "polyeri_rys_nodes",
    "polyeri_rys_block", "polyeri_eri_gradient", "polyeri_eri_gradient_host", "polyeri_one_electron_gradient",
    "polyeri_one_electron_gradient_host", "polyeri_mp2_energy_periodic_mod", "polyeri_mp2_energy_periodic_qp",
    "polyeri_mo_integrals", "polyeri_ccsd", "polyeri_cc",
]


class PolyeriError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"polyeri error {code}: {msg}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [("schwarz_tol", ctypes.c_double), ("prim_tol", ctypes.c_double),
                ("precision", ctypes.c_int), ("mixed_cut", ctypes.c_double),
                ("device", ctypes.c_int)]


class ScfOpts(ctypes.Structure):
    _fields_ = [("scfconv", ctypes.c_double), ("scfrelax", ctypes.c_double),
                ("hfexchange", ctypes.c_double), ("lindep", ctypes.c_double),
                ("temphf", ctypes.c_double), ("scfcycles", ctypes.c_int),
                ("diis", ctypes.c_int), ("charge", ctypes.c_int),
                ("spherical", ctypes.c_int), ("guess", ctypes.c_int),
                ("incremental", ctypes.c_int), ("dens_tol", ctypes.c_double),
                ("solver", ctypes.c_int), ("diis_mode", ctypes.c_int)]


class ScfResult(ctypes.Structure):
    _fields_ = [("e_total", ctypes.c_double), ("e_nuc", ctypes.c_double),
                ("e_kin", ctypes.c_double), ("e_nuc_att", ctypes.c_double),
                ("e_coul", ctypes.c_double), ("e_hfx", ctypes.c_double),
                ("fermi", ctypes.c_double), ("iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("n_indep", ctypes.c_int),
                ("iocc", ctypes.c_int), ("jk_seconds", ctypes.c_double),
                ("fock_seconds", ctypes.c_double), ("scf_seconds", ctypes.c_double),
                ("quartets", ctypes.c_longlong)]


class CcResult(ctypes.Structure):
    _fields_ = [("e2", ctypes.c_double), ("e3", ctypes.c_double),
                ("e_singles", ctypes.c_double), ("e_doubles", ctypes.c_double),
                ("e_corr", ctypes.c_double), ("residual", ctypes.c_double),
                ("iterations", ctypes.c_int), ("converged", ctypes.c_int),
                ("tensor_ms", ctypes.c_double), ("transform_ms", ctypes.c_double),
                ("iter_ms", ctypes.c_double), ("seconds", ctypes.c_double),
                ("iter_gflop", ctypes.c_double)]


MAX_LISTS = 64   # POLYERI_MAX_LISTS


class Stats(ctypes.Structure):
    _fields_ = [("nshell", ctypes.c_int), ("nbf", ctypes.c_int), ("npair", ctypes.c_int),
                ("nprim_pair", ctypes.c_int), ("quartets", ctypes.c_longlong),
                ("prim_quartets", ctypes.c_double), ("model_flops", ctypes.c_double),
                ("last_jk_ms", ctypes.c_double), ("last_quartets", ctypes.c_longlong),
                ("last_class_ms", ctypes.c_double * MAX_LISTS),
                ("class_prim_quartets", ctypes.c_double * MAX_LISTS),
                ("class_flops", ctypes.c_double * MAX_LISTS),
                ("class_quartets", ctypes.c_longlong * MAX_LISTS),
                ("class_l", ctypes.c_int * MAX_LISTS), ("class_prec", ctypes.c_int * MAX_LISTS),
                ("class_kind", ctypes.c_int * MAX_LISTS), ("class_block", ctypes.c_int * MAX_LISTS),
                ("class_grid", ctypes.c_int * MAX_LISTS), ("nclass", ctypes.c_int),
                ("last_mp2_tensor_ms", ctypes.c_double), ("last_mp2_transform_ms", ctypes.c_double),
                ("last_mp2_flops", ctypes.c_double)]


_LIB = None


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise PolyeriError(-1, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    p, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    L.polyeri_basis_read.argtypes = [ctypes.c_char_p, i, p, p] + [p] * 17
    L.polyeri_rys_nodes.argtypes = [i, ctypes.c_double, p, p]
    L.polyeri_rys_block.argtypes = [i, i, i, i, p, p, p]
    L.polyeri_ctx_create.argtypes = [i, i, p, p, p, p, p, p, p, p, i, p, p, p, ctypes.POINTER(p)]
    L.polyeri_ctx_destroy.argtypes = [p]
    L.polyeri_ctx_create_periodic.argtypes = [i, i, p, p, p, p, p, p, p, p, i, p, p, d, i, i, p, ctypes.POINTER(p)]
    L.polyeri_jk_build_cells.argtypes = [p, p, p, p]
    L.polyeri_scf_periodic.argtypes = [p, p, i, p, p, p, p, p, p]
    L.polyeri_mp2_energy_periodic.argtypes = [p, i, p, p, p, i, i, i, i, p, p, p]
    L.polyeri_mp2_energy_periodic_mod.argtypes = [p, i, i, p, p, p, i, i, i, i, p, p, p]
    L.polyeri_mp2_energy_periodic_qp.argtypes = [p, i, i, p, p, p, i, i, i, i, i, p, p, p, p]
    L.polyeri_mp2_energy_qp.argtypes = [p, p, p, i, i, i, i, i, i, p, p, p]
    L.polyeri_mo_integrals.argtypes = [p, p, i, i, p]
    L.polyeri_ccsd.argtypes = [p, p, p, i, i, i, i, d, i, i, ctypes.POINTER(CcResult)]
    L.polyeri_cc.argtypes = [p, ctypes.c_char_p, p, p, i, i, i, i, d, i, i, ctypes.POINTER(CcResult)]
    L.polyeri_store_fourindex_eri.argtypes = [p, ctypes.c_char_p, d, ctypes.POINTER(ctypes.c_longlong)]
    L.polyeri_restore_fourindex_eri.argtypes = [ctypes.c_char_p, i, p, ctypes.POINTER(ctypes.c_longlong)]
    L.polyeri_restore_fourindex_eri_cells.argtypes = [ctypes.c_char_p, i, i, i, p, ctypes.POINTER(ctypes.c_longlong)]
    L.polyeri_load_fourindex_eri.argtypes = [p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_longlong)]
    L.polyeri_set_direct.argtypes = [p, i]
    L.polyeri_eri_gradient.argtypes = [p, p, d, p]
    L.polyeri_one_electron_gradient.argtypes = [p, p, p, p]
    L.polyeri_one_electron_gradient_host.argtypes = [i, i, p, p, p, p, p, p, p, p, i, p, p, p, p, p]
    L.polyeri_eri_gradient_host.argtypes = [i, i, p, p, p, p, p, p, p, p, i, p, p, d, p, d, p]
    L.polyeri_get_launch_config.argtypes = [p, i, p]
    L.polyeri_set_launch_config.argtypes = [p, i, p]
    L.polyeri_eri_tensor_cells.argtypes = [p, p]
    L.polyeri_reduced_cel1.argtypes = [p]
    L.polyeri_ctx_destroy.restype = None
    L.polyeri_jk_build.argtypes = [p, p, p, p]
    L.polyeri_jk_build_screened.argtypes = [p, p, ctypes.c_double, p, p]
    L.polyeri_autotune.argtypes = [p, ctypes.c_int]
    L.polyeri_set_graph.argtypes = [p, ctypes.c_int]
    L.polyeri_set_general_contraction.argtypes = [p, ctypes.c_int]
    L.polyeri_jk_build_device.argtypes = [p, p, p, p]
    L.polyeri_jk_build_device_noalloc.argtypes = [p, p, p, p]
    L.polyeri_sync.argtypes = [p]
    L.polyeri_one_electron.argtypes = [p, p, p, p, p]
    L.polyeri_eri_tensor.argtypes = [p, p]
    L.polyeri_scf.argtypes = [p, p, p, p, p, p, p]
    L.polyeri_mp2_energy.argtypes = [p, p, p, i, i, i, i, p, p]
    L.polyeri_mp2_energy_ex.argtypes = [p, p, p, i, i, i, i, i, p, p]
    L.polyeri_get_stats.argtypes = [p, p]
    L.polyeri_set_profiling.argtypes = [p, i]
    L.polyeri_last_error.restype = ctypes.c_char_p
    L.polyeri_device_alloc.argtypes = [p, ctypes.c_size_t, ctypes.POINTER(p)]
    L.polyeri_device_free.argtypes = [p, p]
    L.polyeri_memcpy.argtypes = [p, p, p, ctypes.c_size_t, i]
    L.polyeri_fock_build_device.argtypes = [p, p, d, p, p, p]
    L.polyeri_device_copy.argtypes = [p, p, p, ctypes.c_size_t]
    L.polyeri_mlp_create.argtypes = [i, p, ctypes.c_ulonglong, ctypes.POINTER(p)]
    L.polyeri_mlp_destroy.argtypes = [p]
    L.polyeri_mlp_destroy.restype = None
    L.polyeri_mlp_nparam.argtypes = [p]
    L.polyeri_mlp_params.argtypes = [p, p, i]
    L.polyeri_mlp_grads.argtypes = [p, p]
    L.polyeri_mlp_forward.argtypes = [p, p, i, p]
    L.polyeri_mlp_train_step.argtypes = [p, p, p, p, i, ctypes.c_float, i, ctypes.POINTER(ctypes.c_float)]
    lg = ctypes.c_long
    L.polyeri_gemm_bf16.argtypes = [i, i, i, p, lg, lg, p, lg, lg, ctypes.c_float, ctypes.c_float, p, lg]
    L.polyeri_gemm_f64.argtypes = [i, i, i, i, p, lg, lg, p, lg, lg, p, lg, lg]
    for name in EXPORTS:
        if name not in ("polyeri_ctx_destroy", "polyeri_last_error", "polyeri_mlp_destroy"):
            getattr(L, name).restype = ctypes.c_int
    L.polyeri_mlp_nparam.restype = ctypes.c_long
    _LIB = L
    return L


def check(rc):
    if rc != 0:
        msg = lib().polyeri_last_error()
        raise PolyeriError(rc, msg.decode() if msg else "")
    return rc
