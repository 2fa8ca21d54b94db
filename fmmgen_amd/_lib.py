"""ctypes binding of libfmm_hip.so (the C ABI in include/fmm_hip.h).

The library is loaded from ``fmmgen_amd/_build`` only; there is no CPU or
PyTorch fallback.  A missing library raises ``FmmLibraryMissing`` with the
build command.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfmm_hip.so")

FMM_OK = 0
MODE_PLAIN_F64 = 0
MODE_MIXED_F32 = 1
MODE_MIXED_BF16X3 = 2
MODE_LEARNED_F32 = 3
MODE_LEARNED_BF16X3 = 4
MODES = {"plain": MODE_PLAIN_F64, "f32": MODE_MIXED_F32, "bf16x3": MODE_MIXED_BF16X3,
         "learned": MODE_LEARNED_F32, "learned_bf16x3": MODE_LEARNED_BF16X3}

MLP_FEATURES, MLP_HIDDEN, MLP_OUTPUTS, MLP_ORDER0, MLP_PARAMS = 12, 64, 11, 2, 5707
MLP_NOALLOC_MAX_BATCH = 8192

PHASE_P2M, PHASE_M2M, PHASE_M2L, PHASE_P2P, PHASE_L2L, PHASE_L2P = 1, 2, 4, 8, 16, 32
PHASE_ALL = 63

# operand forms of the M2L (fmm_tree_set_m2l_form)
M2L_FORMS = {"auto": 0, "full": 1, "harmonic": 2}
OPS = {"P2M": 0, "M2M": 1, "M2L": 2, "L2L": 3, "L2P": 4, "M2P": 5, "P2P": 6, "P2M_S": 7}

# Every symbol include/fmm_hip.h declares.
EXPORTS = (
    "fmm_last_error", "fmm_abi_version", "fmm_variant_orders", "fmm_tree_build", "fmm_tree_destroy",
    "fmm_tree_info", "fmm_tree_set_sources", "fmm_evaluate", "fmm_evaluate_phases", "fmm_evaluate_bh",
    "fmm_direct_sampled",
    "fmm_op_batch", "fmm_export_cells", "fmm_export_particles", "fmm_export_lists", "fmm_export_expansions",
    "fmm_timing_enable", "fmm_timing_read",
    "fmm_mlp_forward", "fmm_mlp_infer_f16", "fmm_mlp_work_floats", "fmm_mlp_grad", "fmm_mlp_grad_noalloc", "fmm_mlp_adam",
    "fmm_tree_set_order", "fmm_tree_set_m2l_form", "fmm_tree_leaf_features", "fmm_tree_leaf_error", "fmm_tree_leaves",
    "fmm_tree_set_model", "fmm_tree_learned_stats", "fmm_tree_leaf_orders", "fmm_tree_field_scale",
)
PHASE_NAMES = ("prep", "P2M", "M2M", "M2L", "P2P", "L2L", "L2P", "order", "judge", "bh")


class FmmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"fmm error {code}: {msg}")
        self.code = code


class FmmLibraryMissing(ImportError):
    pass


_lib = None

vp = ctypes.c_void_p
sz = ctypes.c_size_t
ci = ctypes.c_int
cd = ctypes.c_double
ll_p = ctypes.POINTER(ctypes.c_longlong)
d_p = ctypes.POINTER(ctypes.c_double)
i_p = ctypes.POINTER(ctypes.c_int)


def load(path=None):
    """Loads (once) and returns the library handle."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise FmmLibraryMissing(f"{path} is missing: build it with `python -m fmmgen_amd.build`")
    L = ctypes.CDLL(path)
    L.fmm_last_error.restype = ctypes.c_char_p
    L.fmm_last_error.argtypes = []
    L.fmm_abi_version.argtypes = []
    L.fmm_variant_orders.argtypes = [ci, ci, i_p, i_p]
    L.fmm_tree_build.argtypes = [vp, vp, sz, sz, ci, cd, ci, ci, vp, ctypes.POINTER(vp)]
    L.fmm_tree_destroy.argtypes = [vp]
    L.fmm_tree_info.argtypes = [vp, ll_p]
    L.fmm_tree_set_sources.argtypes = [vp, vp]
    L.fmm_evaluate.argtypes = [vp, ci, vp, vp]
    L.fmm_evaluate_phases.argtypes = [vp, ci, ci, vp, vp]
    L.fmm_evaluate_bh.argtypes = [vp, vp, vp]
    L.fmm_direct_sampled.argtypes = [vp, vp, sz, vp, sz, ci, ci, vp, vp]
    L.fmm_op_batch.argtypes = [ci, ci, ci, ci, sz, vp, vp, ci, vp, ci, vp]
    L.fmm_export_cells.argtypes = [vp, d_p, ll_p]
    L.fmm_export_particles.argtypes = [vp, ll_p, ll_p]
    L.fmm_export_lists.argtypes = [vp, ll_p, ll_p]
    L.fmm_export_expansions.argtypes = [vp, d_p, d_p]
    L.fmm_timing_enable.argtypes = [vp, ci]
    L.fmm_timing_read.argtypes = [vp, d_p, ll_p]
    cf = ctypes.c_float
    L.fmm_mlp_forward.argtypes = [vp, ci, vp, vp, vp]
    L.fmm_mlp_infer_f16.argtypes = [vp, ci, vp, vp, vp, cd, cd, ci, ci, vp]
    L.fmm_mlp_work_floats.argtypes = [ci]
    L.fmm_mlp_grad.argtypes = [vp, vp, vp, ci, vp, vp, vp, vp, vp]
    L.fmm_mlp_grad_noalloc.argtypes = [vp, vp, vp, ci, vp, vp, vp, vp]
    L.fmm_mlp_adam.argtypes = [vp, vp, vp, vp, cf, cf, cf, cf, ci, vp]
    L.fmm_tree_set_order.argtypes = [vp, ci]
    L.fmm_tree_set_m2l_form.argtypes = [vp, ci]
    L.fmm_tree_leaf_features.argtypes = [vp, vp, vp]
    L.fmm_tree_leaf_error.argtypes = [vp, vp, vp, vp, vp]
    L.fmm_tree_leaves.argtypes = [vp, ll_p]
    L.fmm_tree_set_model.argtypes = [vp, vp, cd, cd]
    L.fmm_tree_learned_stats.argtypes = [vp, d_p]
    L.fmm_tree_leaf_orders.argtypes = [vp, i_p]
    L.fmm_tree_field_scale.argtypes = [vp, d_p, vp]
    for nm in EXPORTS:
        if nm not in ("fmm_last_error",):
            getattr(L, nm).restype = ci
    L.fmm_mlp_work_floats.restype = sz
    _lib = L
    return L


def check(code):
    if code != FMM_OK:
        msg = load().fmm_last_error()
        raise FmmError(code, msg.decode() if msg else "")
    return code
