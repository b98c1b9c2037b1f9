"""ctypes binding of the C ABI in include/bh.h (libmtl_bh.so, built in-tree).

There is no CPU fallback: if the library is missing or the GPU is absent, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmtl_bh.so")

BH_OK = 0
BH_ERR_ARG = 1
BH_ERR_HIP = 2
BH_ERR_OOM = 3
BH_ERR_CAPACITY = 4
BH_ERR_STATE = 5
STATUS_NAMES = {0: "BH_OK", 1: "BH_ERR_ARG", 2: "BH_ERR_HIP", 3: "BH_ERR_OOM",
                4: "BH_ERR_CAPACITY", 5: "BH_ERR_STATE"}

KERNEL_GRAVITY = 0
KERNEL_TSNE = 1
TREE_QUADTREE = 0
TREE_MORTON_FIXED = 1
TREE_GPU_REF = 2
NUM_PHASES = 7
PHASE_NAMES = ("bounds+morton", "sort", "gather", "tree", "upward", "traversal", "integrate")


class BHConfig(C.Structure):
    _fields_ = [("n_max", C.c_uint32), ("kernel", C.c_int32), ("tree_mode", C.c_int32),
                ("max_children", C.c_int32), ("fixed_depth", C.c_int32),
                ("code_bits", C.c_int32), ("G", C.c_float), ("softening", C.c_float),
                ("theta", C.c_double)]


class TsneParams(C.Structure):
    _fields_ = [("learning_rate", C.c_double), ("momentum_early", C.c_double),
                ("momentum_late", C.c_double), ("exag_early", C.c_double),
                ("exag_late", C.c_double), ("theta_min", C.c_double), ("theta_max", C.c_double),
                ("momentum_switch", C.c_int32), ("exag_switch", C.c_int32),
                ("schedule_iters", C.c_int32), ("pad", C.c_int32),
                ("repulsion_scale", C.c_double)]


class BHError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


_vp = C.c_void_p
_u32, _i32, _u64, _d, _f = C.c_uint32, C.c_int32, C.c_uint64, C.c_double, C.c_float

# name -> (restype, argtypes); the exported symbol set of include/bh.h
SIGNATURES = {
    "bh_config_default": (_i32, [C.POINTER(BHConfig)]),
    "bh_create": (_i32, [C.POINTER(_vp), C.POINTER(BHConfig)]),
    "bh_destroy": (None, [_vp]),
    "bh_last_error": (C.c_char_p, [_vp]),
    "bh_set_particles": (_i32, [_vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    "bh_set_particles_device": (_i32, [_vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    "bh_get_particles": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "bh_build": (_i32, [_vp, _vp]),
    "bh_forces": (_i32, [_vp, _d, _vp, _vp, _vp]),
    "bh_forces_device": (_i32, [_vp, _d, _vp, _vp, _vp]),
    "bh_leapfrog": (_i32, [_vp, C.c_int, _f, _d]),
    "bh_direct_fp64": (_i32, [_vp, _vp, _u32, _vp, _vp, _vp]),
    "bh_tsne_params_default": (_i32, [C.POINTER(TsneParams), _u32, _d]),
    "bh_tsne_params_gpu_reference": (_i32, [C.POINTER(TsneParams)]),
    "bh_tsne_theta": (_d, [C.POINTER(TsneParams), _i32]),
    "bh_tsne_set_p": (_i32, [_vp, _u32, _vp, _vp, _vp, _u32]),
    "bh_tsne_set_embedding": (_i32, [_vp, _vp, _vp]),
    "bh_tsne_get_embedding": (_i32, [_vp, _vp, _vp]),
    "bh_tsne_step": (_i32, [_vp, C.POINTER(TsneParams), _i32, _vp, _vp, _vp]),
    "bh_tsne_run": (_i32, [_vp, C.POINTER(TsneParams), _i32, _i32]),
    "bh_tsne_kl": (_i32, [_vp, _vp, _vp]),
    "bh_tsne_affinities": (_i32, [_vp, _vp, _u32, _u32, _d, _u32, _vp, _vp, _vp, _vp, _vp, _vp,
                                  _vp]),
    "bh_mlp_param_count": (_u32, []),
    "bh_mlp_create": (_i32, [_vp, _u32, _u64]),
    "bh_mlp_destroy": (None, [_vp]),
    "bh_mlp_last_error": (C.c_char_p, [_vp]),
    "bh_mlp_set_params": (_i32, [_vp, _vp, _u32]),
    "bh_mlp_get_params": (_i32, [_vp, _vp, _u32]),
    "bh_mlp_set_terms": (_i32, [_vp, _i32]),
    "bh_mlp_get_terms": (_i32, [_vp]),
    "bh_mlp_forward": (_i32, [_vp, _vp, _u32, _vp, _i32]),
    "bh_mlp_loss_grad": (_i32, [_vp, _vp, _vp, _u32, _vp, _vp]),
    "bh_mlp_adam_step": (_i32, [_vp, _vp, _f]),
    "bh_mlp_dataset_clear": (_i32, [_vp]),
    "bh_mlp_dataset_size": (_i32, [_vp, _vp]),
    "bh_mlp_dataset_get": (_i32, [_vp, _vp, _vp, _u32]),
    "bh_mlp_dataset_set": (_i32, [_vp, _vp, _vp, _u32]),
    "bh_mlp_train": (_i32, [_vp, _i32, _u32, _f, _u64, _vp]),
    "bh_learned_enable": (_i32, [_vp, _i32]),
    "bh_learned_collect": (_i32, [_vp, _vp, _d, _d, _vp, _u32, _u32, _vp]),
    "bh_forces_learned": (_i32, [_vp, _vp, _d, _d, _vp, _vp, _vp]),
    "bh_learned_export": (_i32, [_vp, _vp, _vp, _vp, _u32]),
    "bh_create_batched": (_i32, [_vp, _vp, _vp, _u32]),
    "bh_batch_info": (_i32, [_vp, _vp, _vp]),
    "bh_leapfrog_graph": (_i32, [_vp, _i32, _f, _d]),
    "bh_get_bounds": (_i32, [_vp, _vp, _vp]),
    "bh_get_quantizer": (_i32, [_vp, _vp, _vp, _vp]),
    "bh_get_morton": (_i32, [_vp, _vp, _vp]),
    "bh_tree_info": (_i32, [_vp, _vp, _vp, _vp]),
    "bh_get_tree": (_i32, [_vp] + [_vp] * 8),
    "bh_get_tree_slots": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "bh_counters": (_i32, [_vp, _vp, _vp]),
    "bh_profile_enable": (_i32, [_vp, C.c_int]),
    "bh_profile_phases": (_i32, [_vp, C.c_uint32]),
    "bh_profile_read": (_i32, [_vp, _vp, _vp, C.c_int]),
    "bh_synchronize": (_i32, [_vp]),
    "bh_traverse_variant_count": (_i32, []),
    "bh_set_traverse_variant": (_i32, [_vp, C.c_int]),
    "bh_set_traverse_split": (_i32, [_vp, C.c_int]),
    "bh_set_walk_priority": (_i32, [_vp, C.c_int]),
    "bh_batch_order_by_cost": (_i32, [_vp]),
    "bh_get_walk_priority": (_i32, [_vp, C.POINTER(C.c_int)]),
    "bh_set_solver": (_i32, [_vp, C.c_int]),
    "bh_set_sort_spin_limit": (_i32, [_vp, _u32]),
    "bh_direct_fp64_noalloc_async": (_i32, [_vp, _vp, _u32, _vp, _vp, _vp]),
    "bh_energy": (_i32, [_vp, _d, _vp, _vp, _vp, _vp]),
}

_lib = None


def load():
    """Load libmtl_bh.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        # PyTorch ships its own HIP runtime (soname libamdhip64.so.7) and asks for it by the
        # unversioned file name. Loaded after this library, it would come in as a second runtime
        # next to /opt/rocm's, and the runtime initialised second sees no devices. Importing
        # torch first makes this library bind to torch's copy: one runtime per process.
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


def check(status, ctx=None):
    if status != BH_OK:
        msg = ""
        if ctx:
            m = load().bh_last_error(ctx)
            msg = m.decode() if m else ""
        raise BHError(status, msg)
