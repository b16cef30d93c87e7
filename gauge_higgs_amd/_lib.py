"""Loads libghx.so (built in-tree by __graft_entry__.build()) and declares the C ABI.

There is no CPU fallback: a missing library raises, and every entry point runs
HIP kernels on the current device.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GHX_LIB") or os.path.join(HERE, "libghx.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    vp, i, d, u64, dp = C.c_void_p, C.c_int, C.c_double, C.c_uint64, C.POINTER(C.c_double)
    sig = {
        "ghx_create": [C.POINTER(C.c_void_p), i, i, i, C.POINTER(C.c_int), i],
        "ghx_destroy": [vp],
        "ghx_set_rep": [vp, C.c_char_p, C.POINTER(C.c_int)],
        "ghx_rep_info": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_long)],
        "ghx_rep_generators": [vp, dp],
        "ghx_rep_basis": [vp, dp],
        "ghx_set_couplings": [vp, i, d, d, d, d],
        "ghx_set_integrator": [vp, i, i, i],
        "ghx_set_seed": [vp, i, u64],
        "ghx_set_nested": [vp, i, i],
        "ghx_set_nested_inner": [vp, i],
        "ghx_init_fields": [vp, i, i, d, i, u64, d],
        "ghx_set_counters": [vp, i, u64, u64],
        "ghx_get_counters": [vp, i, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), dp],
        "ghx_refresh_momenta": [vp],
        "ghx_kick": [vp, d],
        "ghx_drift": [vp, d],
        "ghx_md_evolve": [vp],
        "ghx_hamiltonian": [vp, dp],
        "ghx_trajectory": [vp, C.POINTER(C.c_int), dp],
        "ghx_trajectories": [vp, i, i],
        "ghx_observables": [vp, dp],
        "ghx_max_unitarity": [vp, i, dp],
        "ghx_wilson_loop": [vp, C.c_char_p, i, i, dp],
        "ghx_series_analysis": [dp, C.c_long, i, dp],
        "ghx_autocorr_function": [dp, C.c_long, i, dp, C.POINTER(C.c_long)],
        "ghx_polyakov_loop_rep": [vp, C.c_char_p, i, dp],
        "ghx_reunitarize": [vp],
        "ghx_get_fields": [vp, i, dp, dp, dp, dp],
        "ghx_set_fields": [vp, i, dp, dp, dp, dp],
        "ghx_forces": [vp, i, dp, dp],
        "ghx_replica_swap": [vp, i, i, d, C.POINTER(C.c_int), dp],
        "ghx_sync": [vp],
        "ghx_time_md_steps": [vp, i, dp],
        "ghx_set_ladders": [vp, i, C.POINTER(C.c_int), C.POINTER(C.c_int), u64],
        "ghx_set_tempering": [vp, i],
        "ghx_swap_step": [vp],
        "ghx_swap_stats": [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_int)],
        "ghx_set_tuning": [vp, i, i],
        "ghx_get_tuning": [vp, i, C.POINTER(C.c_int)],
        "ghx_autotune": [vp, i, C.POINTER(C.c_int), dp],
        "ghx_set_frozen_phi": [vp, i],
        "ghx_casimir_channels": [i, C.c_char_p, dp, C.POINTER(C.c_int)],
        "ghx_set_multi_potential": [vp, dp, i, d],
        "ghx_save_checkpoint": [vp, C.c_char_p],
        "ghx_load_checkpoint": [vp, C.c_char_p],
        "ghx_normalize_phi": [vp],
        "ghx_surrogate_init": [vp, u64],
        "ghx_rng_gauss": [u64, u64, C.c_int64, dp],
        "ghx_rng_gauss_noalloc": [u64, u64, C.c_int64, dp],
        "ghx_surrogate_enable": [vp, i],
        "ghx_surrogate_nparams": [],
        "ghx_surrogate_params": [vp, i, C.POINTER(C.c_float)],
        "ghx_surrogate_gradient": [vp, C.POINTER(C.c_float), dp],
        "ghx_surrogate_train": [vp, i, d, dp, dp],
        "ghx_surrogate_snapshot": [vp, i, C.POINTER(C.c_float), C.POINTER(C.c_float), dp, dp],
        "ghx_rep_query": [i, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_long), dp, dp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = i
    L.ghx_last_error.argtypes = [vp]
    L.ghx_last_error.restype = C.c_char_p
    L.ghx_volume.argtypes = [vp]
    L.ghx_volume.restype = C.c_int64
    _lib = L
    return L
