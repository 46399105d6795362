"""Loader for the in-tree HIP library (qhg4_amd/libqhg4_amd.so).

There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
import ctypes as C
import os

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqhg4_amd.so")
_lib = None

# every symbol declared in include/qhg_gpu.h
EXPORTS = [
    "qhg_gpu_abi_version", "qhg_gpu_last_error", "qhg_gpu_genome_words", "qhg_gpu_binomial_table",
    "qhg_gpu_create", "qhg_gpu_destroy", "qhg_gpu_upload_agents", "qhg_gpu_num_alive",
    "qhg_gpu_download_agents", "qhg_gpu_pre_loop", "qhg_gpu_initialize_step", "qhg_gpu_do_actions",
    "qhg_gpu_finalize_step", "qhg_gpu_run", "qhg_gpu_set_env", "qhg_gpu_env_event", "qhg_gpu_set_navigation", "qhg_gpu_sync", "qhg_gpu_stats",
    "qhg_gpu_cell_counts", "qhg_gpu_capacities", "qhg_gpu_env_weights", "qhg_gpu_compact", "qhg_gpu_state_digest",
    "qhg_gpu_set_timing", "qhg_gpu_kernel_times", "qhg_gpu_kernel_name", "qhg_gpu_device_bytes",
    "qhg_gpu_set_tuning", "qhg_gpu_get_tuning", "qhg_gpu_autotune",
    "qhg_params_from_modules", "qhg_params_get_attribute", "qhg_params_set_attribute", "qhg_gpu_modify_attribute",
    "qhg_gpu_set_clock", "qhg_gpu_get_clock", "qhg_gpu_get_params", "qhg_params_from_attribute_map",
    "qhg_params_attribute_name", "qhg_params_weight_name", "qhg_params_polyline_name", "qhg_params_event_flags",
    "qhg_params_class_action", "qhg_params_evaluator_name",
    "qhg_batch_create", "qhg_batch_destroy", "qhg_batch_add", "qhg_batch_remove", "qhg_batch_size",
    "qhg_batch_run", "qhg_batch_sync", "qhg_batch_set_timing", "qhg_batch_kernel_times",
    "qhg_batch_kernel_name", "qhg_batch_set_graph", "qhg_batch_set_tuning", "qhg_batch_get_tuning",
    "qhg_batch_autotune", "qhg_env_tables_noalloc",
]


class QhgError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise QhgError("%s not built: run `python -m qhg4_amd.build` (or __graft_entry__.build())" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    h = C.c_void_p
    sig = {
        "qhg_gpu_abi_version": ([], C.c_int),
        "qhg_gpu_last_error": ([], C.c_char_p),
        "qhg_gpu_genome_words": ([C.c_int32], C.c_int),
        "qhg_gpu_binomial_table": ([C.c_double, C.c_int32, C.c_double, P(C.c_double), P(C.c_int32)], C.c_int),
        "qhg_gpu_create": ([P(abi.GridDesc), P(abi.EnvDesc), P(abi.Params), C.c_int64, P(h)], C.c_int),
        "qhg_gpu_destroy": ([h], C.c_int),
        "qhg_gpu_upload_agents": ([h, P(abi.Agents)], C.c_int),
        "qhg_gpu_num_alive": ([h, P(C.c_int64)], C.c_int),
        "qhg_gpu_download_agents": ([h, P(abi.Agents)], C.c_int),
        "qhg_gpu_pre_loop": ([h, C.c_double], C.c_int),
        "qhg_gpu_initialize_step": ([h, C.c_double], C.c_int),
        "qhg_gpu_do_actions": ([h, C.c_uint32, C.c_double], C.c_int),
        "qhg_gpu_finalize_step": ([h], C.c_int),
        "qhg_gpu_run": ([h, C.c_int32], C.c_int),
        "qhg_gpu_set_env": ([h, P(abi.EnvDesc)], C.c_int),
        "qhg_gpu_env_event": ([h, P(abi.EnvDesc), C.c_int32], C.c_int),
        "qhg_gpu_set_navigation": ([h, P(abi.NavDesc)], C.c_int),
        "qhg_params_event_flags": ([C.c_char_p, C.c_int32], C.c_int),
        "qhg_gpu_sync": ([h], C.c_int),
        "qhg_gpu_stats": ([h, P(abi.Stats)], C.c_int),
        "qhg_gpu_cell_counts": ([h, P(C.c_uint32)], C.c_int),
        "qhg_gpu_capacities": ([h, P(C.c_double)], C.c_int),
        "qhg_gpu_env_weights": ([h, P(C.c_double)], C.c_int),
        "qhg_gpu_compact": ([h], C.c_int),
        "qhg_gpu_state_digest": ([h, P(C.c_uint64)], C.c_int),
        "qhg_gpu_set_timing": ([h, C.c_int32], C.c_int),
        "qhg_gpu_kernel_times": ([h, P(C.c_double), P(C.c_int64), C.c_int32, P(C.c_int32)], C.c_int),
        "qhg_gpu_kernel_name": ([C.c_int32], C.c_char_p),
        "qhg_gpu_device_bytes": ([h, P(C.c_int64)], C.c_int),
        "qhg_gpu_set_tuning": ([h, P(abi.Tuning)], C.c_int),
        "qhg_gpu_get_tuning": ([h, P(abi.Tuning)], C.c_int),
        "qhg_gpu_autotune": ([h, C.c_int32, P(abi.Tuning), P(C.c_int32)], C.c_int),
        "qhg_params_from_modules": ([C.c_char_p, P(abi.ModuleAttrs), C.c_int32, P(C.c_char_p), P(C.c_int32),
                                     C.c_int32, C.c_uint64, P(abi.Params)], C.c_int),
        "qhg_params_get_attribute": ([P(abi.Params), C.c_char_p, P(C.c_double)], C.c_int),
        "qhg_params_set_attribute": ([P(abi.Params), C.c_char_p, C.c_double], C.c_int),
        "qhg_gpu_modify_attribute": ([h, C.c_char_p, C.c_double], C.c_int),
        "qhg_gpu_set_clock": ([h, C.c_int64, C.c_double, C.c_int64], C.c_int),
        "qhg_gpu_get_clock": ([h, P(C.c_int64), P(C.c_double), P(C.c_int64)], C.c_int),
        "qhg_gpu_get_params": ([h, P(abi.Params)], C.c_int),
        "qhg_params_from_attribute_map": ([C.c_char_p, C.c_int32, P(C.c_char_p), P(C.c_char_p), P(C.c_char_p),
                                           P(C.c_int32), C.c_int32, C.c_uint64, P(abi.Params)], C.c_int),
        "qhg_params_attribute_name": ([C.c_int32, C.c_int32], C.c_char_p),
        "qhg_params_weight_name": ([C.c_char_p, C.c_int32], C.c_char_p),
        "qhg_params_polyline_name": ([C.c_char_p, C.c_int32], C.c_char_p),
        "qhg_params_class_action": ([C.c_char_p, C.c_int32, P(C.c_int32)], C.c_char_p),
        "qhg_params_evaluator_name": ([C.c_char_p, C.c_int32], C.c_char_p),
        "qhg_batch_create": ([P(h)], C.c_int),
        "qhg_batch_destroy": ([h], C.c_int),
        "qhg_batch_add": ([h, h], C.c_int),
        "qhg_batch_remove": ([h, h], C.c_int),
        "qhg_batch_size": ([h, P(C.c_int32)], C.c_int),
        "qhg_batch_run": ([h, C.c_int32], C.c_int),
        "qhg_batch_sync": ([h], C.c_int),
        "qhg_batch_set_timing": ([h, C.c_int32], C.c_int),
        "qhg_batch_kernel_times": ([h, P(C.c_double), P(C.c_int64), C.c_int32, P(C.c_int32)], C.c_int),
        "qhg_batch_kernel_name": ([C.c_int32], C.c_char_p),
        "qhg_batch_set_graph": ([h, C.c_int32], C.c_int),
        "qhg_batch_set_tuning": ([h, P(abi.Tuning)], C.c_int),
        "qhg_batch_get_tuning": ([h, P(abi.Tuning)], C.c_int),
        "qhg_batch_autotune": ([h, C.c_int32, P(abi.Tuning), P(C.c_int32)], C.c_int),
        "qhg_env_tables_noalloc": ([P(abi.GridDesc), P(abi.EnvDesc), P(abi.Params), C.c_void_p, C.c_void_p,
                                    C.c_void_p], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc, what):
    if rc != 0:
        raise QhgError("%s failed (%d): %s" % (what, rc, lib().qhg_gpu_last_error().decode()), rc)
