"""ctypes binding of libmifft.so (declarations mirror include/mifft.h)."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmifft.so")

MAX_RANK = 8
MAX_FORCED = 8

u32, u64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32
vp, cp = ctypes.c_void_p, ctypes.c_char_p


class Tuning(ctypes.Structure):
    _fields_ = [
        ("workgroup_size", u32),
        ("fused_workgroup_size", u32),
        ("rader_max_prime", u64),
        ("direct_max_prime", u64),
        ("force_rader_axes", i32 * MAX_FORCED),
        ("n_force_rader_axes", i32),
        ("force_bluestein_axes", i32 * MAX_FORCED),
        ("n_force_bluestein_axes", i32),
        ("fuse_long_axes", i32),
        ("fuse_small_volumes", i32),
        ("fused_min_convolution_length", u64),
    ]


class Config(ctypes.Structure):
    _fields_ = [
        ("shape", u64 * MAX_RANK),
        ("rank", i32),
        ("axes", i32 * MAX_RANK),
        ("naxes", i32),
        ("batch", u64),
        ("direction", i32),
        ("normalization", i32),
        ("precision", i32),
        ("tuning", Tuning),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("n", u64),
        ("batch", u64),
        ("direction", i32),
        ("normalization", i32),
        ("algorithm", i32),
        ("reserved", i32),
        ("in_ptr", vp),
        ("out_ptr", vp),
    ]


class LogicalLayout(ctypes.Structure):
    _fields_ = [("element_offset", u64), ("element_stride", u64), ("batch_stride", u64)]


class ServerRequest(ctypes.Structure):
    _fields_ = [
        ("kind", i32),
        ("precision", i32),
        ("rank", i32),
        ("direction", i32),
        ("normalization", i32),
        ("reserved", i32),
        ("shape", u64 * 2),
        ("batch", u64),
        ("in_ptr", vp),
        ("out_ptr", vp),
        ("in_bytes", u64),
        ("out_bytes", u64),
    ]


class ServerCompletion(ctypes.Structure):
    _fields_ = [
        ("ticket", u64),
        ("submit_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("status", i32),
        ("path", i32),
    ]


class ServerResult(ctypes.Structure):
    _fields_ = [
        ("arrival_s", ctypes.c_double),
        ("admit_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("status", i32),
        ("path", i32),
    ]


# (name, restype, argtypes) for every exported function of include/mifft.h.
SIGNATURES = [
    ("mifft_tuning_init", None, [ctypes.POINTER(Tuning)]),
    ("mifft_config_init", i32, [ctypes.POINTER(Config), i32, ctypes.POINTER(u64)]),
    ("mifft_config_validate", i32, [ctypes.POINTER(Config)]),
    ("mifft_config_scale", i32, [ctypes.POINTER(Config), ctypes.POINTER(ctypes.c_double)]),
    ("mifft_config_total_len", i32, [ctypes.POINTER(Config), ctypes.POINTER(u64)]),
    ("mifft_plan_create", i32, [ctypes.POINTER(Config), i32, ctypes.POINTER(vp)]),
    ("mifft_plan_destroy", None, [vp]),
    ("mifft_plan_create_with_diagnostics", i32, [ctypes.POINTER(Config), i32, ctypes.POINTER(vp), ctypes.c_char_p,
                                                 ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("mifft_last_error_diagnostics", i32, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("mifft_plan_diagnostics_for_workspace", i32, [vp, vp, u64, vp, u64, vp, u64, ctypes.c_char_p, ctypes.c_size_t,
                                                   ctypes.POINTER(ctypes.c_size_t)]),
    ("mifft_plan_diagnostics_for_views", i32, [vp, vp, u64, ctypes.POINTER(LogicalLayout), vp, u64,
                                               ctypes.POINTER(LogicalLayout), ctypes.POINTER(u64), ctypes.c_char_p,
                                               ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("mifft_plan_kind", i32, [vp]),
    ("mifft_plan_config", i32, [vp, ctypes.POINTER(Config)]),
    ("mifft_plan_factors", i32, [vp, ctypes.POINTER(u64), i32]),
    ("mifft_plan_axis_factors", i32, [vp, i32, ctypes.POINTER(u64), i32]),
    ("mifft_plan_axis_kinds", i32, [vp, ctypes.POINTER(i32), i32]),
    ("mifft_plan_route", i32, [vp]),
    ("mifft_plan_workspace_bytes", u64, [vp]),
    ("mifft_plan_required_input_bytes", u64, [vp]),
    ("mifft_plan_required_output_bytes", u64, [vp]),
    ("mifft_plan_required_bytes", u64, [vp]),
    ("mifft_plan_packed_shape", i32, [vp, ctypes.POINTER(u64), i32]),
    ("mifft_plan_supports_in_place", i32, [vp]),
    ("mifft_plan_launch_count", i32, [vp]),
    ("mifft_plan_describe", i32, [vp, ctypes.c_char_p, ctypes.c_size_t]),
    ("mifft_plan_diagnostics", i32, [vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("mifft_plan_launch_describe", i32, [vp, i32, ctypes.c_char_p, ctypes.c_size_t]),
    ("mifft_execute_launch", i32, [vp, vp, i32, vp, u64, vp, u64, vp, u64]),
    ("mifft_plan_launch_tuning", i32, [vp, i32, ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]),
    ("mifft_plan_table_count", i32, [vp]),
    ("mifft_plan_table", i32, [vp, i32, ctypes.c_char_p, ctypes.c_size_t, vp, u64, ctypes.POINTER(u64)]),
    ("mifft_twiddle_table", i32, [vp, u64, u64, i32, vp]),
    ("mifft_registry_count", i32, []),
    ("mifft_registry_name", i32, [i32, ctypes.c_char_p, ctypes.c_size_t]),
    ("mifft_execute", i32, [vp, vp, vp, u64, vp, u64]),
    ("mifft_execute_with_workspace", i32, [vp, vp, vp, u64, vp, u64, vp, u64]),
    ("mifft_execute_in_place", i32, [vp, vp, vp, u64]),
    ("mifft_execute_static_scratch", i32, [vp, vp, vp, u64, vp, u64]),
    ("mifft_logical_layout_init", None, [ctypes.POINTER(LogicalLayout)]),
    ("mifft_logical_layout_span", i32, [ctypes.POINTER(LogicalLayout), u64, u64, ctypes.POINTER(u64)]),
    ("mifft_plan_logical_workspace_bytes", i32, [vp, ctypes.POINTER(LogicalLayout), ctypes.POINTER(LogicalLayout),
                                                 ctypes.POINTER(u64)]),
    ("mifft_execute_logical", i32, [vp, vp, vp, u64, ctypes.POINTER(LogicalLayout), vp, u64,
                                    ctypes.POINTER(LogicalLayout)]),
    ("mifft_execute_logical_with_workspace", i32, [vp, vp, vp, u64, ctypes.POINTER(LogicalLayout), vp, u64,
                                                   ctypes.POINTER(LogicalLayout), vp, u64]),
    ("mifft_record_logical", i32, [vp, vp, vp, u64, ctypes.POINTER(LogicalLayout), vp, u64,
                                   ctypes.POINTER(LogicalLayout)]),
    ("mifft_recorder_begin", i32, [vp, i32, ctypes.POINTER(vp)]),
    ("mifft_recorder_stream", vp, [vp]),
    ("mifft_record", i32, [vp, vp, vp, u64, vp, u64]),
    ("mifft_record_in_place", i32, [vp, vp, vp, u64]),
    ("mifft_recorder_end", i32, [vp]),
    ("mifft_recorder_launch", i32, [vp, vp]),
    ("mifft_recorder_destroy", None, [vp]),
    ("mifft_batch_create", i32, [ctypes.POINTER(BatchItem), i32, ctypes.POINTER(Tuning), ctypes.POINTER(vp)]),
    ("mifft_batch_execute", i32, [vp, vp]),
    ("mifft_batch_capture", i32, [vp, vp]),
    ("mifft_batch_launch", i32, [vp, vp]),
    ("mifft_batch_info", i32, [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64),
                               ctypes.POINTER(u64), ctypes.POINTER(u32)]),
    ("mifft_batch_destroy", None, [vp]),
    ("mifft_server_create", i32, [u64, i32, i32, ctypes.POINTER(vp)]),
    ("mifft_server_set_batch_limit", i32, [vp, u64]),
    ("mifft_server_prepare", i32, [vp, ctypes.POINTER(ServerRequest), i32]),
    ("mifft_server_replay", i32, [vp, ctypes.POINTER(ServerRequest), ctypes.POINTER(ctypes.c_double), i32, i32,
                                  ctypes.POINTER(ServerResult)]),
    ("mifft_server_destroy", None, [vp]),
    ("mifft_server_submit", i32, [vp, ctypes.POINTER(ServerRequest), ctypes.POINTER(u64)]),
    ("mifft_server_poll", i32, [vp, i32, ctypes.POINTER(ServerCompletion), i32, ctypes.POINTER(i32)]),
    ("mifft_server_pending", i32, [vp, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64)]),
    ("mifft_server_set_plan_cache", i32, [vp, u64]),
    ("mifft_server_cached_shapes", u64, [vp]),
    ("mifft_status_string", cp, [i32]),
    ("mifft_last_error_detail", cp, []),
    ("mifft_version", cp, []),
    ("mifft_autotune", i32, [ctypes.POINTER(Config), i32, i32, vp, ctypes.POINTER(i32), ctypes.c_char_p, u64]),
    ("mifft_autotune_sequence", i32, [ctypes.POINTER(Config), ctypes.POINTER(i32), i32, i32, vp,
                                       ctypes.POINTER(i32), ctypes.c_char_p, u64]),
    ("mifft_tuning_table_save", i32, [ctypes.c_char_p]),
    ("mifft_tuning_table_load", i32, [ctypes.c_char_p]),
    ("mifft_tuning_table_parse", i32, [ctypes.c_char_p, ctypes.POINTER(i32)]),
    ("mifft_tuning_table_clear", None, []),
    ("mifft_tuning_table_size", i32, []),
    ("mifft_cpu_plan_create", i32, [ctypes.POINTER(Config), i32, ctypes.POINTER(vp)]),
    ("mifft_cpu_plan_destroy", None, [vp]),
    ("mifft_cpu_plan_kind", i32, [vp]),
    ("mifft_cpu_plan_config", i32, [vp, ctypes.POINTER(Config)]),
    ("mifft_cpu_plan_packed_shape", i32, [vp, ctypes.POINTER(u64), i32]),
    ("mifft_cpu_plan_required_input_len", u64, [vp]),
    ("mifft_cpu_plan_required_output_len", u64, [vp]),
    ("mifft_cpu_execute", i32, [vp, vp, u64, vp, u64]),
    ("mifft_cpu_execute_f64", i32, [vp, vp, u64, vp, u64]),
    ("mifft_cpu_execute_in_place", i32, [vp, vp, u64]),
    ("mifft_cpu_execute_in_place_f64", i32, [vp, vp, u64]),
    ("mifft_cpu_set_max_threads", None, [i32]),
    ("mifft_cpu_max_threads", i32, []),
    ("mifft_cpu_isa", cp, []),
    ("mifft_fill_uniform_f32", i32, [vp, u64, u64, vp]),
    ("mifft_fill_normal_f32", i32, [vp, u64, u64, vp]),
]

_lib = None


class LibraryMissing(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Loads the in-tree libmifft.so; raises LibraryMissing if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is missing: build it with `python -m wgpufft_amd.build` "
            "(there is no CPU fallback)")
    # torch ships its own libamdhip64.so.7 (and HSA runtime); libmifft.so
    # needs the same soname.  Loading torch first makes both use torch's copy:
    # loaded the other way round, torch would run on a mismatched HIP runtime
    # and see no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
