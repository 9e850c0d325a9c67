"""ctypes binding of libbmfr_hip.so (the C ABI declared in include/bmfr_hip.h).

The shared library is built in-tree by ``bmfr_amd/csrc/Makefile``.  Loading
fails loudly when it is missing: there is no CPU fallback for any stage.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbmfr_hip.so")

c_int32 = ctypes.c_int32
c_float_p = ctypes.POINTER(ctypes.c_float)


class BmfrParams(ctypes.Structure):
    """struct bmfr_params (include/bmfr_hip.h); the reference's -D options
    of opencl/bmfr.cpp:205-232 as runtime fields."""
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("position_limit_squared", ctypes.c_float),
        ("normal_limit_squared", ctypes.c_float),
        ("blend_alpha", ctypes.c_float),
        ("second_blend_alpha", ctypes.c_float),
        ("taa_blend_alpha", ctypes.c_float),
        ("noise_amount", ctypes.c_double),
        ("storage", ctypes.c_int32),
        ("tmp_storage", ctypes.c_int32),
        ("qr_mode", ctypes.c_int32),
        ("refactor_threshold", ctypes.c_float),
        ("feature_count", ctypes.c_int32),          # 0: the reference's set
        ("features_not_scaled", ctypes.c_int32),
        ("feature_factors", (ctypes.c_uint8 * 2) * 16),
    ]


class BmfrSizes(ctypes.Structure):
    _fields_ = [
        ("workset_width", ctypes.c_int32),
        ("workset_height", ctypes.c_int32),
        ("margin_width", ctypes.c_int32),
        ("margin_height", ctypes.c_int32),
        ("blocks_x", ctypes.c_int32),
        ("blocks_y", ctypes.c_int32),
        ("blocks", ctypes.c_int32),
        ("pixels", ctypes.c_int64),
        ("padded_pixels", ctypes.c_int64),
        ("device_bytes", ctypes.c_int64),
    ]


class BmfrFrameRequest(ctypes.Structure):
    """struct bmfr_frame_request: one frame of a server stream for the
    request loop (bmfr_server_submit)."""
    _fields_ = [
        ("stream", ctypes.c_void_p),
        ("frame_number", ctypes.c_int32),
        ("prev_frame_camera_matrix", ctypes.c_float * 16),
        ("pixel_offset", ctypes.c_float * 2),
        ("albedo", ctypes.c_void_p),
        ("normals", ctypes.c_void_p),
        ("positions", ctypes.c_void_p),
        ("noisy", ctypes.c_void_p),
        ("surface", ctypes.c_void_p),
        ("surface_format", ctypes.c_int32),
        ("surface_row_pitch", ctypes.c_int64),
        ("producer_stream", ctypes.c_void_p),
        ("t_arrival", ctypes.c_double),
        ("tag", ctypes.c_uint64),
    ]


class BmfrServerEvent(ctypes.Structure):
    """struct bmfr_server_event (bmfr_server_poll)."""
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("stream", ctypes.c_void_p),
        ("tag", ctypes.c_uint64),
        ("frame_number", ctypes.c_int32),
        ("step_streams", ctypes.c_int32),
        ("step_id", ctypes.c_int64),
        ("t_arrival", ctypes.c_double),
        ("t_launch", ctypes.c_double),
        ("t_done", ctypes.c_double),
        ("step_gpu_ms", ctypes.c_float),
    ]


# (name, restype, argtypes) of every exported entry point.
SIGNATURES = [
    ("bmfr_api_version", c_int32, []),
    ("bmfr_status_string", ctypes.c_char_p, [c_int32]),
    ("bmfr_last_error", ctypes.c_char_p, []),
    ("bmfr_device_count", c_int32, []),
    ("bmfr_measure_lds_read_bandwidth", c_int32, [c_int32, ctypes.POINTER(ctypes.c_double)]),
    ("bmfr_default_params", c_int32, [c_int32, c_int32, ctypes.POINTER(BmfrParams)]),
    ("bmfr_parse_feature_buffers", c_int32, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(BmfrParams)]),
    ("bmfr_compute_sizes", c_int32, [ctypes.POINTER(BmfrParams), ctypes.POINTER(BmfrSizes)]),
    ("bmfr_stream_open", c_int32, [ctypes.POINTER(BmfrParams), ctypes.POINTER(ctypes.c_void_p)]),
    ("bmfr_stream_close", c_int32, [ctypes.c_void_p]),
    ("bmfr_stream_sizes", c_int32, [ctypes.c_void_p, ctypes.POINTER(BmfrSizes)]),
    ("bmfr_stream_reset", c_int32, [ctypes.c_void_p]),
    ("bmfr_set_fitter_config", c_int32, [ctypes.c_void_p, c_int32, c_int32]),
    ("bmfr_upload_frame", c_int32, [ctypes.c_void_p, c_float_p, c_float_p, c_float_p, c_float_p]),
    ("bmfr_copy_frame_inputs_device", c_int32, [ctypes.c_void_p] * 5),
    ("bmfr_bind_device_inputs", c_int32, [ctypes.c_void_p] * 5),
    ("bmfr_bind_output_surface", c_int32, [ctypes.c_void_p, ctypes.c_void_p, c_int32, ctypes.c_int64]),
    ("bmfr_stream_hip_stream", c_int32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("bmfr_accumulate_noisy_data", c_int32, [ctypes.c_void_p, c_int32, c_float_p, c_float_p]),
    ("bmfr_fitter", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_weighted_sum", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_fitter_fp64_weights", c_int32, [ctypes.c_void_p, c_float_p]),
    ("bmfr_refactor_count", c_int32, [ctypes.c_void_p, ctypes.POINTER(c_int32)]),
    ("bmfr_fit_blocks", c_int32, [ctypes.c_void_p, c_int32, c_int32, c_int32, ctypes.c_double, c_int32,
                                  ctypes.c_float, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("bmfr_accumulate_filtered_data", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_taa", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_end_frame", c_int32, [ctypes.c_void_p]),
    ("bmfr_set_fused", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_run_frame", c_int32, [ctypes.c_void_p, c_int32, c_float_p, c_float_p]),
    ("bmfr_download_result", c_int32, [ctypes.c_void_p, c_float_p]),
    ("bmfr_synchronize", c_int32, [ctypes.c_void_p]),
    ("bmfr_buffer_bytes", c_int32, [ctypes.c_void_p, c_int32, ctypes.POINTER(ctypes.c_size_t)]),
    ("bmfr_read_buffer", c_int32, [ctypes.c_void_p, c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    ("bmfr_write_buffer", c_int32, [ctypes.c_void_p, c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    ("bmfr_buffer_device_ptr", c_int32, [ctypes.c_void_p, c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    ("bmfr_set_timing", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_stage_times", c_int32, [ctypes.c_void_p, c_float_p]),
    ("bmfr_record_marker", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_marker_elapsed", c_int32, [ctypes.c_void_p, c_int32, c_int32, c_float_p]),
    ("bmfr_batch_create", c_int32, [ctypes.POINTER(ctypes.c_void_p), c_int32,
                                    ctypes.POINTER(ctypes.c_void_p)]),
    ("bmfr_batch_destroy", c_int32, [ctypes.c_void_p]),
    ("bmfr_batch_run", c_int32, [ctypes.c_void_p, ctypes.POINTER(c_int32), c_float_p, c_float_p,
                                 c_int32]),
    ("bmfr_batch_synchronize", c_int32, [ctypes.c_void_p]),
    ("bmfr_batch_set_graph", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_batch_set_fused", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_batch_set_lanes", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_batch_set_fitter_config", c_int32, [ctypes.c_void_p, c_int32, c_int32, c_int32]),
    ("bmfr_batch_stage_times", c_int32, [ctypes.c_void_p, c_float_p]),
    ("bmfr_batch_autotune", c_int32, [ctypes.c_void_p, c_int32, ctypes.c_char_p, ctypes.c_size_t]),
    ("bmfr_batch_record_marker", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_batch_marker_elapsed", c_int32, [ctypes.c_void_p, c_int32, c_int32, c_float_p]),
    ("bmfr_server_create", c_int32, [ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]),
    ("bmfr_server_destroy", c_int32, [ctypes.c_void_p]),
    ("bmfr_server_usage", c_int32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                    ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(c_int32)]),
    ("bmfr_server_open_stream", c_int32, [ctypes.c_void_p, ctypes.POINTER(BmfrParams),
                                          ctypes.POINTER(ctypes.c_void_p)]),
    ("bmfr_server_close_stream", c_int32, [ctypes.c_void_p, ctypes.c_void_p]),
    ("bmfr_server_request_stream", c_int32, [ctypes.c_void_p, ctypes.POINTER(BmfrParams), ctypes.c_uint64,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    ("bmfr_server_cancel_request", c_int32, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(c_int32)]),
    ("bmfr_server_step", c_int32, [ctypes.c_void_p, c_int32, ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.POINTER(c_int32), c_float_p, c_float_p,
                                   ctypes.POINTER(ctypes.c_int64)]),
    ("bmfr_server_step_query", c_int32, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(c_int32),
                                         c_float_p]),
    ("bmfr_server_synchronize", c_int32, [ctypes.c_void_p]),
    ("bmfr_server_set_fused", c_int32, [ctypes.c_void_p, c_int32]),
    ("bmfr_server_set_policy", c_int32, [ctypes.c_void_p, c_int32, c_int32]),
    ("bmfr_clock_seconds", ctypes.c_double, []),
    ("bmfr_server_submit", c_int32, [ctypes.c_void_p, ctypes.POINTER(BmfrFrameRequest)]),
    ("bmfr_server_poll", c_int32, [ctypes.c_void_p, ctypes.POINTER(BmfrServerEvent), c_int32, c_int32,
                                   ctypes.POINTER(c_int32)]),
    ("bmfr_scene_create", c_int32, [ctypes.c_uint64, c_int32, ctypes.c_float, c_int32, c_int32,
                                    c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    ("bmfr_scene_destroy", c_int32, [ctypes.c_void_p]),
    ("bmfr_scene_camera", c_int32, [ctypes.c_void_p, c_int32, c_float_p, c_float_p]),
    ("bmfr_scene_render_host", c_int32, [ctypes.c_void_p, c_int32, c_float_p, c_float_p,
                                         c_float_p, c_float_p]),
    ("bmfr_scene_render_to_stream", c_int32, [ctypes.c_void_p, c_int32, ctypes.c_void_p]),
    ("bmfr_scene_render_device", c_int32, [ctypes.c_void_p, c_int32, c_int32] + [ctypes.c_void_p] * 5),
    ("bmfr_exr_info", c_int32, [ctypes.c_char_p, ctypes.POINTER(c_int32), ctypes.POINTER(c_int32),
                                ctypes.POINTER(c_int32)]),
    ("bmfr_exr_read_rgb", c_int32, [ctypes.c_char_p, c_int32, c_int32, c_float_p]),
    ("bmfr_exr_write_rgb", c_int32, [ctypes.c_char_p, c_int32, c_int32, c_float_p, ctypes.c_int64, c_int32]),
    ("bmfr_png_write_rgb", c_int32, [ctypes.c_char_p, c_int32, c_int32, c_float_p, ctypes.c_int64]),
    ("bmfr_camera_matrices_load", c_int32, [ctypes.c_char_p, c_int32, c_float_p, c_float_p, c_float_p,
                                            c_float_p]),
]

_lib = None


def load():
    """Load libbmfr_hip.so once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C bmfr_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
