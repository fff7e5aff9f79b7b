"""ctypes binding of libsigninum_j2k_hip.so (the C ABI in include/signinum_j2k_hip.h).

The library is built in-tree by signinum_amd._build.  Loading fails loudly
when it is missing: there is no CPU fallback for the decode path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsigninum_j2k_hip.so")

SGN_OK = 0
SGN_E_ARG = -1
SGN_E_UNSUPPORTED = -2
SGN_E_MALFORMED = -3
SGN_E_BLOCK = -4
SGN_E_BUFFER = -5
SGN_E_NOMEM = -6
SGN_E_DEVICE = -7
SGN_E_REGION = -8

MODE_PLAIN = 0
MODE_BATCHED = 1
MODE_GRAPH = 2

STAGE_BLOCKS = 0
STAGE_LEVELS = 1
STAGE_FINAL = 2

c_size_t = ctypes.c_size_t
c_u8p = ctypes.POINTER(ctypes.c_uint8)


class J2kInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "width", "height", "num_components", "bit_depth", "num_levels", "reversible", "mct",
        "high_throughput", "tiles_x", "tiles_y", "cb_width", "cb_height", "progression",
        "colorspace", "container")]


class TileJob(ctypes.Structure):
    _fields_ = [("bytes", ctypes.c_void_p), ("len", c_size_t), ("dst", ctypes.c_void_p),
                ("pitch", c_size_t), ("capacity", c_size_t)]


class TileReq(ctypes.Structure):
    _fields_ = [("bytes", ctypes.c_void_p), ("len", c_size_t), ("dst", ctypes.c_void_p),
                ("pitch", c_size_t), ("capacity", c_size_t), ("fmt", ctypes.c_int32),
                ("denominator", ctypes.c_uint32), ("has_roi", ctypes.c_uint32),
                ("roi_x", ctypes.c_uint32), ("roi_y", ctypes.c_uint32),
                ("roi_w", ctypes.c_uint32), ("roi_h", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("window_us", ctypes.c_uint32), ("hbm_budget", ctypes.c_uint64),
                ("max_batch_tiles", ctypes.c_uint32), ("max_inflight", ctypes.c_uint32)]


class ServeTile(ctypes.Structure):
    _fields_ = [("bytes", ctypes.c_void_p), ("len", c_size_t), ("fmt", ctypes.c_int32),
                ("denominator", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("first_fail", ctypes.c_uint32),
                ("tiles", ctypes.c_uint32), ("pixels", ctypes.c_uint64),
                ("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in (
        "requests_submitted", "requests_done", "requests_failed", "batches", "tiles", "pending",
        "running_batches", "hbm_in_use")]


class PlanStats(ctypes.Structure):
    _fields_ = [("tiles", ctypes.c_uint64), ("code_blocks", ctypes.c_uint64),
                ("coded_bytes", ctypes.c_uint64), ("output_bytes", ctypes.c_uint64),
                ("output_pixels", ctypes.c_uint64), ("coefficient_bytes", ctypes.c_uint64),
                ("launches_batched", ctypes.c_uint32), ("narrow_samples", ctypes.c_uint32),
                ("host_parse_ms", ctypes.c_float), ("host_merge_ms", ctypes.c_float),
                ("host_stage_ms", ctypes.c_float), ("host_copy_ms", ctypes.c_float),
                ("pruned_code_blocks", ctypes.c_uint64)]


class HtCodeBlockJob(ctypes.Structure):
    _fields_ = [("data_offset", ctypes.c_uint64), ("cleanup_length", ctypes.c_uint32),
                ("refinement_length", ctypes.c_uint32), ("width", ctypes.c_uint32),
                ("height", ctypes.c_uint32), ("output_offset", ctypes.c_uint32),
                ("output_stride", ctypes.c_uint32), ("missing_bit_planes", ctypes.c_uint8),
                ("number_of_coding_passes", ctypes.c_uint8), ("num_bitplanes", ctypes.c_uint8),
                ("stripe_causal", ctypes.c_uint8), ("dequantization_step", ctypes.c_float)]


class EncodeOptions(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_uint32) for n in (
        "levels", "cb_width_log2", "cb_height_log2", "reversible", "use_mct", "progression",
        "guard_bits")] + [("base_step", ctypes.c_float)] +
        [(n, ctypes.c_uint32) for n in ("precinct_log2", "sop", "eph", "tile_width",
                                         "tile_height", "image_x0", "image_y0", "tile_x0",
                                         "tile_y0", "packed_headers", "block_coder",
                                         "cb_style", "ht_refine", "subsampling")])


class HtCodeBlockBatchJob(ctypes.Structure):
    _fields_ = [("output_x", ctypes.c_uint32), ("output_y", ctypes.c_uint32),
                ("code_block", HtCodeBlockJob)]


# name -> (restype, argtypes)
SIGNATURES = {
    "sgn_last_error_message": (ctypes.c_char_p, []),
    "sgn_j2k_inspect": (ctypes.c_int, [ctypes.c_void_p, c_size_t, ctypes.POINTER(J2kInfo)]),
    "sgn_j2k_scaled_dims": (ctypes.c_int, [ctypes.c_void_p, c_size_t, ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_uint32),
                                           ctypes.POINTER(ctypes.c_uint32)]),
    "sgn_hip_session_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "sgn_hip_session_destroy": (None, [ctypes.c_void_p]),
    "sgn_hip_session_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "sgn_hip_session_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "sgn_hip_session_dispatches": (ctypes.c_uint64, [ctypes.c_void_p]),
    "sgn_device_alloc": (ctypes.c_int, [ctypes.c_void_p, c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "sgn_device_free": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "sgn_device_fill": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t, ctypes.c_uint8]),
    "sgn_device_copy": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, c_size_t]),
    "sgn_device_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           c_size_t]),
    "sgn_device_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         c_size_t]),
    "sgn_device_download_2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t,
                                              ctypes.c_void_p, c_size_t, c_size_t, c_size_t]),
    "sgn_host_alloc": (ctypes.c_int, [c_size_t, ctypes.POINTER(ctypes.c_void_p)]),
    "sgn_host_free": (None, [ctypes.c_void_p]),
    "sgn_hip_event_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "sgn_hip_event_destroy": (None, [ctypes.c_void_p]),
    "sgn_hip_event_record": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "sgn_hip_event_elapsed_ms": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.POINTER(ctypes.c_float)]),
    "sgn_j2k_decode_into": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t, ctypes.c_int,
                                           ctypes.c_uint32, ctypes.c_void_p, c_size_t, c_size_t]),
    "sgn_j2k_decode_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t,
                                             ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p,
                                             c_size_t, c_size_t]),
    "sgn_j2k_decode_tiles_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TileJob),
                                                   c_size_t, ctypes.c_int, ctypes.c_uint32,
                                                   ctypes.c_int, ctypes.POINTER(c_size_t)]),
    "sgn_j2k_decode_tiles_device_noalloc": (ctypes.c_int, [ctypes.c_void_p,
                                                           ctypes.POINTER(TileJob), c_size_t,
                                                           ctypes.c_int, ctypes.c_uint32,
                                                           ctypes.POINTER(c_size_t)]),
    "sgn_j2k_noalloc_scratch_bytes": (ctypes.c_uint64, []),
    "sgn_j2k_region_dims": (ctypes.c_int, [ctypes.c_void_p, c_size_t] + [ctypes.c_uint32] * 5 +
                            [ctypes.POINTER(ctypes.c_uint32)] * 4),
    "sgn_j2k_decode_region_into": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t,
                                                  ctypes.c_int] + [ctypes.c_uint32] * 5 +
                                   [ctypes.c_void_p, c_size_t, c_size_t]),
    "sgn_j2k_decode_region_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t,
                                                    ctypes.c_int] + [ctypes.c_uint32] * 5 +
                                     [ctypes.c_void_p, c_size_t, c_size_t]),
    "sgn_j2k_decode_tiles_region_device": (ctypes.c_int, [ctypes.c_void_p,
                                                          ctypes.POINTER(TileReq), c_size_t,
                                                          ctypes.POINTER(c_size_t)]),
    "sgn_dicom_extract_frames": (ctypes.c_int, [ctypes.c_void_p, c_size_t, c_size_t,
                                                ctypes.c_void_p, c_size_t,
                                                ctypes.POINTER(c_size_t),
                                                ctypes.POINTER(c_size_t), c_size_t,
                                                ctypes.POINTER(c_size_t),
                                                ctypes.POINTER(c_size_t)]),
    "sgn_j2k_plan_create_mixed": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TileReq), c_size_t,
                                                  ctypes.POINTER(ctypes.c_void_p),
                                                  ctypes.POINTER(ctypes.c_int32)]),
    "sgn_ht_lane_variant_count": (ctypes.c_int, []),
    "sgn_ht_lane_variant_name": (ctypes.c_char_p, [ctypes.c_int]),
    "sgn_j2k_plan_set_lane_variant": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "sgn_j2k_plan_get_lane_variant": (ctypes.c_int, [ctypes.c_void_p]),
    "sgn_j2k_plan_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_float), c_size_t,
                                             ctypes.POINTER(ctypes.c_int)]),
    "sgn_dwt_shape_count": (ctypes.c_int, []),
    "sgn_dwt_shape_name": (ctypes.c_char_p, [ctypes.c_int]),
    "sgn_j2k_plan_set_dwt_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "sgn_j2k_plan_get_dwt_shape": (ctypes.c_int, [ctypes.c_void_p]),
    "sgn_j2k_plan_autotune_dwt": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.POINTER(ctypes.c_float), c_size_t,
                                                 ctypes.POINTER(ctypes.c_int)]),
    "sgn_cost_model_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "sgn_cost_model_destroy": (None, [ctypes.c_void_p]),
    "sgn_cost_model_param_count": (ctypes.c_int, []),
    "sgn_cost_model_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t]),
    "sgn_cost_model_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t]),
    "sgn_cost_model_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, c_size_t, ctypes.c_void_p]),
    "sgn_cost_model_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, c_size_t, ctypes.c_uint32,
                                            ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
    "sgn_cost_model_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              c_size_t, ctypes.c_void_p]),
    "sgn_j2k_plan_time_code_blocks": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t,
                                                     c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.POINTER(c_size_t)]),
    "sgn_j2k_plan_order_code_blocks": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                                      ctypes.c_int, ctypes.c_void_p]),
    "sgn_j2k_plan_autotune_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_void_p]),
    "sgn_server_default_config": (None, [ctypes.POINTER(ServerConfig)]),
    "sgn_server_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ServerConfig),
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "sgn_server_destroy": (None, [ctypes.c_void_p]),
    "sgn_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServeTile), c_size_t,
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "sgn_server_pump": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_uint32)]),
    "sgn_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(RequestInfo), c_size_t,
                                       ctypes.POINTER(c_size_t)]),
    "sgn_server_output": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, c_size_t,
                                         ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(c_size_t),
                                         ctypes.POINTER(ctypes.c_uint32),
                                         ctypes.POINTER(ctypes.c_uint32)]),
    "sgn_server_release": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64]),
    "sgn_server_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServerStats)]),
    "sgn_j2k_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(TileJob), c_size_t,
                                           ctypes.c_int, ctypes.c_uint32,
                                           ctypes.POINTER(ctypes.c_void_p),
                                           ctypes.POINTER(ctypes.c_int32)]),
    "sgn_j2k_plan_destroy": (None, [ctypes.c_void_p]),
    "sgn_j2k_plan_set_output": (ctypes.c_int, [ctypes.c_void_p, c_size_t, ctypes.c_void_p,
                                               c_size_t, c_size_t]),
    "sgn_j2k_plan_decode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "sgn_j2k_plan_decode_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    "sgn_j2k_plan_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(PlanStats)]),
    "sgn_j2k_plan_tile_dims": (ctypes.c_int, [ctypes.c_void_p, c_size_t,
                                              ctypes.POINTER(ctypes.c_uint32),
                                              ctypes.POINTER(ctypes.c_uint32),
                                              ctypes.POINTER(c_size_t)]),
    "sgn_j2k_code_blocks": (ctypes.c_int, [ctypes.c_void_p, c_size_t,
                                           ctypes.POINTER(HtCodeBlockJob), c_size_t,
                                           ctypes.POINTER(c_size_t)]),
    "sgn_ht_decode_code_blocks": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t,
                                                 ctypes.POINTER(HtCodeBlockJob), c_size_t,
                                                 ctypes.c_void_p, c_size_t]),
    "sgn_ht_decode_code_blocks_raw": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t,
                                                     ctypes.POINTER(HtCodeBlockJob), c_size_t,
                                                     ctypes.c_void_p, c_size_t]),
    "sgn_idwt_level": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_uint32] * 4 +
                       [ctypes.c_void_p] * 5),
    "sgn_ht_decode_sub_band": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t,
                                              ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.POINTER(HtCodeBlockBatchJob), c_size_t,
                                              ctypes.c_void_p, c_size_t]),
    "sgn_store_component": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, c_size_t] +
                            [ctypes.c_uint32] * 5 + [ctypes.c_void_p, c_size_t] +
                            [ctypes.c_uint32] * 3 + [ctypes.c_float]),
    "sgn_inverse_mct": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, c_size_t,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    "sgn_j2k_encode_default_options": (None, [ctypes.POINTER(EncodeOptions)]),
    "sgn_j2k_encode_ht": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, ctypes.c_uint32, c_size_t,
                                         ctypes.POINTER(EncodeOptions), ctypes.c_void_p, c_size_t,
                                         ctypes.POINTER(c_size_t)]),
}

_lib = None


def load():
    """Returns the loaded library, raising if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with signinum_amd._build.build() "
            "(the HTJ2K decode path has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    msg = load().sgn_last_error_message()
    return msg.decode("utf-8", "replace") if msg else ""
