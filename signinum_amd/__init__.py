"""signinum-j2k-hip: MI355X-native HTJ2K tile decode.

A drop-in for the HTJ2K decode surface of frames-sg/signinum
(crates/signinum-j2k/src/view.rs ``J2kDecoder``, batch.rs
``decode_tiles_into`` and the device adapter shape of
crates/signinum-j2k-cuda).  Host code parses codestreams into flat
code-block tables; hand-written gfx950 kernels do HT block decoding,
dequantisation, inverse DWT, inverse colour transform and packing.

Everything here is a thin layer over the C ABI of libsigninum_j2k_hip.so.
"""
import ctypes
import enum

import numpy as np

from . import _lib
from ._lib import (MODE_BATCHED, MODE_GRAPH, MODE_PLAIN, EncodeOptions, HtCodeBlockJob, J2kInfo,
                   PlanStats, TileJob)

__all__ = [
    "PixelFormat", "Downscale", "BackendRequest", "J2kError", "UnsupportedError",
    "MalformedError", "CodeBlockDecodeError", "BufferError", "DeviceError", "TileBatchError",
    "UnsupportedBackendError", "HipSession", "HipEvent", "J2kDecoder", "DeviceSurface", "TilePlan",
    "decode_tiles_into", "decode_tiles_to_device", "encode_j2k_lossless", "encode_j2k_ht",
    "encode_j2k_classic",
    "HtCodeBlockDecoder", "code_block_jobs", "CostModel", "MODE_PLAIN", "MODE_BATCHED",
    "MODE_GRAPH", "extract_dicom_encapsulated_frames", "InvalidRegionError", "region_dims",
    "decode_tiles_region_scaled_into",
]


class PixelFormat(enum.IntEnum):
    """PixelFormat (crates/signinum-core/src/pixel.rs:13-58)."""
    GRAY8 = 0
    RGB8 = 1
    RGBA8 = 2
    GRAY16 = 3
    RGB16 = 4

    @property
    def channels(self):
        return {0: 1, 1: 3, 2: 4, 3: 1, 4: 3}[int(self)]

    @property
    def bytes_per_sample(self):
        return 2 if self in (PixelFormat.GRAY16, PixelFormat.RGB16) else 1

    @property
    def bytes_per_pixel(self):
        return self.channels * self.bytes_per_sample


class Downscale(enum.IntEnum):
    """Downscale (crates/signinum-core/src/scale.rs:3-30); value = denominator."""
    NONE = 1
    HALF = 2
    QUARTER = 4
    EIGHTH = 8

    @property
    def denominator(self):
        return int(self)


class BackendRequest(enum.Enum):
    """BackendRequest (crates/signinum-core/src/backend.rs:15-22) with a HIP variant."""
    AUTO = "auto"
    CPU = "cpu"
    HIP = "hip"


class J2kError(Exception):
    """J2kError (crates/signinum-j2k/src/error.rs:5-74)."""

    def __init__(self, status, message):
        super().__init__(f"{message} (status {status})")
        self.status = status
        self.message = message


class UnsupportedError(J2kError):
    pass


class MalformedError(J2kError):
    pass


class CodeBlockDecodeError(J2kError):
    pass


class BufferError(J2kError):
    pass


class DeviceError(J2kError):
    pass


class UnsupportedBackendError(J2kError):
    """Error::UnsupportedBackend (crates/signinum-j2k-cuda/src/error.rs:6-41)."""


class InvalidRegionError(J2kError):
    """J2kError::InvalidRegion (crates/signinum-j2k/src/error.rs:24-31)."""


class TileBatchError(Exception):
    """TileBatchError (crates/signinum-j2k/src/batch.rs:47-67)."""

    def __init__(self, index, source):
        super().__init__(f"tile {index} decode failed: {source}")
        self.index = index
        self.source = source


_ERRORS = {
    _lib.SGN_E_UNSUPPORTED: UnsupportedError,
    _lib.SGN_E_MALFORMED: MalformedError,
    _lib.SGN_E_BLOCK: CodeBlockDecodeError,
    _lib.SGN_E_BUFFER: BufferError,
    _lib.SGN_E_DEVICE: DeviceError,
    _lib.SGN_E_REGION: InvalidRegionError,
    _lib.SGN_E_NOMEM: J2kError,
    _lib.SGN_E_ARG: J2kError,
}


def _check(rc):
    if rc != _lib.SGN_OK:
        raise _ERRORS.get(rc, J2kError)(rc, _lib.last_error())


def _buf(data):
    """(pointer, length, keepalive) of a bytes-like object."""
    if isinstance(data, (bytes, bytearray)):
        arr = np.frombuffer(data, dtype=np.uint8)
    else:
        arr = np.ascontiguousarray(np.asarray(data).view(np.uint8).reshape(-1))
    return arr.ctypes.data, arr.size, arr


def _route(backend):
    """Route rules of the device adapters (crates/signinum-j2k-cuda/src/runtime.rs:10-59)."""
    if backend in (None, BackendRequest.AUTO, BackendRequest.HIP):
        return
    raise UnsupportedBackendError(_lib.SGN_E_UNSUPPORTED,
                                  f"backend request {backend} is not supported by signinum-j2k-hip"
                                  " (it decodes on the device)")


class HipSession:
    """One HIP stream and its device tables (CudaSession,
    crates/signinum-j2k-cuda/src/session.rs:11-49).  Use from one thread at
    a time."""

    def __init__(self, device=0):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _check(self._lib.sgn_hip_session_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    @property
    def stream(self):
        return self._lib.sgn_hip_session_stream(self.handle)

    @property
    def dispatches(self):
        return int(self._lib.sgn_hip_session_dispatches(self.handle))

    def synchronize(self):
        _check(self._lib.sgn_hip_session_synchronize(self.handle))

    def close(self):
        if self.handle:
            self._lib.sgn_hip_session_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceSurface:
    """DeviceSurface of the device adapters (crates/signinum-j2k-cuda/src/
    surface.rs:11-127): a pitched device image in memory owned by the
    library (sgn_device_alloc), released with the surface."""

    def __init__(self, session, width, height, pixel_format, pitch_bytes=None, _block=None,
                 _offset=0):
        self.session = session
        self.width = int(width)
        self.height = int(height)
        self.pixel_format = PixelFormat(pixel_format)
        self.pitch_bytes = int(pitch_bytes or self.width * self.pixel_format.bytes_per_pixel)
        # a surface of a batch lives in the batch's shared allocation
        self._block = _block
        if _block is not None:
            self._ptr = _block.ptr + int(_offset)
            return
        p = ctypes.c_void_p()
        _check(_lib.load().sgn_device_alloc(session.handle, max(1, self.byte_len), ctypes.byref(p)))
        self._ptr = p.value

    @property
    def device_ptr(self):
        return self._ptr

    @property
    def byte_len(self):
        return self.pitch_bytes * self.height

    @property
    def row_bytes(self):
        return self.width * self.pixel_format.bytes_per_pixel

    def fill(self, value=0):
        _check(_lib.load().sgn_device_fill(self.session.handle, self._ptr, self.byte_len, value))

    def copy_from(self, other):
        """Device-to-device copy of another surface of the same byte length
        (asynchronous on the session stream)."""
        if other.byte_len != self.byte_len:
            raise BufferError(_lib.SGN_E_BUFFER, "surfaces differ in size")
        _check(_lib.load().sgn_device_copy(self.session.handle, self._ptr, other._ptr,
                                           self.byte_len))

    def download_into(self, out, stride):
        dst = np.asarray(out)
        if not dst.flags["C_CONTIGUOUS"] or not dst.flags["WRITEABLE"]:
            raise BufferError(_lib.SGN_E_BUFFER, "output must be a writable contiguous array")
        if self.height and dst.nbytes < stride * (self.height - 1) + self.row_bytes:
            raise BufferError(_lib.SGN_E_BUFFER, "output buffer too small")
        _check(_lib.load().sgn_device_download_2d(self.session.handle, dst.ctypes.data, stride,
                                                  self._ptr, self.pitch_bytes, self.row_bytes,
                                                  self.height))

    def to_numpy(self):
        """Tightly packed (H, W, C) array (uint8 or little-endian uint16)."""
        out = np.zeros(max(1, self.row_bytes * self.height), np.uint8)
        self.download_into(out, self.row_bytes)
        out = out[:self.row_bytes * self.height]
        if self.pixel_format.bytes_per_sample == 2:
            out = out.view("<u2")
        return out.reshape(self.height, self.width, self.pixel_format.channels)

    def close(self):
        if getattr(self, "_block", None) is not None:
            self._block = None  # the shared allocation goes with its last surface
            self._ptr = None
        elif getattr(self, "_ptr", None):
            _lib.load().sgn_device_free(self.session.handle, self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _alloc_surface(session, width, height, fmt):
    return DeviceSurface(session, width, height, fmt)


class _DeviceBlock:
    """One device allocation shared by the surfaces of a batch; freed when
    the last of them is released."""

    def __init__(self, session, nbytes):
        self.session = session
        p = ctypes.c_void_p()
        _check(_lib.load().sgn_device_alloc(session.handle, max(1, int(nbytes)), ctypes.byref(p)))
        self.ptr = p.value

    def __del__(self):
        try:
            _lib.load().sgn_device_free(self.session.handle, self.ptr)
        except Exception:
            pass


def _alloc_surfaces(session, dims, fmts):
    """Surfaces of (width, height) and format each, carved from one device
    allocation at 256-byte aligned offsets (one allocation instead of one
    per tile)."""
    sizes = [max(1, w * PixelFormat(f).bytes_per_pixel * h) for (w, h), f in zip(dims, fmts)]
    offs, total = [], 0
    for n in sizes:
        offs.append(total)
        total += (n + 255) & ~255
    block = _DeviceBlock(session, total)
    return [DeviceSurface(session, w, h, f, _block=block, _offset=o)
            for (w, h), f, o in zip(dims, fmts, offs)]


class HipEvent:
    """A HIP event recorded on a session's stream."""

    def __init__(self):
        e = ctypes.c_void_p()
        _check(_lib.load().sgn_hip_event_create(ctypes.byref(e)))
        self.handle = e.value

    def record(self, session):
        _check(_lib.load().sgn_hip_event_record(session.handle, self.handle))

    def elapsed_ms(self, end):
        ms = ctypes.c_float()
        _check(_lib.load().sgn_hip_event_elapsed_ms(self.handle, end.handle, ctypes.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            _lib.load().sgn_hip_event_destroy(self.handle)
        except Exception:
            pass


def scaled_dims(data, scale=Downscale.NONE):
    ptr, n, _keep = _buf(data)
    w = ctypes.c_uint32()
    h = ctypes.c_uint32()
    _check(_lib.load().sgn_j2k_scaled_dims(ptr, n, int(scale), ctypes.byref(w), ctypes.byref(h)))
    return w.value, h.value


def extract_dicom_encapsulated_frames(data, limit=None):
    """extract_dicom_encapsulated_frames(_with_limit)
    (crates/signinum-j2k-metal/src/dicom.rs:22-89): the codestream bytes of
    each frame of DICOM encapsulated Pixel Data, at most `limit` frames."""
    L = _lib.load()
    ptr, n, _keep = _buf(data)
    lim = (1 << 63) if limit is None else int(limit)
    count = ctypes.c_size_t()
    total = ctypes.c_size_t()
    _check(L.sgn_dicom_extract_frames(ptr, n, lim, None, 0, None, None, 0, ctypes.byref(count),
                                      ctypes.byref(total)))
    k = count.value
    out = np.zeros(max(1, total.value), np.uint8)
    offs = (ctypes.c_size_t * max(1, k))()
    lens = (ctypes.c_size_t * max(1, k))()
    _check(L.sgn_dicom_extract_frames(ptr, n, lim, out.ctypes.data, out.size, offs, lens, k,
                                      ctypes.byref(count), ctypes.byref(total)))
    return [out[offs[i]:offs[i] + lens[i]].tobytes() for i in range(k)]


def region_dims(data, roi, scale=Downscale.NONE):
    """(x, y, w, h) of the output of a region decode: Rect::scaled_covering
    (crates/signinum-core/src/types.rs:69-83) of roi = (x, y, w, h)."""
    ptr, n, _keep = _buf(data)
    o = [ctypes.c_uint32() for _ in range(4)]
    x, y, w, h = (int(v) for v in roi)
    _check(_lib.load().sgn_j2k_region_dims(ptr, n, int(Downscale(scale)), x, y, w, h,
                                           *[ctypes.byref(v) for v in o]))
    return tuple(v.value for v in o)


class J2kDecoder:
    """J2kDecoder (crates/signinum-j2k/src/view.rs:84-311; device adapter
    crates/signinum-j2k-cuda/src/decoder.rs:17-280)."""

    def __init__(self, data, session=None):
        self._data = bytes(data)
        self._info = J2kDecoder.inspect(self._data)
        self._session = session
        self._own_session = None

    @staticmethod
    def inspect(data):
        ptr, n, _keep = _buf(data)
        info = J2kInfo()
        _check(_lib.load().sgn_j2k_inspect(ptr, n, ctypes.byref(info)))
        return {f: getattr(info, f) for f, _ in J2kInfo._fields_}

    @property
    def info(self):
        return dict(self._info)

    @property
    def dimensions(self):
        return self._info["width"], self._info["height"]

    def _sess(self):
        if self._session is None:
            if self._own_session is None:
                self._own_session = HipSession()
            return self._own_session
        return self._session

    def decode_into(self, out, stride, fmt):
        return self.decode_scaled_into(None, out, stride, fmt, Downscale.NONE)

    def decode_into_with_scratch(self, pool, out, stride, fmt):
        return self.decode_scaled_into(pool, out, stride, fmt, Downscale.NONE)

    def decode_scaled_into(self, pool, out, stride, fmt, scale):
        """Decodes on the device and copies to host `out` (row stride in bytes)."""
        fmt = PixelFormat(fmt)
        scale = Downscale(scale)
        dst = np.asarray(out)
        if not dst.flags["C_CONTIGUOUS"] or not dst.flags["WRITEABLE"]:
            raise BufferError(_lib.SGN_E_BUFFER, "output must be a writable contiguous array")
        s = self._sess()
        ptr, n, _keep = _buf(self._data)
        _check(_lib.load().sgn_j2k_decode_into(s.handle, ptr, n, int(fmt), int(scale),
                                               dst.ctypes.data, dst.nbytes, int(stride)))
        w, h = scaled_dims(self._data, scale)
        return {"decoded": (0, 0, w, h)}

    def submit_to_device(self, session, fmt, backend=BackendRequest.HIP, scale=Downscale.NONE):
        """ImageDecodeSubmit::submit_to_device / submit_scaled_to_device."""
        _route(backend)
        fmt = PixelFormat(fmt)
        w, h = scaled_dims(self._data, scale)
        surf = _alloc_surface(session, w, h, fmt)
        ptr, n, _keep = _buf(self._data)
        _check(_lib.load().sgn_j2k_decode_device(session.handle, ptr, n, int(fmt), int(scale),
                                                 surf.device_ptr, surf.pitch_bytes,
                                                 surf.byte_len))
        return surf

    def decode_region_into(self, pool, out, stride, fmt, roi):
        """decode_region_into (crates/signinum-j2k/src/view.rs:224-256)."""
        return self.decode_region_scaled_into(pool, out, stride, fmt, roi, Downscale.NONE)

    def decode_region_scaled_into(self, pool, out, stride, fmt, roi, scale):
        """decode_region_scaled_into (crates/signinum-j2k/src/view.rs:289-311):
        the region's covering rectangle at 1/scale into host `out`."""
        fmt = PixelFormat(fmt)
        scale = Downscale(scale)
        dst = np.asarray(out)
        if not dst.flags["C_CONTIGUOUS"] or not dst.flags["WRITEABLE"]:
            raise BufferError(_lib.SGN_E_BUFFER, "output must be a writable contiguous array")
        s = self._sess()
        ptr, n, _keep = _buf(self._data)
        x, y, w, h = (int(v) for v in roi)
        _check(_lib.load().sgn_j2k_decode_region_into(s.handle, ptr, n, int(fmt), int(scale),
                                                      x, y, w, h, dst.ctypes.data, dst.nbytes,
                                                      int(stride)))
        return {"decoded": region_dims(self._data, roi, scale)}

    def decode_region_to_numpy(self, fmt, roi, scale=Downscale.NONE):
        fmt = PixelFormat(fmt)
        _x, _y, w, h = region_dims(self._data, roi, scale)
        out = np.zeros(max(1, h * w * fmt.bytes_per_pixel), np.uint8)
        self.decode_region_scaled_into(None, out, w * fmt.bytes_per_pixel, fmt, roi, scale)
        out = out[:h * w * fmt.bytes_per_pixel]
        arr = out.view("<u2") if fmt.bytes_per_sample == 2 else out
        return arr.reshape(h, w, fmt.channels)

    def decode_to_numpy(self, fmt, scale=Downscale.NONE):
        fmt = PixelFormat(fmt)
        w, h = scaled_dims(self._data, scale)
        out = np.zeros(h * w * fmt.bytes_per_pixel, np.uint8)
        self.decode_scaled_into(None, out, w * fmt.bytes_per_pixel, fmt, scale)
        arr = out.view("<u2") if fmt.bytes_per_sample == 2 else out
        return arr.reshape(h, w, fmt.channels)


class TilePlan:
    """A parsed batch resident in device memory: codestream bytes, the
    code-block table and the synthesis/packing jobs (the device form of
    J2kDirectColorPlan, crates/signinum-j2k-native/src/direct_plan.rs:12-205).
    decode() re-runs the whole batch without host work.

    `fmt` and `scale` are one value for the whole batch or one per tile
    (a heterogeneous batch of formats and reduced resolutions)."""

    def __init__(self, session, inputs, fmt, scale=Downscale.NONE, surfaces=None, rois=None):
        self.session = session
        self._inputs = [bytes(b) for b in inputs]
        n = len(self._inputs)
        fmts = list(fmt) if isinstance(fmt, (list, tuple)) else [fmt] * n
        scales = list(scale) if isinstance(scale, (list, tuple)) else [scale] * n
        self.rois = list(rois) if rois is not None else [None] * n
        if len(fmts) != n or len(scales) != n or len(self.rois) != n:
            raise ValueError("one format, scale and region per tile")
        self.formats = [PixelFormat(f) for f in fmts]
        self.scales = [Downscale(s) for s in scales]
        self.fmt = self.formats[0] if n else PixelFormat(fmt if not isinstance(fmt, (list, tuple)) else 0)
        self.scale = self.scales[0] if n else Downscale.NONE
        self._keep = [np.frombuffer(b, np.uint8) for b in self._inputs]
        if surfaces is None:
            dims = []
            for i, b in enumerate(self._inputs):
                try:
                    if self.rois[i] is None:
                        dims.append(scaled_dims(b, self.scales[i]))
                    else:
                        dims.append(region_dims(b, self.rois[i], self.scales[i])[2:])
                except J2kError as e:
                    raise TileBatchError(i, e)
            surfaces = _alloc_surfaces(session, dims, self.formats)
        self.surfaces = surfaces
        reqs = (_lib.TileReq * max(1, n))()
        for i, (k, s) in enumerate(zip(self._keep, surfaces)):
            roi = self.rois[i]
            reqs[i] = _lib.TileReq(k.ctypes.data, k.size, s.device_ptr, s.pitch_bytes, s.byte_len,
                                   int(self.formats[i]), int(self.scales[i]),
                                   0 if roi is None else 1,
                                   *((0, 0, 0, 0) if roi is None else (int(v) for v in roi)))
        status = (ctypes.c_int32 * max(1, n))()
        h = ctypes.c_void_p()
        rc = _lib.load().sgn_j2k_plan_create_mixed(session.handle, reqs, n, ctypes.byref(h), status)
        if rc != _lib.SGN_OK:
            bad = next((i for i in range(n) if status[i] != 0), 0)
            raise TileBatchError(bad, _ERRORS.get(rc, J2kError)(rc, _lib.last_error()))
        self.handle = h

    def decode(self, mode=MODE_GRAPH):
        _check(_lib.load().sgn_j2k_plan_decode(self.session.handle, self.handle, int(mode)))

    def decode_stage(self, stage):
        """One stage of the batched decode: 0 blocks, 1 levels, 2 final."""
        _check(_lib.load().sgn_j2k_plan_decode_stage(self.session.handle, self.handle, int(stage)))

    @staticmethod
    def lane_variants():
        """Names of the HT lane-kernel variants the autotuner chooses from."""
        L = _lib.load()
        return [L.sgn_ht_lane_variant_name(i).decode() for i in range(L.sgn_ht_lane_variant_count())]

    @property
    def lane_variant(self):
        return int(_lib.load().sgn_j2k_plan_get_lane_variant(self.handle))

    @lane_variant.setter
    def lane_variant(self, v):
        _check(_lib.load().sgn_j2k_plan_set_lane_variant(self.handle, int(v)))

    def autotune(self, reps=3):
        """Times every lane-kernel variant on this plan, keeps the fastest and
        returns {variant name: mean ms}."""
        L = _lib.load()
        n = L.sgn_ht_lane_variant_count()
        ms = (ctypes.c_float * n)()
        chosen = ctypes.c_int()
        _check(L.sgn_j2k_plan_autotune(self.session.handle, self.handle, int(reps), ms, n,
                                       ctypes.byref(chosen)))
        names = self.lane_variants()
        return {names[i]: float(ms[i]) for i in range(n)}

    @staticmethod
    def dwt_shapes():
        """Names of the synthesis-level launch shapes the autotuner chooses
        from."""
        L = _lib.load()
        return [L.sgn_dwt_shape_name(i).decode() for i in range(L.sgn_dwt_shape_count())]

    @property
    def dwt_shape(self):
        return int(_lib.load().sgn_j2k_plan_get_dwt_shape(self.handle))

    @dwt_shape.setter
    def dwt_shape(self, v):
        _check(_lib.load().sgn_j2k_plan_set_dwt_shape(self.handle, int(v)))

    def autotune_dwt(self, reps=5):
        """Times the level stage in every synthesis-level shape, keeps the
        fastest and returns {shape name: mean ms}."""
        L = _lib.load()
        n = L.sgn_dwt_shape_count()
        ms = (ctypes.c_float * n)()
        chosen = ctypes.c_int()
        _check(L.sgn_j2k_plan_autotune_dwt(self.session.handle, self.handle, int(reps), ms, n,
                                           ctypes.byref(chosen)))
        names = self.dwt_shapes()
        return {names[i]: float(ms[i]) for i in range(n)}

    ORDER_LENGTH, ORDER_INPUT, ORDER_MODEL = 0, 1, 2
    ORDER_GEOMETRY, ORDER_GEOMETRY_COST, ORDER_GEOMETRY_MODEL = 3, 4, 5

    def order_code_blocks(self, order, model=None):
        """Orders the lane code-blocks: ORDER_LENGTH (compressed-length
        heuristic), ORDER_INPUT, ORDER_MODEL (predicted decode cost),
        ORDER_GEOMETRY (block shape, input order within a shape),
        ORDER_GEOMETRY_COST / ORDER_GEOMETRY_MODEL (block shape, then tiles by
        summed heuristic / predicted block cost)."""
        _check(_lib.load().sgn_j2k_plan_order_code_blocks(
            self.session.handle, self.handle, int(order), model.handle if model else None))

    def autotune_order(self, reps=3):
        """Autotunes code-block order and lane-kernel variant together; keeps
        the fastest pair and returns ({order name: best ms}, chosen order)."""
        ms = (ctypes.c_float * 4)()
        chosen = ctypes.c_int()
        _check(_lib.load().sgn_j2k_plan_autotune_order(self.session.handle, self.handle,
                                                       int(reps), ms, ctypes.byref(chosen)))
        names = ("length", "input", "geometry", "geometry-cost")
        return {names[i]: float(ms[i]) for i in range(4)}, chosen.value

    def time_code_blocks(self, stride=1, max_samples=1 << 20):
        """(features [n, 5] float32, cycles [n] float32) of code-blocks decoded
        alone, one per wave, by the lane kernel."""
        st = self.stats()
        cap = min(max_samples, int(st["code_blocks"]) // max(1, stride) + 1)
        feats = np.zeros((cap, 5), np.float32)
        cyc = np.zeros(cap, np.float32)
        n = ctypes.c_size_t()
        _check(_lib.load().sgn_j2k_plan_time_code_blocks(
            self.session.handle, self.handle, int(stride), cap, feats.ctypes.data,
            cyc.ctypes.data, ctypes.byref(n)))
        return feats[:n.value].copy(), cyc[:n.value].copy()

    def stats(self):
        st = PlanStats()
        _check(_lib.load().sgn_j2k_plan_get_stats(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in PlanStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().sgn_j2k_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CostModel:
    """Code-block decode-cost model (5-32-32-1 ReLU MLP) trained and run on
    the device by the library's own kernels."""

    PARAMS = 1281

    def __init__(self, session, seed=1):
        self.session = session
        h = ctypes.c_void_p()
        _check(_lib.load().sgn_cost_model_create(session.handle, int(seed), ctypes.byref(h)))
        self.handle = h

    @property
    def params(self):
        out = np.zeros(self.PARAMS, np.float32)
        _check(_lib.load().sgn_cost_model_get_params(self.handle, out.ctypes.data, out.size))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.PARAMS:
            raise ValueError("1281 parameters expected")
        _check(_lib.load().sgn_cost_model_set_params(self.handle, p.ctypes.data, p.size))

    @staticmethod
    def _xy(X, Y=None):
        X = np.ascontiguousarray(X, np.float32).reshape(-1, 5)
        if Y is None:
            return X, None
        Y = np.ascontiguousarray(Y, np.float32).reshape(-1)
        if Y.size != X.shape[0]:
            raise ValueError("one target per sample")
        return X, Y

    def gradient(self, X, Y):
        X, Y = self._xy(X, Y)
        g = np.zeros(self.PARAMS, np.float32)
        _check(_lib.load().sgn_cost_model_gradient(self.session.handle, self.handle,
                                                   X.ctypes.data, Y.ctypes.data, X.shape[0],
                                                   g.ctypes.data))
        return g

    def train(self, X, Y, epochs=500, lr=1e-2):
        """Full-batch Adam on the device; returns the final mean loss."""
        X, Y = self._xy(X, Y)
        loss = ctypes.c_float()
        _check(_lib.load().sgn_cost_model_train(self.session.handle, self.handle, X.ctypes.data,
                                                Y.ctypes.data, X.shape[0], int(epochs),
                                                float(lr), ctypes.byref(loss)))
        return float(loss.value)

    def predict(self, X):
        X, _ = self._xy(X)
        out = np.zeros(X.shape[0], np.float32)
        _check(_lib.load().sgn_cost_model_predict(self.session.handle, self.handle,
                                                  X.ctypes.data, X.shape[0], out.ctypes.data))
        return out

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().sgn_cost_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decode_tiles_to_device(session, inputs, fmt, scale=Downscale.NONE, mode=MODE_BATCHED,
                           backend=BackendRequest.HIP):
    """TileBatchDecodeManyDevice::decode_tiles_to_device
    (crates/signinum-j2k-cuda/src/codec.rs:190-210): decodes every input in
    one batch and returns one DeviceSurface per input, in input order."""
    _route(backend)
    fmt = PixelFormat(fmt)
    scale = Downscale(scale)
    inputs = [bytes(b) for b in inputs]
    keep = [np.frombuffer(b, np.uint8) for b in inputs]
    dims = []
    for b in inputs:
        try:
            dims.append(scaled_dims(b, scale))
        except J2kError as e:
            raise TileBatchError(len(dims), e)
    surfaces = _alloc_surfaces(session, dims, [fmt] * len(inputs))
    n = len(inputs)
    jobs = (TileJob * max(1, n))()
    for i, (k, s) in enumerate(zip(keep, surfaces)):
        jobs[i] = TileJob(k.ctypes.data, k.size, s.device_ptr, s.pitch_bytes, s.byte_len)
    ff = ctypes.c_size_t(0)
    rc = _lib.load().sgn_j2k_decode_tiles_device(session.handle, jobs, n, int(fmt), int(scale),
                                                 int(mode), ctypes.byref(ff))
    if rc != _lib.SGN_OK:
        raise TileBatchError(ff.value, _ERRORS.get(rc, J2kError)(rc, _lib.last_error()))
    return surfaces


def decode_tiles_region_scaled_into(jobs, fmt, session=None):
    """decode_tiles_region_scaled_into (crates/signinum-j2k/src/batch.rs:
    151-195): jobs are (input_bytes, out_array, stride, roi, scale) with roi
    = (x, y, w, h) in source pixels; every region is decoded in one batch and
    written to its array.  Raises TileBatchError with the first failing
    index."""
    fmt = PixelFormat(fmt)
    own = session is None
    if own:
        session = HipSession()
    try:
        keep, dims = [], []
        for i, (inp, _out, _stride, roi, scale) in enumerate(jobs):
            keep.append(np.frombuffer(bytes(inp), np.uint8))
            try:
                dims.append(region_dims(keep[-1], roi, scale)[2:])
            except J2kError as e:
                raise TileBatchError(i, e)
        surfs = _alloc_surfaces(session, dims, [fmt] * len(jobs))
        n = len(jobs)
        reqs = (_lib.TileReq * max(1, n))()
        for i, ((_inp, _out, _stride, roi, scale), k, s) in enumerate(zip(jobs, keep, surfs)):
            reqs[i] = _lib.TileReq(k.ctypes.data, k.size, s.device_ptr, s.pitch_bytes, s.byte_len,
                                   int(fmt), int(Downscale(scale)), 1, *(int(v) for v in roi))
        ff = ctypes.c_size_t(0)
        rc = _lib.load().sgn_j2k_decode_tiles_region_device(session.handle, reqs, n,
                                                            ctypes.byref(ff))
        if rc != _lib.SGN_OK:
            raise TileBatchError(ff.value, _ERRORS.get(rc, J2kError)(rc, _lib.last_error()))
        for (_inp, out, stride, _roi, _scale), s in zip(jobs, surfs):
            if s.byte_len:
                s.download_into(out, stride)
    finally:
        if own:
            session.close()


def decode_tiles_into(jobs, fmt, session=None, scale=Downscale.NONE, mode=MODE_BATCHED):
    """decode_tiles_into (crates/signinum-j2k/src/batch.rs:110-147): jobs are
    (input_bytes, out_array, stride) triples; outputs are written in place.
    Raises TileBatchError with the first failing index."""
    own = session is None
    if own:
        session = HipSession()
    try:
        surfs = decode_tiles_to_device(session, [j[0] for j in jobs], fmt, scale, mode)
        for (inp, out, stride), s in zip(jobs, surfs):
            s.download_into(out, stride)
    finally:
        if own:
            session.close()


def default_encode_options():
    o = EncodeOptions()
    _lib.load().sgn_j2k_encode_default_options(ctypes.byref(o))
    return o


def encode_j2k_ht(samples, bit_depth=None, **options):
    """HTJ2K encode of an (H, W) or (H, W, C) uint8/uint16 array.  Options
    mirror sgn_j2k_encode_options: levels, cb_width_log2, cb_height_log2,
    reversible, use_mct, progression, guard_bits, base_step, precinct_log2,
    sop, eph, tile_width, tile_height."""
    arr = np.ascontiguousarray(samples)
    if arr.ndim == 2:
        arr = arr[:, :, None]
    h, w, c = arr.shape
    if bit_depth is None:
        bit_depth = 8 if arr.dtype == np.uint8 else 16
    if arr.dtype not in (np.uint8, np.uint16) or (bit_depth > 8) != (arr.dtype == np.uint16):
        raise J2kError(_lib.SGN_E_ARG, "samples must be uint8 (<= 8 bits) or uint16")
    o = default_encode_options()
    if "subsampling" in options and not isinstance(options["subsampling"], int):
        # [(dx, dy), ...] per component
        options = dict(options)
        options["subsampling"] = sum(((dx & 0xF) | (dy & 0xF) << 4) << (8 * c)
                                     for c, (dx, dy) in enumerate(options["subsampling"]))
    for k, v in options.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown encode option {k}")
        setattr(o, k, v)
    cap = arr.nbytes * 2 + 65536
    out = np.empty(cap, np.uint8)
    n = ctypes.c_size_t(0)
    lib = _lib.load()
    rc = lib.sgn_j2k_encode_ht(arr.ctypes.data, w, h, c, bit_depth, w * c * arr.itemsize,
                               ctypes.byref(o), out.ctypes.data, cap, ctypes.byref(n))
    if rc == _lib.SGN_E_BUFFER and n.value > cap:
        out = np.empty(n.value, np.uint8)
        rc = lib.sgn_j2k_encode_ht(arr.ctypes.data, w, h, c, bit_depth, w * c * arr.itemsize,
                                   ctypes.byref(o), out.ctypes.data, n.value, ctypes.byref(n))
    _check(rc)
    return out[:n.value].tobytes()


def encode_j2k_classic(samples, bit_depth=None, cb_style=0, **options):
    """Classic (EBCOT / MQ, Part 1) encode, every coding pass; cb_style is
    the code-block style (0x01 selective bypass, 0x02 context reset, 0x04
    termination on each pass, 0x08 vertically causal context, 0x20
    segmentation symbols).  Other options as encode_j2k_ht."""
    return encode_j2k_ht(samples, bit_depth=bit_depth, block_coder=1, cb_style=cb_style,
                         **options)


def encode_j2k_lossless(samples, bit_depth=None, levels=5, cb_log2=6, progression=0):
    """encode_j2k_lossless with HTJ2K block coding, reversible 5/3 and RCT
    (crates/signinum-j2k/src/encode.rs:166-182)."""
    return encode_j2k_ht(samples, bit_depth=bit_depth, levels=levels, cb_width_log2=cb_log2,
                         cb_height_log2=cb_log2, reversible=1, use_mct=1,
                         progression=progression)


def code_block_jobs(data):
    """HT code-block jobs of a full-resolution decode (sgn_j2k_code_blocks)."""
    ptr, n, _keep = _buf(data)
    lib = _lib.load()
    cnt = ctypes.c_size_t(0)
    _check(lib.sgn_j2k_code_blocks(ptr, n, None, 0, ctypes.byref(cnt)))
    arr = (HtCodeBlockJob * max(1, cnt.value))()
    _check(lib.sgn_j2k_code_blocks(ptr, n, arr, cnt.value, ctypes.byref(cnt)))
    return list(arr[:cnt.value])


class HtCodeBlockDecoder:
    """Stage hooks of HtCodeBlockDecoder (crates/signinum-j2k-native/src/
    lib.rs:832-909), executed on the device with host buffers."""

    def __init__(self, session):
        self.session = session

    def decode_code_blocks(self, data, jobs, out_len, raw=False):
        arr = (HtCodeBlockJob * max(1, len(jobs)))(*jobs)
        ptr, n, _keep = _buf(data)
        lib = _lib.load()
        if raw:
            out = np.zeros(out_len, np.uint32)
            _check(lib.sgn_ht_decode_code_blocks_raw(self.session.handle, ptr, n, arr, len(jobs),
                                                     out.ctypes.data, out_len))
        else:
            out = np.zeros(out_len, np.float32)
            _check(lib.sgn_ht_decode_code_blocks(self.session.handle, ptr, n, arr, len(jobs),
                                                 out.ctypes.data, out_len))
        return out

    def decode_sub_band(self, data, width, height, jobs, out=None):
        """decode_sub_band (lib.rs:864): jobs are (output_x, output_y,
        HtCodeBlockJob) of one width x height sub-band; returns the band as
        float32 (height, width), samples outside the blocks taken from `out`
        (zeros by default)."""
        arr = (_lib.HtCodeBlockBatchJob * max(1, len(jobs)))()
        for i, (x, y, j) in enumerate(jobs):
            arr[i] = _lib.HtCodeBlockBatchJob(int(x), int(y), j)
        band = np.zeros(width * height, np.float32) if out is None else \
            np.ascontiguousarray(out, np.float32).reshape(-1).copy()
        ptr, n, _keep = _buf(data)
        _check(_lib.load().sgn_ht_decode_sub_band(self.session.handle, ptr, n, int(width),
                                                  int(height), arr, len(jobs),
                                                  band.ctypes.data, band.size))
        return band.reshape(height, width)

    def decode_store_component(self, input, input_width, source_x, source_y, copy_width,
                               copy_height, output, output_width, output_x, output_y, addend):
        """decode_store_component (lib.rs:899, J2kStoreComponentJob
        lib.rs:805-828): returns a copy of `output` with the rectangle
        stored."""
        src = np.ascontiguousarray(input, np.float32).reshape(-1)
        dst = np.ascontiguousarray(output, np.float32).reshape(-1).copy()
        _check(_lib.load().sgn_store_component(
            self.session.handle, src.ctypes.data, src.size, int(input_width), int(source_x),
            int(source_y), int(copy_width), int(copy_height), dst.ctypes.data, dst.size,
            int(output_width), int(output_x), int(output_y), float(addend)))
        return dst

    def decode_single_decomposition_idwt(self, reversible, rect, ll, hl, lh, hh):
        x0, y0, x1, y1 = rect
        out = np.zeros((y1 - y0) * (x1 - x0), np.float32)
        bands = [np.ascontiguousarray(b, dtype=np.float32) for b in (ll, hl, lh, hh)]
        _check(_lib.load().sgn_idwt_level(self.session.handle, int(bool(reversible)), x0, y0, x1,
                                          y1, *[b.ctypes.data for b in bands], out.ctypes.data))
        return out.reshape(y1 - y0, x1 - x0)

    def decode_inverse_mct(self, reversible, p0, p1, p2, addends):
        planes = [np.ascontiguousarray(p, dtype=np.float32).copy() for p in (p0, p1, p2)]
        _check(_lib.load().sgn_inverse_mct(self.session.handle, int(bool(reversible)),
                                           *[p.ctypes.data for p in planes], planes[0].size,
                                           *[float(a) for a in addends]))
        return planes
