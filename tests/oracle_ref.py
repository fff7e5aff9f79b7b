"""ctypes binding of the CPU oracle (oracle/liboracle_j2k.so).

Test infrastructure: the oracle is the checker for the HIP path, never the
thing measured or shipped.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_L = None

FMT_CHANNELS = {0: 1, 1: 3, 2: 4, 3: 1, 4: 3}
FMT_BPS = {0: 1, 1: 1, 2: 1, 3: 2, 4: 2}


class Info(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint32) for n in (
        "width", "height", "num_components", "bit_depth", "num_levels", "reversible", "mct",
        "high_throughput", "tiles_x", "tiles_y", "cb_width", "cb_height", "progression")]


def lib():
    global _L
    if _L is None:
        L = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle_j2k.so"))
        sz = ctypes.c_size_t
        vp = ctypes.c_void_p
        u32 = ctypes.c_uint32
        L.oracle_decode.argtypes = [vp, sz, ctypes.c_int, u32, vp, sz, sz]
        L.oracle_decode_planes.argtypes = [vp, sz, u32, ctypes.c_int, vp, sz]
        L.oracle_scaled_dims.argtypes = [vp, sz, u32, ctypes.POINTER(u32), ctypes.POINTER(u32)]
        L.oracle_inspect.argtypes = [vp, sz, ctypes.POINTER(Info)]
        L.oracle_decode_batch.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(sz), sz, ctypes.c_int,
                                          u32, ctypes.POINTER(vp), ctypes.POINTER(sz),
                                          ctypes.POINTER(sz), ctypes.c_int, ctypes.POINTER(sz)]
        L.oracle_ht_decode_block.argtypes = [vp, u32, u32, u32, u32, u32, u32, ctypes.c_int,
                                             ctypes.c_int, vp, u32]
        L.oracle_coefficient_to_i32.argtypes = [u32, u32]
        L.oracle_coefficient_to_i32.restype = ctypes.c_int32
        L.oracle_last_stats.argtypes = [ctypes.POINTER(u32)] * 3
        L.oracle_set_zero_refinement.argtypes = [ctypes.c_int]
        L.oracle_idwt_level.argtypes = [ctypes.c_int, u32, u32, u32, u32, vp, vp, vp, vp, vp]
        L.oracle_inverse_mct.argtypes = [ctypes.c_int, vp, vp, vp, sz, ctypes.c_float,
                                         ctypes.c_float, ctypes.c_float]
        L.oracle_get_tables.argtypes = [vp, vp, vp, vp]
        L.oracle_mlp_predict.argtypes = [vp, vp]
        L.oracle_mlp_predict.restype = ctypes.c_double
        L.oracle_mlp_loss.argtypes = [vp, vp, vp, sz]
        L.oracle_mlp_loss.restype = ctypes.c_double
        L.oracle_mlp_loss_grad.argtypes = [vp, vp, vp, sz, vp]
        L.oracle_mlp_loss_grad.restype = ctypes.c_double
        _L = L
    return _L


def _ptr(b):
    a = np.frombuffer(b, np.uint8)
    return a.ctypes.data, a.size, a


def scaled_dims(cs, den=1):
    p, n, _k = _ptr(cs)
    w = ctypes.c_uint32()
    h = ctypes.c_uint32()
    rc = lib().oracle_scaled_dims(p, n, den, ctypes.byref(w), ctypes.byref(h))
    assert rc == 0, rc
    return w.value, h.value


def inspect(cs):
    p, n, _k = _ptr(cs)
    info = Info()
    rc = lib().oracle_inspect(p, n, ctypes.byref(info))
    assert rc == 0, rc
    return {f: getattr(info, f) for f, _ in Info._fields_}


def decode(cs, fmt, den=1):
    """(H, W, C) uint8 / uint16 array, or raises on a decode error."""
    w, h = scaled_dims(cs, den)
    ch, bps = FMT_CHANNELS[fmt], FMT_BPS[fmt]
    out = np.zeros(w * h * ch * bps, np.uint8)
    p, n, _k = _ptr(cs)
    rc = lib().oracle_decode(p, n, fmt, den, out.ctypes.data, out.size, w * ch * bps)
    if rc != 0:
        raise RuntimeError(f"oracle decode failed: {rc}")
    if bps == 2:
        out = out.view("<u2")
    return out.reshape(h, w, ch)


def decode_rc(cs, fmt, den=1):
    w, h = scaled_dims(cs, den)
    ch, bps = FMT_CHANNELS[fmt], FMT_BPS[fmt]
    out = np.zeros(max(1, w * h * ch * bps), np.uint8)
    p, n, _k = _ptr(cs)
    return lib().oracle_decode(p, n, fmt, den, out.ctypes.data, out.size, w * ch * bps)


def decode_planes(cs, den=1, use_f64=False):
    info = inspect(cs)
    w, h = scaled_dims(cs, den)
    c = info["num_components"]
    out = np.zeros(w * h * c, np.float64)
    p, n, _k = _ptr(cs)
    rc = lib().oracle_decode_planes(p, n, den, int(use_f64), out.ctypes.data, out.size)
    assert rc == 0, rc
    return out.reshape(c, h, w)


def last_stats():
    v = [ctypes.c_uint32() for _ in range(3)]
    lib().oracle_last_stats(*[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def set_zero_refinement(flag):
    lib().oracle_set_zero_refinement(int(bool(flag)))


def ht_decode_block(data, lcup, lref, w, h, zblk, passes, stripe_causal=False, phase_limit=2):
    out = np.zeros(max(1, w * h), np.uint32)
    buf = np.frombuffer(bytes(data), np.uint8)
    rc = lib().oracle_ht_decode_block(buf.ctypes.data, lcup, lref, w, h, zblk, passes,
                                      int(stripe_causal), phase_limit, out.ctypes.data, w)
    return rc, out[:w * h].reshape(h, w)


def coefficient_to_i32(v, k):
    return lib().oracle_coefficient_to_i32(v, k)


def idwt_level(reversible, rect, ll, hl, lh, hh):
    x0, y0, x1, y1 = rect
    out = np.zeros((y1 - y0) * (x1 - x0), np.float32)
    bands = [np.ascontiguousarray(b, np.float32) for b in (ll, hl, lh, hh)]
    rc = lib().oracle_idwt_level(int(reversible), x0, y0, x1, y1,
                                 *[b.ctypes.data for b in bands], out.ctypes.data)
    assert rc == 0
    return out.reshape(y1 - y0, x1 - x0)


def inverse_mct(reversible, p0, p1, p2, addends):
    planes = [np.ascontiguousarray(p, np.float32).copy() for p in (p0, p1, p2)]
    lib().oracle_inverse_mct(int(reversible), *[p.ctypes.data for p in planes], planes[0].size,
                             *[float(a) for a in addends])
    return planes


def decode_batch(streams, fmt, den=1, workers=1):
    n = len(streams)
    keep = [np.frombuffer(s, np.uint8) for s in streams]
    def dims_or_zero(s):
        try:
            return scaled_dims(s, den)
        except AssertionError:
            return (0, 0)
    dims = [dims_or_zero(s) for s in streams]
    ch, bps = FMT_CHANNELS[fmt], FMT_BPS[fmt]
    outs = [np.zeros(max(1, w * h * ch * bps), np.uint8) for (w, h) in dims]
    VP = ctypes.c_void_p * n
    SZ = ctypes.c_size_t * n
    datas = VP(*[k.ctypes.data for k in keep])
    lens = SZ(*[k.size for k in keep])
    optrs = VP(*[o.ctypes.data for o in outs])
    olens = SZ(*[o.size for o in outs])
    strides = SZ(*[w * ch * bps for (w, h) in dims])
    ff = ctypes.c_size_t(0)
    rc = lib().oracle_decode_batch(datas, lens, n, fmt, den, optrs, olens, strides, workers,
                                   ctypes.byref(ff))
    return rc, ff.value, outs


def tables():
    v0 = np.zeros(1024, np.uint16)
    v1 = np.zeros(1024, np.uint16)
    u0 = np.zeros(320, np.uint16)
    u1 = np.zeros(256, np.uint16)
    lib().oracle_get_tables(v0.ctypes.data, v1.ctypes.data, u0.ctypes.data, u1.ctypes.data)
    return v0, v1, u0, u1


class OracleCb(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_uint32) for n in ("tile", "comp", "res", "band", "x0", "y0", "x1",
                                                 "y1")] + [("offset", ctypes.c_int64)] +
                [(n, ctypes.c_uint32) for n in ("lcup", "lref", "zblk", "passes")])


def code_blocks(cs):
    L = lib()
    L.oracle_code_blocks.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(OracleCb),
                                     ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    p, n, _k = _ptr(cs)
    cnt = ctypes.c_size_t()
    rc = L.oracle_code_blocks(p, n, None, 0, ctypes.byref(cnt))
    assert rc == 0, rc
    arr = (OracleCb * max(1, cnt.value))()
    rc = L.oracle_code_blocks(p, n, arr, cnt.value, ctypes.byref(cnt))
    assert rc == 0, rc
    return list(arr[:cnt.value])


def mlp_loss_grad(params, X, Y):
    """(mean loss, gradient) of the decode-cost MLP in double precision."""
    p = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 5)
    Y = np.ascontiguousarray(Y, np.float32).reshape(-1)
    g = np.zeros(1281, np.float64)
    loss = lib().oracle_mlp_loss_grad(p.ctypes.data, X.ctypes.data, Y.ctypes.data, X.shape[0],
                                      g.ctypes.data)
    return loss, g


def mlp_loss(params, X, Y):
    p = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 5)
    Y = np.ascontiguousarray(Y, np.float32).reshape(-1)
    return lib().oracle_mlp_loss(p.ctypes.data, X.ctypes.data, Y.ctypes.data, X.shape[0])


def mlp_predict(params, X):
    p = np.ascontiguousarray(params, np.float64)
    X = np.ascontiguousarray(X, np.float32).reshape(-1, 5)
    return np.array([lib().oracle_mlp_predict(p.ctypes.data, X[i].ctypes.data)
                     for i in range(X.shape[0])])
