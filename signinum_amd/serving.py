"""Continuous-batching server (sgn_server_*) and an open-loop Poisson load
generator for it."""
import ctypes
import time

import numpy as np

from . import _lib, _check, PixelFormat, Downscale, HipSession


class Server:
    """Requests of 1..n tiles are admitted as they arrive, batched over a
    window, decoded on several streams at once, with outputs kept in device
    memory until release().  Drive it with pump()."""

    def __init__(self, device=0, window_us=2000, hbm_budget=64 << 30, max_batch_tiles=4096,
                 max_inflight=4):
        self._lib = _lib.load()
        cfg = _lib.ServerConfig(int(window_us), int(hbm_budget), int(max_batch_tiles),
                                int(max_inflight))
        h = ctypes.c_void_p()
        _check(self._lib.sgn_server_create(int(device), ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.device = device
        self._sess = None

    def submit(self, tiles):
        """tiles: iterable of (codestream bytes, PixelFormat, Downscale).
        Returns the request id.  The server copies the codestreams."""
        tiles = list(tiles)
        keep = [np.frombuffer(t[0], np.uint8) for t in tiles]
        arr = (_lib.ServeTile * len(tiles))()
        for i, ((_b, fmt, den), k) in enumerate(zip(tiles, keep)):
            arr[i] = _lib.ServeTile(k.ctypes.data, k.size, int(PixelFormat(fmt)),
                                    int(Downscale(den)))
        rid = ctypes.c_uint64()
        _check(self._lib.sgn_server_submit(self.handle, arr, len(tiles), ctypes.byref(rid)))
        return int(rid.value)

    def pump(self, flush=False):
        n = ctypes.c_uint32()
        _check(self._lib.sgn_server_pump(self.handle, 1 if flush else 0, ctypes.byref(n)))
        return int(n.value)

    def poll(self, cap=4096):
        out = (_lib.RequestInfo * cap)()
        n = ctypes.c_size_t()
        _check(self._lib.sgn_server_poll(self.handle, out, cap, ctypes.byref(n)))
        return [{f: getattr(out[i], f) for f, _ in _lib.RequestInfo._fields_}
                for i in range(n.value)]

    def output(self, rid, tile):
        """(device pointer, pitch, width, height) of a finished tile."""
        p, pitch = ctypes.c_void_p(), ctypes.c_size_t()
        w, h = ctypes.c_uint32(), ctypes.c_uint32()
        _check(self._lib.sgn_server_output(self.handle, int(rid), int(tile), ctypes.byref(p),
                                           ctypes.byref(pitch), ctypes.byref(w), ctypes.byref(h)))
        return p.value, pitch.value, w.value, h.value

    def download(self, rid, tile, fmt):
        """Host copy (H, W, C) of a finished tile."""
        fmt = PixelFormat(fmt)
        p, pitch, w, h = self.output(rid, tile)
        if self._sess is None:
            self._sess = HipSession(self.device)
        out = np.zeros(h * w * fmt.bytes_per_pixel, np.uint8)
        if out.size:
            _check(self._lib.sgn_device_download_2d(self._sess.handle, out.ctypes.data,
                                                    w * fmt.bytes_per_pixel, p, pitch,
                                                    w * fmt.bytes_per_pixel, h))
        if fmt.bytes_per_sample == 2:
            out = out.view("<u2")
        return out.reshape(h, w, fmt.channels)

    def release(self, rid):
        _check(self._lib.sgn_server_release(self.handle, int(rid)))

    def stats(self):
        st = _lib.ServerStats()
        _check(self._lib.sgn_server_get_stats(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in _lib.ServerStats._fields_}

    def wait(self, ids, timeout_s=60.0):
        """Pumps (flushing) until every id in `ids` finished; returns their records."""
        want, got = set(ids), {}
        t_end = time.time() + timeout_s
        while want - set(got):
            self.pump(flush=True)
            for r in self.poll():
                got[r["id"]] = r
            if time.time() > t_end:
                raise TimeoutError("requests did not finish")
        return got

    def close(self):
        if getattr(self, "handle", None):
            self._lib.sgn_server_destroy(self.handle)
            self.handle = None
        if self._sess is not None:
            self._sess.close()
            self._sess = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_poisson(server, pool, rate_per_s, duration_s, seed=0, max_tiles=64, release=True):
    """Open-loop load: requests arrive as a Poisson process at `rate_per_s`;
    each takes 1..max_tiles tiles drawn from `pool` (list of (bytes, fmt,
    scale, ...)).  A request's latency runs from its scheduled arrival (not
    from the moment the loop got round to submitting it) to the pump that
    observed its completion.  Returns (records with "latency_ms", wall time
    from the first arrival to the last completion)."""
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    next_at = t0 + rng.exponential(1.0 / rate_per_s)
    t_stop = t0 + duration_s
    records, arrival = [], {}
    submitted = 0
    while True:
        now = time.perf_counter()
        if now >= next_at and next_at < t_stop:
            k = int(rng.integers(1, max_tiles + 1))
            pick = rng.integers(0, len(pool), k)
            rid = server.submit([pool[i][:3] for i in pick])
            arrival[rid] = next_at
            submitted += 1
            next_at += rng.exponential(1.0 / rate_per_s)
            continue
        server.pump(flush=now >= t_stop)
        done = server.poll()
        t_done = time.perf_counter()
        for r in done:
            r["server_latency_ms"] = r["latency_ms"]
            r["latency_ms"] = (t_done - arrival[r["id"]]) * 1e3
            records.append(r)
            if release:
                server.release(r["id"])
        if next_at >= t_stop and len(records) == submitted:
            break
    return records, time.perf_counter() - t0
