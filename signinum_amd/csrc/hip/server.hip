// Continuous-batching tile-decode server.
//
// No counterpart in the reference (its decode entry points are per call).
// Requests (1..n tiles of any size, bit depth, format and scale) are
// admitted as they arrive.  A batch is formed when the oldest waiting
// request has waited `window_us`, or when enough tiles wait; it takes
// waiting requests in arrival order while their device footprint (output
// surfaces, codestream arena, coefficient arena) fits the HBM budget.  Each
// batch is one plan decoded on one of `max_inflight` streams, so batches
// admitted while others run overlap on the device.  Outputs stay in device
// memory, owned by the server until the request is released.  The caller
// drives the server with sgn_server_pump(); nothing runs on hidden threads
// apart from the plan builder's own parse workers.
#include <chrono>
#include <deque>
#include <unordered_map>

namespace {

using ServerClock = std::chrono::steady_clock;

struct ServeRequest {
    uint64_t id = 0;
    std::vector<std::vector<uint8_t>> bytes;  // copies of the tile codestreams
    std::vector<int32_t> fmt;
    std::vector<uint32_t> den;
    std::vector<uint32_t> width, height;
    std::vector<size_t> row_bytes;
    std::vector<void *> out;  // device surfaces (one per tile), carved from out_block
    void *out_block = nullptr;
    uint64_t footprint = 0;   // admission estimate, bytes
    uint64_t out_bytes = 0;
    bool out_counted = false; // out_bytes included in the server's device usage
    ServerClock::time_point arrival, start, done;
    int32_t status = 0;
    uint32_t first_fail = 0;
    enum State { kPending, kRunning, kDone } state = kPending;
};

struct ServeBatch {
    std::unique_ptr<sgn_j2k_plan> plan;
    std::vector<uint64_t> ids;         // requests in plan order
    std::vector<size_t> tile_base;     // first plan tile of each request
    int slot = -1;
    hipEvent_t done = nullptr;
    uint64_t plan_bytes = 0;
};

uint64_t plan_device_bytes(const sgn_j2k_plan &P) { return P.device_bytes; }

}  // namespace

struct sgn_server {
    sgn_server_config cfg{};
    int device = 0;
    // plan and output blocks, reused across batches (declared before the
    // members whose blocks it takes back)
    std::unique_ptr<DeviceBlocks> blocks;
    std::vector<sgn_hip_session *> slots;  // one stream per in-flight batch
    std::vector<bool> slot_busy;
    std::unordered_map<uint64_t, std::unique_ptr<ServeRequest>> reqs;
    std::deque<uint64_t> pending;
    std::vector<std::unique_ptr<ServeBatch>> running;
    std::vector<sgn_request_info> finished;  // not yet polled
    uint64_t next_id = 1;
    uint64_t hbm_in_use = 0;
    sgn_server_stats stats{};

    ~sgn_server() {
        for (auto &b : running) {
            if (b->done) {
                (void)hipEventSynchronize(b->done);
                (void)hipEventDestroy(b->done);
            }
        }
        running.clear();
        if (blocks)
            for (auto &kv : reqs)
                if (kv.second->out_block) blocks->put(kv.second->out_block);
        reqs.clear();
        blocks.reset();
        for (sgn_hip_session *s : slots) sgn_hip_session_destroy(s);
    }
};

namespace {

double ms_between(ServerClock::time_point a, ServerClock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

void finish_request(sgn_server &S, ServeRequest &r, int32_t status, uint32_t first_fail) {
    r.state = ServeRequest::kDone;
    r.status = status;
    r.first_fail = first_fail;
    r.done = ServerClock::now();
    sgn_request_info info{};
    info.id = r.id;
    info.status = status;
    info.first_fail = first_fail;
    info.tiles = uint32_t(r.bytes.size());
    info.latency_ms = ms_between(r.arrival, r.done);
    info.queue_ms = ms_between(r.arrival, r.start);
    uint64_t px = 0;
    for (size_t t = 0; t < r.width.size(); ++t) px += uint64_t(r.width[t]) * r.height[t];
    info.pixels = px;
    S.finished.push_back(info);
    S.stats.requests_done++;
    if (status) S.stats.requests_failed++;
    // codestream copies are no longer needed; outputs stay until release
    for (auto &b : r.bytes) std::vector<uint8_t>().swap(b);
}

// Completes the batches whose event has fired.  Returns requests finished.
Status reap(sgn_server &S, uint32_t &completed) {
    for (size_t k = 0; k < S.running.size();) {
        ServeBatch &b = *S.running[k];
        const hipError_t q = hipEventQuery(b.done);
        if (q == hipErrorNotReady) {
            ++k;
            continue;
        }
        if (q != hipSuccess) {
            set_last_error(std::string("HIP: ") + hipGetErrorString(q) + " in a serving batch");
            return kErrDevice;
        }
        // per-tile block failures -> the owning request
        std::vector<uint8_t> bad_tile(b.plan->ntiles, 0);
        if (!b.plan->cb_lane.empty()) {
            std::vector<uint8_t> h(b.plan->cb_lane.size());
            HIP_TRY(hipMemcpy(h.data(), b.plan->d_status_lane, h.size(), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < h.size(); ++i)
                if (h[i]) bad_tile[b.plan->cb_lane[i].tile] = 1;
        }
        if (!b.plan->cb_wave.empty()) {
            std::vector<uint8_t> h(b.plan->cb_wave.size());
            HIP_TRY(hipMemcpy(h.data(), b.plan->d_status_wave, h.size(), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < h.size(); ++i)
                if (h[i]) bad_tile[b.plan->cb_wave[i].tile] = 1;
        }
        for (size_t r = 0; r < b.ids.size(); ++r) {
            ServeRequest &q2 = *S.reqs[b.ids[r]];
            const size_t t0 = b.tile_base[r];
            int32_t st = 0;
            uint32_t ff = 0;
            for (size_t t = 0; t < q2.bytes.size(); ++t)
                if (bad_tile[t0 + t]) {
                    st = SGN_E_BLOCK;
                    ff = uint32_t(t);
                    break;
                }
            finish_request(S, q2, st, ff);
            ++completed;
        }
        S.hbm_in_use -= b.plan_bytes;
        S.slot_busy[b.slot] = false;
        (void)hipEventDestroy(b.done);
        S.running.erase(S.running.begin() + long(k));
    }
    return kOk;
}

// Forms and launches one batch if the window or the tile count says so.
Status launch_batch(sgn_server &S, bool force) {
    if (S.pending.empty()) return kOk;
    int slot = -1;
    for (size_t i = 0; i < S.slot_busy.size(); ++i)
        if (!S.slot_busy[i]) {
            slot = int(i);
            break;
        }
    if (slot < 0) return kOk;
    const auto now = ServerClock::now();
    uint64_t waiting_tiles = 0;
    for (uint64_t id : S.pending) waiting_tiles += S.reqs[id]->bytes.size();
    const ServeRequest &oldest = *S.reqs[S.pending.front()];
    const bool due = force || waiting_tiles >= S.cfg.max_batch_tiles ||
                     ms_between(oldest.arrival, now) * 1000.0 >= double(S.cfg.window_us);
    if (!due) return kOk;
    // admission: arrival order, while tiles and device memory allow
    auto B = std::make_unique<ServeBatch>();
    std::vector<sgn_tile_req> tiles;
    uint64_t need = 0;
    size_t taken = 0;
    for (uint64_t id : S.pending) {
        ServeRequest &r = *S.reqs[id];
        if (taken && tiles.size() + r.bytes.size() > S.cfg.max_batch_tiles) break;
        if (S.hbm_in_use + need + r.footprint > S.cfg.hbm_budget) break;
        B->ids.push_back(id);
        B->tile_base.push_back(tiles.size());
        for (size_t t = 0; t < r.bytes.size(); ++t)
            tiles.push_back(sgn_tile_req{r.bytes[t].data(), r.bytes[t].size(), nullptr,
                                         r.row_bytes[t], r.row_bytes[t] * r.height[t], r.fmt[t],
                                         r.den[t]});
        need += r.footprint;
        ++taken;
    }
    if (!taken) return kOk;  // the oldest request waits for memory to free up
    // outputs
    size_t ti = 0;
    for (uint64_t id : B->ids) {
        ServeRequest &r = *S.reqs[id];
        if (!r.out_block) {
            // one allocation per request, surfaces at 256-byte alignment
            size_t total = 0;
            for (size_t t = 0; t < r.bytes.size(); ++t)
                total += (std::max<size_t>(16, r.row_bytes[t] * r.height[t]) + 255) & ~size_t(255);
            HIP_TRY(S.blocks->get(total, &r.out_block));
            size_t o = 0;
            for (size_t t = 0; t < r.bytes.size(); ++t) {
                r.out[t] = static_cast<uint8_t *>(r.out_block) + o;
                o += (std::max<size_t>(16, r.row_bytes[t] * r.height[t]) + 255) & ~size_t(255);
            }
        }
        for (size_t t = 0; t < r.bytes.size(); ++t, ++ti) tiles[ti].dst = r.out[t];
        if (!r.out_counted) S.hbm_in_use += r.out_bytes;
        r.out_counted = true;
    }
    std::vector<int32_t> ts(tiles.size(), 0);
    sgn_hip_session *sess = S.slots[size_t(slot)];
    Status st = make_plan(sess, tiles.data(), tiles.size(), B->plan, ts.data(), S.blocks.get());
    if (st) {
        // parse failures fail their own requests; the rest retry next pump
        size_t r = 0;
        bool any = false;
        for (uint64_t id : B->ids) {
            ServeRequest &q = *S.reqs[id];
            const size_t t0 = B->tile_base[r++];
            for (size_t t = 0; t < q.bytes.size(); ++t)
                if (ts[t0 + t]) {
                    q.start = ServerClock::now();
                    S.pending.erase(std::find(S.pending.begin(), S.pending.end(), id));
                    finish_request(S, q, ts[t0 + t], uint32_t(t));
                    any = true;
                    break;
                }
        }
        if (!any) return st;  // not attributable to a request
        return kOk;
    }
    B->plan_bytes = plan_device_bytes(*B->plan);
    S.hbm_in_use += B->plan_bytes;
    st = decode_plan(sess, *B->plan, SGN_MODE_BATCHED);
    if (st) return st;
    HIP_TRY(hipEventCreateWithFlags(&B->done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(B->done, sess->stream));
    const auto t_start = ServerClock::now();
    for (uint64_t id : B->ids) {
        ServeRequest &r = *S.reqs[id];
        r.state = ServeRequest::kRunning;
        r.start = t_start;
        S.pending.pop_front();  // taken in arrival order from the front
    }
    B->slot = slot;
    S.slot_busy[size_t(slot)] = true;
    S.stats.batches++;
    S.stats.tiles += tiles.size();
    S.running.push_back(std::move(B));
    return kOk;
}

}  // namespace

void sgn_server_default_config(sgn_server_config *c) {
    if (!c) return;
    c->window_us = 2000;
    c->hbm_budget = uint64_t(64) << 30;
    c->max_batch_tiles = 4096;
    c->max_inflight = 4;
}

int sgn_server_create(int device, const sgn_server_config *cfg, sgn_server **out) {
    API_GUARD_BEGIN
    if (!out) return SGN_E_ARG;
    *out = nullptr;
    auto S = std::make_unique<sgn_server>();
    sgn_server_default_config(&S->cfg);
    if (cfg) S->cfg = *cfg;
    if (S->cfg.max_inflight < 1 || S->cfg.max_inflight > 16 || S->cfg.max_batch_tiles < 1) {
        set_last_error("server needs 1..16 in-flight batches and at least one tile per batch");
        return SGN_E_ARG;
    }
    S->device = device;
    // cached free blocks stay within a quarter of the budget
    S->blocks = std::make_unique<DeviceBlocks>(S->cfg.hbm_budget / 4);
    for (uint32_t i = 0; i < S->cfg.max_inflight; ++i) {
        sgn_hip_session *s = nullptr;
        const int rc = sgn_hip_session_create(device, &s);
        if (rc) return rc;
        S->slots.push_back(s);
    }
    S->slot_busy.assign(S->slots.size(), false);
    *out = S.release();
    return SGN_OK;
    API_GUARD_END
}

void sgn_server_destroy(sgn_server *S) { delete S; }

int sgn_server_submit(sgn_server *S, const sgn_serve_tile *tiles, size_t n, uint64_t *id) {
    API_GUARD_BEGIN
    if (!S || !tiles || !n || !id) return SGN_E_ARG;
    auto r = std::make_unique<ServeRequest>();
    r->arrival = ServerClock::now();
    for (size_t t = 0; t < n; ++t) {
        const sgn_serve_tile &q = tiles[t];
        if (!q.bytes && q.len) return SGN_E_ARG;
        Container c;
        Status st = locate_codestream(q.bytes, q.len, c);
        if (st || (st = device_decodable(c))) return st;
        MainHeader h;
        if ((st = read_main_header(c.cs, c.cs_len, h))) return st;
        OutputGeom og;
        if ((st = output_geometry(h, q.fmt, q.denominator, og))) return st;
        r->bytes.emplace_back(c.cs, c.cs + c.cs_len);
        r->fmt.push_back(q.fmt);
        r->den.push_back(q.denominator);
        r->width.push_back(og.width);
        r->height.push_back(og.height);
        r->row_bytes.push_back(og.row_bytes);
        r->out.push_back(nullptr);
        const uint64_t out_b = uint64_t(og.row_bytes) * og.height;
        const uint64_t samples = uint64_t(h.xsiz - h.xo) * (h.ysiz - h.yo) * h.ncomp;
        r->out_bytes += out_b;
        // output + codestream + int32 coefficients (bands and levels) + tables
        r->footprint += out_b + c.cs_len + samples * 4 * 4 / 3 + samples / 8 + 256;
    }
    if (r->footprint > S->cfg.hbm_budget) {
        set_last_error("request larger than the server's device-memory budget");
        return SGN_E_BUFFER;
    }
    r->id = S->next_id++;
    *id = r->id;
    S->pending.push_back(r->id);
    S->stats.requests_submitted++;
    S->reqs[r->id] = std::move(r);
    return SGN_OK;
    API_GUARD_END
}

int sgn_server_pump(sgn_server *S, int flush, uint32_t *completed) {
    API_GUARD_BEGIN
    if (!S) return SGN_E_ARG;
    HIP_TRY(hipSetDevice(S->device));
    uint32_t done = 0;
    Status st = reap(*S, done);
    if (st) return st;
    // several batches may start in one pump when slots are free
    for (size_t k = 0; k < S->slots.size(); ++k) {
        const size_t before = S->running.size();
        if ((st = launch_batch(*S, flush != 0))) return st;
        if (S->running.size() == before) break;
    }
    if (completed) *completed = done;
    return SGN_OK;
    API_GUARD_END
}

int sgn_server_poll(sgn_server *S, sgn_request_info *out, size_t cap, size_t *n) {
    if (!S || !n || (!out && cap)) return SGN_E_ARG;
    const size_t k = std::min(cap, S->finished.size());
    for (size_t i = 0; i < k; ++i) out[i] = S->finished[i];
    S->finished.erase(S->finished.begin(), S->finished.begin() + long(k));
    *n = k;
    return SGN_OK;
}

int sgn_server_output(sgn_server *S, uint64_t id, size_t tile, void **dev, size_t *pitch,
                      uint32_t *width, uint32_t *height) {
    if (!S) return SGN_E_ARG;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end() || tile >= it->second->out.size()) return SGN_E_ARG;
    const ServeRequest &r = *it->second;
    if (r.state != ServeRequest::kDone) {
        set_last_error("request not finished");
        return SGN_E_ARG;
    }
    if (dev) *dev = r.out[tile];
    if (pitch) *pitch = r.row_bytes[tile];
    if (width) *width = r.width[tile];
    if (height) *height = r.height[tile];
    return SGN_OK;
}

int sgn_server_release(sgn_server *S, uint64_t id) {
    if (!S) return SGN_E_ARG;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end() || it->second->state != ServeRequest::kDone) return SGN_E_ARG;
    ServeRequest &r = *it->second;
    (void)hipSetDevice(S->device);
    if (r.out_block) S->blocks->put(r.out_block);
    if (r.out_counted) S->hbm_in_use -= r.out_bytes;
    S->reqs.erase(it);
    return SGN_OK;
}

int sgn_server_get_stats(const sgn_server *S, sgn_server_stats *out) {
    if (!S || !out) return SGN_E_ARG;
    *out = S->stats;
    out->pending = S->pending.size();
    out->running_batches = S->running.size();
    out->hbm_in_use = S->hbm_in_use;
    return SGN_OK;
}
