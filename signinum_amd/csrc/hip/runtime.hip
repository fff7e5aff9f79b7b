// HIP runtime of signinum-j2k-hip: sessions, batch plans, launch sequences
// and the C ABI entry points that need the device.
//
// The role matches the reference's CUDA runtime helper
// (crates/signinum-cuda-runtime/src/lib.rs:281-775: context, buffers,
// launches) and the adapter's decode entry points
// (crates/signinum-j2k-cuda/src/decoder.rs:17-280, codec.rs:18-210), but the
// decode itself runs on the device instead of uploading CPU-decoded pixels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <unordered_map>
#include <new>
#include <mutex>
#include <thread>
#include <vector>

#include "../host/parse.hpp"
#include "../host/sgn_api.hpp"
#include "ht_decode.hip"
#include "ht_coop.hip"
#include "ht_rows.hip"
#include "idwt.hip"
#include "t1_classic.hip"
#include "packets.hpp"
#include "cost_model.hip"

using namespace sgn;

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            set_last_error(std::string("HIP: ") + hipGetErrorString(e_) + " at " #expr); \
            return sgn::kErrDevice;                                                \
        }                                                                          \
    } while (0)

class DeviceBlocks;

// Device memory of sgn_j2k_decode_tiles_device_noalloc: every plan that
// entry point decodes (codestream arena, job tables, status words, work
// counters, coefficient arena) is carved from this array, so the call
// allocates nothing.  Each GPU that loads the library holds its own copy
// until the process ends.  Host code must never pass the array's name to a
// kernel or a copy (on the host it is only a stand-in): sessions resolve
// each device's address with hipGetSymbolAddress (noalloc_scratch()).
constexpr uint64_t kNoAllocScratchBytes = uint64_t(64) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// Pinned host memory (hipHostMalloc) that plan uploads stage their tables
// and non-pinned codestream bytes in, kept by the session and grown on
// demand.  `done` is recorded after the last copy out of it was enqueued;
// a new fill waits for it.
struct PinnedSlot {
    uint8_t *p = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};

struct sgn_hip_session {
    int device = 0;
    hipStream_t stream = nullptr;
    uint16_t *d_tables = nullptr;
    uint64_t dispatches = 0;
    int cus = 0;                    // compute units of the device
    // per lane-kernel variant and sample type (2 v + narrow): resident
    // work-groups (0: unknown)
    std::vector<uint32_t> resident;
    uint32_t classic_resident[4][3] = {};  // resident work-groups per classic kernel (shape, lanes)
    // pipelined batch decode: host-to-device copies run on copy_stream
    // while the previous chunk decodes on `stream`; two staging slots
    hipStream_t copy_stream = nullptr;
    hipStream_t aux_stream = nullptr;  // second decode stream: chunks alternate
    PinnedSlot slot[2];
    DeviceBlocks *blocks = nullptr;  // device blocks of pipelined chunk plans, reused
};

namespace {

// Lane-kernel variants the autotuner chooses from: work-group shape, quad
// pairs per segment (and MagSgn ring slots), and one-shot grid versus
// persistent waves pulling 64-block chunks from a device work counter.
// refine: the list holds blocks with SigProp / MagRef passes (the lane
// kernel handles them in any case; the cooperative one needs its refining
// instance)
using RowsLaunch = void (*)(hipStream_t, uint32_t, const CbJob *, uint32_t, const uint8_t *,
                            void *, const uint16_t *, uint8_t *, uint32_t *, bool);

template <int W, int S, int M, int VS, int LS, bool PERSIST, typename C, int MINW, int CHUNK>
void rows_launch(hipStream_t st, uint32_t grid, const CbJob *jobs, uint32_t n,
                 const uint8_t *arena, void *coef, const uint16_t *tables, uint8_t *status,
                 uint32_t *queue, bool) {
    hipLaunchKernelGGL((sgn_ht_rows<W, S, M, VS, LS, PERSIST, C, MINW, CHUNK>), dim3(grid),
                       dim3(W * 64), 0, st, jobs, n, arena, static_cast<C *>(coef), tables, status,
                       queue);
}

// The cooperative kernel as a lane-list variant: one code-block per wave,
// one wave per work-group (the grid is the block count).
template <typename C>
void coop_launch(hipStream_t st, uint32_t, const CbJob *jobs, uint32_t n, const uint8_t *arena,
                 void *coef, const uint16_t *tables, uint8_t *status, uint32_t *, bool refine) {
    if (refine)
        hipLaunchKernelGGL((sgn_ht_coop<64, true, C>), dim3(n), dim3(64), 0, st, jobs, n, arena,
                           static_cast<C *>(coef), tables, status, 0u);
    else
        hipLaunchKernelGGL((sgn_ht_coop<64, false, C>), dim3(n), dim3(64), 0, st, jobs, n, arena,
                           static_cast<C *>(coef), tables, status, 0u);
}

// launch[0] / kernel[0]: int32 samples; [1]: int16 samples (narrow plans)
struct RowsVariant {
    const char *name;
    int waves;
    bool persist;
    RowsLaunch launch[2];
    const void *kernel[2];
    bool coop = false;  // one code-block per wave (grid = block count)
};

// MINW: waves per SIMD the register allocation must allow (1: no bound);
// CHUNK: code-blocks a persistent wave takes per work-counter grab
#define SGN_ROWS_VARIANT(name, W, S, M, VS, LS, Q, MINW, CHUNK)                              \
    {name,                                                                                 \
     W,                                                                                    \
     Q,                                                                                    \
     {rows_launch<W, S, M, VS, LS, Q, int32_t, MINW, CHUNK>,                               \
      rows_launch<W, S, M, VS, LS, Q, int16_t, MINW, CHUNK>},                              \
     {reinterpret_cast<const void *>(sgn_ht_rows<W, S, M, VS, LS, Q, int32_t, MINW, CHUNK>), \
      reinterpret_cast<const void *>(sgn_ht_rows<W, S, M, VS, LS, Q, int16_t, MINW, CHUNK>)}}

// name: waves per work-group, quad pairs per segment, small rings (3 work-
// groups per CU instead of 2), persistent queue and its chunk
const RowsVariant kRowsVariants[] = {
    SGN_ROWS_VARIANT("w4-seg8", 4, 8, 8, 4, 2, false, 1, 64),
    SGN_ROWS_VARIANT("w4-seg8-queue", 4, 8, 8, 4, 2, true, 1, 64),
    SGN_ROWS_VARIANT("w4-seg4-queue", 4, 4, 4, 4, 2, true, 1, 64),
    SGN_ROWS_VARIANT("w4-seg4-small-queue", 4, 4, 4, 2, 1, true, 1, 64),
    SGN_ROWS_VARIANT("w2-seg16-queue", 2, 16, 16, 4, 2, true, 1, 64),
    SGN_ROWS_VARIANT("w1-seg8-queue", 1, 8, 8, 4, 2, true, 1, 64),
    {"wave-coop",
     1,
     false,
     {coop_launch<int32_t>, coop_launch<int16_t>},
     {reinterpret_cast<const void *>(sgn_ht_coop<64, false, int32_t>),
      reinterpret_cast<const void *>(sgn_ht_coop<64, false, int16_t>)},
     true},
    SGN_ROWS_VARIANT("w4-seg4-small-queue128", 4, 4, 4, 2, 1, true, 1, 128),
    SGN_ROWS_VARIANT("w4-seg4-small-queue256", 4, 4, 4, 2, 1, true, 1, 256),
};
#undef SGN_ROWS_VARIANT
constexpr int kRowsVariantCount = int(sizeof(kRowsVariants) / sizeof(kRowsVariants[0]));
// variant of a plan nobody autotuned: w4-seg4-small-queue, the autotuner's
// choice on every batch measured (small rings, three work-groups per CU)
constexpr int kDefaultLaneVariant = 3;
// the cooperative kernel (one code-block per wave)
constexpr int kCoopVariant = 6;

// Work-group shapes of the synthesis-level launches the autotuner chooses
// from: the 64x64-block kernel (four or two waves per block) and whether
// launches whose levels are all at most 32x32 take the one-wave 32x32
// kernel instead.  Every shape computes the same blocks.
using LevelKernel = void (*)(const LevelJob *, uint32_t, uint32_t, const uint32_t *, uint32_t *);
struct DwtShape {
    const char *name;
    LevelKernel kernel[2];  // [irreversible, reversible]
    uint32_t threads;
    bool small;
};
const DwtShape kDwtShapes[] = {
    {"64x64-w4+32x32-w1", {sgn_idwt_level_irr, sgn_idwt_level_rev}, kIdwtThreads, true},
    {"64x64-w4", {sgn_idwt_level_irr, sgn_idwt_level_rev}, kIdwtThreads, false},
    {"64x64-w2+32x32-w1", {sgn_idwt_level_irr_w2, sgn_idwt_level_rev_w2}, Mid::kThreads, true},
    {"64x64-w2", {sgn_idwt_level_irr_w2, sgn_idwt_level_rev_w2}, Mid::kThreads, false},
};
constexpr int kDwtShapeCount = int(sizeof(kDwtShapes) / sizeof(kDwtShapes[0]));

// Classic code-block kernels by block width (32, 64, 256, 1024) and active
// lanes per wave (64, 32, 16; the 1024-wide shape always takes 16).
using ClassicKernel = void (*)(const ClassicJob *, uint32_t, const uint8_t *, const ClassicSeg *,
                               int32_t *, uint32_t *);
constexpr int kClassicLaneOpts = 3;
constexpr uint32_t kClassicLaneCount[kClassicLaneOpts] = {64, 32, 16};
const ClassicKernel kClassicKernels[4][kClassicLaneOpts] = {
    {sgn_t1_classic<ClassicW32<64>>, sgn_t1_classic<ClassicW32<32>>,
     sgn_t1_classic<ClassicW32<16>>},
    {sgn_t1_classic<ClassicW64<64>>, sgn_t1_classic<ClassicW64<32>>,
     sgn_t1_classic<ClassicW64<16>>},
    {sgn_t1_classic<ClassicW256<64>>, sgn_t1_classic<ClassicW256<32>>,
     sgn_t1_classic<ClassicW256<16>>},
    {sgn_t1_classic<ClassicW1024>, sgn_t1_classic<ClassicW1024>, sgn_t1_classic<ClassicW1024>}};
int classic_shape(uint32_t w) {
    return w <= uint32_t(ClassicW32<64>::kMaxW) ? 0
           : w <= uint32_t(ClassicW64<64>::kMaxW) ? 1
           : w <= uint32_t(ClassicW256<64>::kMaxW) ? 2
                                                    : 3;
}
uint32_t classic_lanes(int v, int opt) {
    return v == 3 ? uint32_t(ClassicW1024::kLanes) : kClassicLaneCount[opt];
}
// Lane-list size up to which plans start with the cooperative kernel.  On
// 256x256 RGB tiles (75 code-blocks each) it took 0.28 against 1.41 ms for
// 16 tiles and 1.21 against 1.61 ms for 256 tiles (19,200 blocks); the lane
// kernel wins once its waves fill the device.
constexpr size_t kCoopMaxBlocks = 24576;

}  // namespace

namespace {

constexpr uint32_t kLaneMaxDim = 64;

// Blocks of the lane-list kernels: at most 64x64, cleanup with or without
// SigProp / MagRef passes.
bool lane_eligible(const CbJob &j) {
    return j.w <= kLaneMaxDim && j.h <= kLaneMaxDim && ((j.flags & kCbZero) || j.passes <= 3);
}

bool lane_refines(const CbJob &j) { return !(j.flags & kCbZero) && j.passes > 1; }

struct LevelRange {
    uint8_t rev, depth;
    uint32_t start, count;
};

template <typename T>
Status dev_upload(T *&dptr, const std::vector<T> &v) {
    dptr = nullptr;
    if (v.empty()) return kOk;
    if (hipMalloc(&dptr, v.size() * sizeof(T)) != hipSuccess) {
        set_last_error("hipMalloc failed for plan tables");
        return kErrDevice;
    }
    if (hipMemcpy(dptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        set_last_error("plan table upload failed");
        return kErrDevice;
    }
    return kOk;
}

}  // namespace

struct sgn_cost_model {
    int device = 0;
    float *d_params = nullptr;             // kMlpParams
    float *d_m = nullptr, *d_v = nullptr;  // Adam moments
    uint32_t steps = 0;
    ~sgn_cost_model() {
        void *ptrs[] = {d_params, d_m, d_v};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
    }
};

// Device blocks kept for reuse instead of freed: a serving loop allocates
// and releases plan and output blocks continuously, and hipFree waits for
// the whole device, so a freed block would stall every batch in flight.
// get() hands out a cached block of at least `bytes` (at most twice that,
// plus 1 MiB) or allocates one; put() caches it, freeing the largest cached
// blocks while more than `cap` bytes are cached.  Not thread-safe: one
// owner (the server) drives it.
class DeviceBlocks {
  public:
    explicit DeviceBlocks(uint64_t cap) : cap_(cap) {}
    ~DeviceBlocks() { trim(0); }
    DeviceBlocks(const DeviceBlocks &) = delete;
    DeviceBlocks &operator=(const DeviceBlocks &) = delete;
    hipError_t get(uint64_t bytes, void **out) {
        bytes = std::max<uint64_t>(bytes, 256);
        auto it = free_.lower_bound(bytes);
        if (it != free_.end() && it->first <= 2 * bytes + (uint64_t(1) << 20)) {
            *out = it->second;
            size_[*out] = it->first;
            cached_ -= it->first;
            free_.erase(it);
            return hipSuccess;
        }
        const hipError_t e = hipMalloc(out, bytes);
        if (e == hipSuccess) size_[*out] = bytes;
        return e;
    }
    void put(void *p) {
        if (!p) return;
        const auto it = size_.find(p);
        if (it == size_.end()) return;
        free_.emplace(it->second, p);
        cached_ += it->second;
        size_.erase(it);
        trim(cap_);
    }

  private:
    void trim(uint64_t keep) {
        while (cached_ > keep && !free_.empty()) {
            auto last = std::prev(free_.end());
            (void)hipFree(last->second);
            cached_ -= last->first;
            free_.erase(last);
        }
    }
    uint64_t cap_, cached_ = 0;
    std::multimap<uint64_t, void *> free_;         // cached: size -> block
    std::unordered_map<void *, uint64_t> size_;   // handed out: block -> size
};

struct sgn_j2k_plan {
    size_t ntiles = 0;
    // host tables
    std::vector<CbJob> cb_lane, cb_lane_plain, cb_wave;
    // lane-list blocks with SigProp / MagRef passes, costliest first: the
    // refinement kernel runs them after a lane-kernel variant (the
    // cooperative kernel runs them itself)
    std::vector<CbJob> cb_refine;
    std::vector<uint32_t> lane_plain_begin, wave_begin;  // per tile, size ntiles+1
    std::vector<LevelJob> lv[2];                          // [rev] sorted by (depth, tile)
    std::vector<uint32_t> lv_group[2];                    // start index of each depth, size D+1
    std::vector<std::vector<LevelRange>> tile_levels;     // per tile
    std::vector<LevelJob> final_lv;
    std::vector<FinalJob> fin[2];
    // work-group (64x64 output block) -> index of its job in lv[rev] / fin[rev]
    std::vector<uint32_t> lv_block_job[2], fin_block_job[2];
    std::vector<std::pair<uint8_t, uint32_t>> tile_final_first;  // (rev, first index)
    std::vector<uint32_t> tile_final_count;
    std::vector<uint32_t> tile_width, tile_height;
    std::vector<size_t> tile_row_bytes;
    uint32_t depth_max = 0;
    uint64_t coef_elems = 0, arena_len = 0, out_bytes = 0, out_pixels = 0;
    uint64_t coded_bytes = 0;  // sum of the codestream lengths
    uint64_t pruned_cbs = 0;   // code-blocks region decodes skip
    // Codestream bytes reach the device arena as runs: tiles whose bytes
    // follow each other in host memory (gaps under 4 KB) form one run and
    // keep their relative offsets, so a batch laid out in one pinned buffer
    // goes up in one DMA straight from the caller's memory.
    struct CopyRun {
        const uint8_t *src;
        uint64_t arena_off, len;
    };
    std::vector<CopyRun> runs;
    // host time of the last build (ms): per-tile parse, merge of the tile
    // plans into the batch tables, staging fill and copy submission
    double t_parse = 0, t_merge = 0, t_stage = 0, t_copy = 0;
    // Code-block samples stored as int16 (every band has at most 15
    // magnitude bit-planes, so every sample fits); intermediate level
    // outputs stay int32.  Each band keeps its int32-sized slot.
    bool narrow = false;
    bool lane_refined = false;  // a lane-list block has SigProp / MagRef passes
    // device
    uint8_t *d_arena = nullptr;
    uint32_t *d_coef = nullptr;
    CbJob *d_cb_lane = nullptr, *d_cb_lane_plain = nullptr, *d_cb_wave = nullptr;
    CbJob *d_cb_refine = nullptr;
    LevelJob *d_lv[2] = {nullptr, nullptr};
    LevelJob *d_final_lv = nullptr;
    FinalJob *d_fin[2] = {nullptr, nullptr};
    uint32_t *d_lv_block_job[2] = {nullptr, nullptr}, *d_fin_block_job[2] = {nullptr, nullptr};
    uint32_t *d_fin_redo = nullptr;  // 5/3 final blocks left to the int32-plane pass
    uint8_t *d_status_lane = nullptr, *d_status_wave = nullptr;
    uint32_t *d_queue = nullptr;  // work counter of the persistent lane kernel
    // packet headers parsed on the device (tiles planned with device
    // packets): tables, tag-tree and length scratch, per-tile status; the
    // lane list then keeps the per-tile order (its jobs are filled on the
    // device) and every lane block goes through the refinement kernel
    bool device_packets = false;
    std::vector<PktTile> pk_tiles;
    std::vector<PktPart> pk_parts;
    std::vector<PktPacket> pk_pkts;
    std::vector<PktBand> pk_bands;
    std::vector<PktCb> pk_cbs;
    uint64_t pk_nodes = 0;
    PktTile *d_pk_tiles = nullptr;
    PktPart *d_pk_parts = nullptr;
    PktPacket *d_pk_pkts = nullptr;
    PktBand *d_pk_bands = nullptr;
    PktCb *d_pk_cbs = nullptr;
    uint32_t *d_pk_nodes = nullptr, *d_pk_lens = nullptr;
    uint8_t *d_pk_status = nullptr;
    // classic (EBCOT) code-blocks, by kernel shape (small, big, wide:
    // [classic_begin[v], classic_begin[v + 1])), costliest first; their
    // segment table; bytes of blocks joined from several packets
    std::vector<ClassicJob> classic;
    uint32_t classic_begin[5] = {0, 0, 0, 0, 0};
    uint32_t classic_grid[4] = {0, 0, 0, 0};  // work-groups launched per shape
    int classic_opt[4] = {0, 0, 0, 0};        // lane option per shape (kClassicLaneCount)
    uint32_t *d_crows = nullptr;              // the lanes' row stores
    std::vector<ClassicSeg> cseg;
    std::vector<uint8_t> joined;
    ClassicJob *d_classic = nullptr;
    ClassicSeg *d_cseg = nullptr;
    // output jobs of tiles whose components are on different sampling
    // grids, their 64x64-block table, and each tile's range
    std::vector<StoreJob> store;
    std::vector<uint32_t> store_block_job;
    std::vector<std::pair<uint32_t, uint32_t>> tile_store;  // (first, count)
    StoreJob *d_store = nullptr;
    uint32_t *d_store_block_job = nullptr;
    int lane_variant = kDefaultLaneVariant;  // index into kRowsVariants
    int dwt_shape = 0;                       // index into kDwtShapes
    hipGraphExec_t graph_exec = nullptr;
    hipGraph_t graph = nullptr;
    uint32_t launches_batched = 0;

    void *d_block = nullptr;  // the one allocation every device pointer above lives in
    uint64_t device_bytes = 0;  // its size
    DeviceBlocks *blocks = nullptr;  // where d_block came from (else hipMalloc)
    bool borrowed = false;           // d_block is the caller's memory: never freed here

    ~sgn_j2k_plan() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (d_block && !borrowed) {
            if (blocks)
                blocks->put(d_block);
            else
                (void)hipFree(d_block);
        }
    }
};

namespace {

// Host worker threads shared by every plan build (created once, on first
// use): parse work of many small plans does not pay a thread start each.
class WorkerPool {
  public:
    static WorkerPool &get() {
        static WorkerPool pool;
        return pool;
    }
    // Runs fn(i) for i in [0, n) on the pool and the calling thread.
    void run(size_t n, const std::function<void(size_t)> &fn) {
        if (n == 0) return;
        if (n == 1 || threads_.empty()) {
            for (size_t i = 0; i < n; ++i) fn(i);
            return;
        }
        std::lock_guard<std::mutex> one(run_mu_);  // one batch of work at a time
        {
            std::lock_guard<std::mutex> g(mu_);
            fn_ = &fn;
            n_ = n;
            next_.store(0);
            busy_ = threads_.size();
            ++gen_;
        }
        cv_.notify_all();
        drain();
        std::unique_lock<std::mutex> g(mu_);
        done_cv_.wait(g, [&] { return busy_ == 0; });
        fn_ = nullptr;
    }

  private:
    WorkerPool() {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const unsigned k = std::min(16u, hw) - 1;
        for (unsigned t = 0; t < k; ++t) threads_.emplace_back([this] { loop(); });
    }
    ~WorkerPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
    }
    void drain() {
        for (size_t i = next_.fetch_add(1); i < n_; i = next_.fetch_add(1)) (*fn_)(i);
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
            }
            drain();
            std::lock_guard<std::mutex> g(mu_);
            if (--busy_ == 0) done_cv_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex mu_, run_mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0, busy_ = 0;
    std::atomic<size_t> next_{0};
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// out = in ordered by cost, largest first, ties in input order (the order
// of a stable sort): a counting sort over the bounded cost range (a
// code-block's cost is its coded bytes plus a quarter of its samples).
void sort_by_cost(const std::vector<CbJob> &in, std::vector<CbJob> &out) {
    uint32_t top = 0;
    for (const CbJob &j : in) top = std::max(top, j.cost);
    if (top > (1u << 20)) {
        out = in;
        std::stable_sort(out.begin(), out.end(),
                         [](const CbJob &a, const CbJob &b) { return a.cost > b.cost; });
        return;
    }
    std::vector<uint32_t> at(size_t(top) + 2, 0);
    for (const CbJob &j : in) at[top - j.cost + 1]++;
    for (size_t k = 1; k < at.size(); ++k) at[k] += at[k - 1];
    out.resize(in.size());
    for (const CbJob &j : in) out[at[top - j.cost]++] = j;
}

// Tiles parsed into their local plans: the parallel half of a plan build.
struct ParsedBatch {
    size_t n = 0;
    std::vector<LocalPlan> lps;
    // each request's bytes narrowed to its codestream (a JP2 file's jp2c
    // payload); only the codestream travels to the device
    std::vector<sgn_tile_req> cs;
    // per tile: code-blocks for the lane list, the wave list and refinement
    std::vector<uint32_t> n_lane, n_wave, n_ref;
    double t_parse = 0;
    bool device_packets = false;
    Status st = kOk;
    std::string msg;  // error of the first failing tile
};

// Parses every tile on the worker pool.  Failures are reported in B.st /
// B.msg (and per tile in tile_status), not in the calling thread's error.
void parse_batch(const sgn_tile_req *reqs, size_t n, ParsedBatch &B, int32_t *tile_status,
                 bool device_packets = false) {
    const auto t0 = std::chrono::steady_clock::now();
    B.n = n;
    B.lps.assign(n, LocalPlan());
    B.cs.assign(reqs, reqs + n);
    B.n_lane.assign(n, 0);
    B.n_wave.assign(n, 0);
    B.n_ref.assign(n, 0);
    B.st = kOk;
    B.msg.clear();
    B.device_packets = device_packets;
    std::vector<int32_t> st(n, 0);
    std::vector<std::string> errs(n);
    std::vector<LocalPlan> &lps = B.lps;
    std::vector<sgn_tile_req> &cs = B.cs;
    const sgn_tile_req *jobs = cs.data();
    WorkerPool::get().run(n, [&](size_t i) {
        Container c;
        st[i] = locate_codestream(reqs[i].bytes, reqs[i].len, c);
        if (!st[i]) st[i] = device_decodable(c);
        if (!st[i]) {
            cs[i].bytes = c.cs;
            cs[i].len = c.cs_len;
            Region roi;
            roi.on = jobs[i].has_roi != 0;
            roi.x = jobs[i].roi_x;
            roi.y = jobs[i].roi_y;
            roi.w = jobs[i].roi_w;
            roi.h = jobs[i].roi_h;
            st[i] = build_local_plan(jobs[i].bytes, jobs[i].len, jobs[i].fmt, jobs[i].denominator,
                                     reinterpret_cast<uint64_t>(jobs[i].dst), jobs[i].pitch,
                                     jobs[i].capacity, lps[i], roi, device_packets);
        }
        if (st[i]) {
            errs[i] = last_error();
            return;
        }
        for (const CbJob &j : lps[i].cbs) {
            if (lane_eligible(j)) {
                ++B.n_lane[i];
                B.n_ref[i] += lane_refines(j) ? 1u : 0u;
            } else {
                ++B.n_wave[i];
            }
        }
    });
    for (size_t i = 0; i < n; ++i) {
        if (tile_status) tile_status[i] = st[i];
        if (st[i] && !B.st) {
            B.st = static_cast<Status>(st[i]);
            B.msg = "tile " + std::to_string(i) + ": " + errs[i];
        }
    }
    B.t_parse = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                    .count();
}

// Merges the parsed tiles into the batch tables.  pool: fill the code-block
// lists on the worker pool (false when another thread is parsing on it).
Status merge_batch(ParsedBatch &B, sgn_j2k_plan &P, bool pool) {
    const size_t n = B.n;
    std::vector<LocalPlan> &lps = B.lps;
    const sgn_tile_req *jobs = B.cs.data();
    const std::vector<uint32_t> &n_lane = B.n_lane, &n_wave = B.n_wave, &n_ref = B.n_ref;
    P.ntiles = n;
    P.t_parse = B.t_parse;
    const auto t1 = std::chrono::steady_clock::now();
    // ---- merge with rebasing ----
    uint64_t arena = 0, coef = 0;
    std::vector<uint64_t> arena_base(n), coef_base(n);
    P.runs.clear();
    P.coded_bytes = 0;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t *b = jobs[i].bytes;
        const uint64_t len = jobs[i].len;
        P.coded_bytes += len;
        bool joined = false;
        if (!P.runs.empty()) {
            sgn_j2k_plan::CopyRun &r = P.runs.back();
            const uint8_t *end = r.src + r.len;
            if (b >= end && uint64_t(b - end) < 4096) {
                arena_base[i] = r.arena_off + uint64_t(b - r.src);
                r.len = uint64_t(b - r.src) + len;
                arena = r.arena_off + r.len;
                joined = true;
            }
        }
        if (!joined) {
            arena = (arena + 15) & ~uint64_t(15);
            arena_base[i] = arena;
            P.runs.push_back({b, arena, len});
            arena += len;
        }
        coef_base[i] = coef;
        coef += lps[i].coef_elems;
    }
    // classic code-blocks whose contributions were joined on the host: one
    // more run after the codestreams
    std::vector<uint64_t> joined_base(n, 0);
    {
        uint64_t total = 0;
        for (const LocalPlan &lp : lps) total += lp.joined.size();
        P.joined.clear();
        if (total) {
            P.joined.reserve(total);
            arena = (arena + 15) & ~uint64_t(15);
            for (size_t i = 0; i < n; ++i) {
                joined_base[i] = arena + P.joined.size();
                P.joined.insert(P.joined.end(), lps[i].joined.begin(), lps[i].joined.end());
            }
            P.runs.push_back({P.joined.data(), arena, total});
            arena += total;
        }
    }
    // the HT kernels read whole aligned 16-byte units and the classic kernel
    // a window of three dwords: pad the arena so that both stay inside the
    // allocation
    arena = (arena + 31) & ~uint64_t(15);
    if (arena >= (1ull << 32) || coef >= (1ull << 32)) {
        set_last_error("batch too large for 32-bit plan offsets; split it");
        return kErrArg;
    }
    P.arena_len = arena;
    P.coef_elems = coef;
    {
        uint32_t kmax = 0;
        for (const LocalPlan &lp : lps) kmax = std::max(kmax, lp.max_k);
        const char *env = std::getenv("SGN_COEF_NARROW");
        const bool allow = !(env && env[0] == '0');
        bool classic = false;
        for (const LocalPlan &lp : lps) classic |= !lp.classic.empty();
        // classic blocks OR magnitude bits into int32 samples
        P.narrow = allow && !classic && kmax <= 15 && 2 * coef < (1ull << 32);
    }
    // code-block jobs: each tile's slice of the three lists from prefix sums
    // of its counts, then every tile's jobs rebased into place in parallel
    P.lane_plain_begin.assign(n + 1, 0);
    P.wave_begin.assign(n + 1, 0);
    std::vector<uint32_t> ref_begin(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
        P.lane_plain_begin[i + 1] = P.lane_plain_begin[i] + n_lane[i];
        P.wave_begin[i + 1] = P.wave_begin[i] + n_wave[i];
        ref_begin[i + 1] = ref_begin[i] + n_ref[i];
    }
    P.cb_lane_plain.resize(P.lane_plain_begin[n]);
    P.cb_wave.resize(P.wave_begin[n]);
    P.cb_refine.resize(ref_begin[n]);
    // device packet tables: each tile's slice, rebased with its jobs below
    struct PkBase {
        uint32_t tile, part, pkt, band, cb;
        uint64_t node;
    };
    std::vector<PkBase> pkb(n + 1, PkBase{0, 0, 0, 0, 0, 0});
    for (size_t i = 0; i < n; ++i) {
        const LocalPlan &lp = lps[i];
        pkb[i + 1] = {pkb[i].tile + uint32_t(lp.pk_tiles.size()),
                      pkb[i].part + uint32_t(lp.pk_parts.size()),
                      pkb[i].pkt + uint32_t(lp.pk_pkts.size()),
                      pkb[i].band + uint32_t(lp.pk_bands.size()),
                      pkb[i].cb + uint32_t(lp.pk_cbs.size()), pkb[i].node + lp.pk_nodes};
    }
    P.device_packets = B.device_packets && pkb[n].tile > 0;
    if (pkb[n].node * 2 >= (uint64_t(1) << 32)) {
        set_last_error("batch too large for the device packet tables; split it");
        return kErrArg;
    }
    P.pk_tiles.resize(pkb[n].tile);
    P.pk_parts.resize(pkb[n].part);
    P.pk_pkts.resize(pkb[n].pkt);
    P.pk_bands.resize(pkb[n].band);
    P.pk_cbs.resize(pkb[n].cb);
    P.pk_nodes = pkb[n].node;
    auto fill = [&](size_t i) {
        const LocalPlan &lp = lps[i];
        const uint32_t cbase = static_cast<uint32_t>(coef_base[i]);
        uint32_t a = P.lane_plain_begin[i], w = P.wave_begin[i], r = ref_begin[i];
        // list position of each local job (device packet tables)
        std::vector<uint32_t> pos(lp.pk_cbs.empty() ? 0 : lp.cbs.size());
        for (size_t k = 0; k < lp.cbs.size(); ++k) {
            if (!pos.empty())
                pos[k] = lane_eligible(lp.cbs[k]) ? a : (kPktWave | w);
            CbJob j = lp.cbs[k];
            j.src += static_cast<uint32_t>(arena_base[i]);
            // int16 index of a narrow sample: twice its band's int32 offset
            // plus the offset inside the band
            j.dst = P.narrow ? j.dst + lp.cb_band[k] + 2 * cbase : j.dst + cbase;
            j.tile = static_cast<uint32_t>(i);
            if (lane_eligible(j)) {
                P.cb_lane_plain[a++] = j;
                if (lane_refines(j)) P.cb_refine[r++] = j;
            } else {
                P.cb_wave[w++] = j;
            }
        }
        const PkBase &pb = pkb[i];
        for (size_t k = 0; k < lp.pk_tiles.size(); ++k) {
            PktTile t = lp.pk_tiles[k];
            t.part0 += pb.part;
            t.pkt0 += pb.pkt;
            t.cb0 += pb.cb;
            t.node0 += static_cast<uint32_t>(pb.node);
            P.pk_tiles[pb.tile + k] = t;
        }
        for (size_t k = 0; k < lp.pk_parts.size(); ++k) {
            PktPart q = lp.pk_parts[k];
            q.off += static_cast<uint32_t>(arena_base[i]);
            P.pk_parts[pb.part + k] = q;
        }
        for (size_t k = 0; k < lp.pk_pkts.size(); ++k) {
            PktPacket q = lp.pk_pkts[k];
            q.band0 += pb.band;
            P.pk_pkts[pb.pkt + k] = q;
        }
        for (size_t k = 0; k < lp.pk_bands.size(); ++k) {
            PktBand q = lp.pk_bands[k];
            q.cb0 += pb.cb;
            q.incl += static_cast<uint32_t>(pb.node);
            q.zbp += static_cast<uint32_t>(pb.node);
            P.pk_bands[pb.band + k] = q;
        }
        for (size_t k = 0; k < lp.pk_cbs.size(); ++k) {
            PktCb q = lp.pk_cbs[k];
            if (q.job != kPktNoJob) q.job = q.lane_pos = pos[q.job];
            P.pk_cbs[pb.cb + k] = q;
        }
    };
    if (pool)
        WorkerPool::get().run(n, fill);
    else
        for (size_t i = 0; i < n; ++i) fill(i);
    P.tile_levels.assign(n, {});
    P.tile_final_first.assign(n, {0, 0});
    P.tile_final_count.assign(n, 0);
    P.tile_width.resize(n);
    P.tile_height.resize(n);
    P.tile_row_bytes.resize(n);
    for (size_t i = 0; i < n; ++i) {
        LocalPlan &lp = lps[i];
        P.tile_width[i] = lp.width;
        P.tile_height[i] = lp.height;
        P.tile_row_bytes[i] = lp.row_bytes;
        P.out_bytes += lp.out_bytes;
        P.out_pixels += uint64_t(lp.width) * lp.height;
        P.pruned_cbs += lp.pruned_cbs;
        const uint32_t cbase = static_cast<uint32_t>(coef_base[i]);
        P.tile_store.push_back({static_cast<uint32_t>(P.store.size()),
                                static_cast<uint32_t>(lp.stores.size())});
        for (StoreJob sj : lp.stores) {
            sj.plane += cbase;
            sj.block_base = static_cast<uint32_t>(P.store_block_job.size());
            for (uint32_t b = 0; b < uint32_t(sj.nbx) * sj.nby; ++b)
                P.store_block_job.push_back(static_cast<uint32_t>(P.store.size()));
            P.store.push_back(sj);
        }
        const uint32_t seg_base = static_cast<uint32_t>(P.cseg.size());
        P.cseg.insert(P.cseg.end(), lp.cseg.begin(), lp.cseg.end());
        for (ClassicJob j : lp.classic) {
            j.src += static_cast<uint32_t>((j.style & kClassicJoined) ? joined_base[i] : arena_base[i]);
            j.dst += cbase;
            j.seg += seg_base;
            j.tile = static_cast<uint32_t>(i);
            P.classic.push_back(j);
        }
        auto rebase = [&](LevelJob &l) {
            if (!P.narrow) l.narrow = 0;
            for (int b = 0; b < 4; ++b)
                l.band[b] = ((l.narrow >> b) & 1) ? 2 * (l.band[b] + cbase) : l.band[b] + cbase;
        };
        for (LevelJob &l : lp.levels) {
            rebase(l);
            l.out += cbase;
        }
        for (LevelJob &l : lp.finals_lv) rebase(l);
        for (size_t d = 0; d < lp.level_depth.size(); ++d)
            P.depth_max = std::max<uint32_t>(P.depth_max, lp.level_depth[d]);
    }
    // code-blocks for the batched lane kernel, ordered by predicted cost so
    // that a wave's lanes finish together
    if (P.device_packets) {
        // lengths are known on the device only: the batched lane list is
        // ordered by block shape (largest first, per-tile order inside a
        // shape), and each packet-table entry learns its block's place
        const size_t nl = P.cb_lane_plain.size();
        std::vector<uint32_t> idx(nl), at(nl);
        {
            // counting sort by area (lane blocks are at most 64 x 64)
            constexpr uint32_t kTop = kLaneMaxDim * kLaneMaxDim;
            std::vector<uint32_t> start(kTop + 2, 0);
            auto area = [&](size_t k) {
                return uint32_t(P.cb_lane_plain[k].w) * P.cb_lane_plain[k].h;
            };
            for (size_t k = 0; k < nl; ++k) start[kTop - area(k) + 1]++;
            for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
            for (size_t k = 0; k < nl; ++k) idx[start[kTop - area(k)]++] = static_cast<uint32_t>(k);
        }
        P.cb_lane.resize(idx.size());
        for (size_t k = 0; k < idx.size(); ++k) {
            P.cb_lane[k] = P.cb_lane_plain[idx[k]];
            at[idx[k]] = static_cast<uint32_t>(k);
        }
        for (PktCb &q : P.pk_cbs)
            if (q.job != kPktNoJob && !(q.job & kPktWave)) q.lane_pos = at[q.job];
        P.cb_refine.clear();
        P.lane_refined = true;
    } else {
        sort_by_cost(P.cb_lane_plain, P.cb_lane);
        P.lane_refined = !P.cb_refine.empty();
    }
    std::stable_sort(P.cb_refine.begin(), P.cb_refine.end(), [](const CbJob &a, const CbJob &b) {
        return uint32_t(a.lref) + a.lcup > uint32_t(b.lref) + b.lcup;
    });
    // classic blocks: by kernel shape, costliest first within each
    {
        auto shape = [](const ClassicJob &j) { return classic_shape(j.w); };
        std::stable_sort(P.classic.begin(), P.classic.end(),
                         [&](const ClassicJob &a, const ClassicJob &b) {
                             const int sa = shape(a), sb = shape(b);
                             return sa != sb ? sa < sb : a.cost > b.cost;
                         });
        P.classic_begin[0] = 0;
        for (int v = 0; v < 4; ++v) {
            uint32_t e = P.classic_begin[v];
            while (e < P.classic.size() && shape(P.classic[e]) == v) ++e;
            P.classic_begin[v + 1] = e;
        }
    }
    // kernel of a plan nobody autotuned: below kCoopMaxBlocks code-blocks the
    // lane kernel cannot fill the device (each wave takes 64 blocks) and its
    // time is one block's serial decode, so one block per wave finishes first
    P.lane_variant = P.cb_lane.size() <= kCoopMaxBlocks ? kCoopVariant : kDefaultLaneVariant;
    // synthesis levels grouped by (transform, depth), tiles in order: the
    // levels of each depth counted, then placed tile by tile
    const uint32_t D = P.depth_max;
    for (int rev = 0; rev < 2; ++rev) {
        std::vector<uint32_t> start(D + 2, 0);  // start[d]: first index of depth d
        for (const LocalPlan &lp : lps)
            for (size_t k = 0; k < lp.levels.size(); ++k)
                if (lp.levels[k].reversible == rev) start[lp.level_depth[k] + 1]++;
        for (uint32_t d = 1; d <= D + 1; ++d) start[d] += start[d - 1];
        P.lv_group[rev].assign(start.begin() + 1, start.end());
        P.lv[rev].resize(start[D + 1]);
        std::vector<uint32_t> pos(start), first(D + 1, 0), count(D + 1, 0);
        for (size_t i = 0; i < n; ++i) {
            const LocalPlan &lp = lps[i];
            for (size_t k = 0; k < lp.levels.size(); ++k) {
                if (lp.levels[k].reversible != rev) continue;
                const uint32_t d = lp.level_depth[k];
                if (!count[d]) first[d] = pos[d];
                P.lv[rev][pos[d]++] = lp.levels[k];
                count[d]++;
            }
            for (uint32_t d = 1; d <= D; ++d) {
                if (!count[d]) continue;
                P.tile_levels[i].push_back(
                    {static_cast<uint8_t>(rev), static_cast<uint8_t>(d), first[d], count[d]});
                count[d] = 0;
            }
        }
        uint32_t blocks = 0;  // work-groups in launch order
        for (LevelJob &l : P.lv[rev]) {
            l.block_base = blocks;
            blocks += uint32_t(l.nbx) * l.nby;
        }
    }
    // finals
    uint32_t fblocks[2] = {0, 0};
    for (size_t i = 0; i < n; ++i) {
        LocalPlan &lp = lps[i];
        const uint32_t lv_base = static_cast<uint32_t>(P.final_lv.size());
        P.final_lv.insert(P.final_lv.end(), lp.finals_lv.begin(), lp.finals_lv.end());
        for (size_t f = 0; f < lp.finals.size(); ++f) {
            FinalJob fj = lp.finals[f];
            for (int c = 0; c < 4; ++c) fj.level[c] += lv_base;
            const int rev = fj.reversible ? 1 : 0;
            for (int c = 0; c < fj.ncomp && c < 3; ++c)
                if (P.final_lv[fj.level[c]].reversible != rev) {
                    set_last_error("tile " + std::to_string(i) +
                                   ": components with different wavelet transforms");
                    return kErrUnsupported;
                }
            fj.block_base = fblocks[rev];
            fblocks[rev] += uint32_t(fj.nbx) * fj.nby;
            if (f == 0) P.tile_final_first[i] = {static_cast<uint8_t>(rev),
                                                 static_cast<uint32_t>(P.fin[rev].size())};
            P.fin[rev].push_back(fj);
            P.tile_final_count[i]++;
        }
    }
    // block -> job tables: one load per work-group instead of a search
    for (int rev = 0; rev < 2; ++rev) {
        P.lv_block_job[rev].clear();
        for (size_t j = 0; j < P.lv[rev].size(); ++j)
            P.lv_block_job[rev].insert(P.lv_block_job[rev].end(),
                                       size_t(P.lv[rev][j].nbx) * P.lv[rev][j].nby, uint32_t(j));
        P.fin_block_job[rev].clear();
        for (size_t j = 0; j < P.fin[rev].size(); ++j)
            P.fin_block_job[rev].insert(P.fin_block_job[rev].end(),
                                        size_t(P.fin[rev][j].nbx) * P.fin[rev][j].nby, uint32_t(j));
    }
    P.t_merge = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1)
                    .count();
    return kOk;
}

// Packet headers of HT tiles parsed on the device (SGN_DEVICE_PACKETS=1).
bool device_packets_enabled() {
    const char *e = std::getenv("SGN_DEVICE_PACKETS");
    return e && e[0] == '1';
}

Status build_plan(const sgn_tile_req *reqs, size_t n, sgn_j2k_plan &P, int32_t *tile_status) {
    ParsedBatch B;
    parse_batch(reqs, n, B, tile_status, device_packets_enabled());
    if (B.st) {
        set_last_error(B.msg);
        return B.st;
    }
    return merge_batch(B, P, true);
}

// True when p lies in page-locked host memory (hipHostMalloc or
// registered), which the DMA engines read directly.  A query on pageable
// memory fails; its error is cleared so later launch checks do not see it.
bool host_pinned(const void *p) {
    hipPointerAttribute_t a{};
    const hipError_t e = hipPointerGetAttributes(&a, p);
    (void)hipGetLastError();
    return e == hipSuccess && a.type == hipMemoryTypeHost;
}

// Reserves `bytes` of the session's pinned staging slot k, after the copies
// out of its previous contents have finished.
Status slot_reserve(sgn_hip_session *s, int k, size_t bytes, uint8_t **out) {
    PinnedSlot &sl = s->slot[k];
    if (sl.pending) {
        HIP_TRY(hipEventSynchronize(sl.done));
        sl.pending = false;
    }
    if (!sl.done) HIP_TRY(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    if (sl.cap < bytes) {
        if (sl.p) HIP_TRY(hipHostFree(sl.p));
        sl.p = nullptr;
        sl.cap = 0;
        const size_t cap = (std::max<size_t>(bytes, size_t(16) << 20) + (size_t(2) << 20) - 1) &
                           ~((size_t(2) << 20) - 1);
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&sl.p), cap, hipHostMallocDefault));
        sl.cap = cap;
    }
    *out = sl.p;
    return kOk;
}

// One device allocation per plan: codestream arena, job tables, status
// words, work counter and coefficient arena, carved at 256-byte alignment.
// Codestream runs of at least 64 KB in pinned host memory are copied
// straight from the caller's memory; the other runs and the tables are
// staged in the session's pinned slot `slot` and go up in as few copies as
// the run layout allows.  Copies are enqueued on `copy`; the caller orders
// the decode after them (same stream or an event).  With sync the call
// returns after the copies completed.
// With `fixed` (fixed_cap bytes of device memory the caller owns) nothing is
// allocated: the plan is carved from it, every run and table is copied
// straight from host memory (no pinned staging), and the call returns after
// the copies completed.
uint32_t classic_resident(sgn_hip_session *s, int v, int opt);
uint32_t session_cus(sgn_hip_session *s);

Status upload_plan(sgn_hip_session *s, const sgn_tile_req *jobs, sgn_j2k_plan &P,
                   hipStream_t copy, int slot, bool sync, void *fixed = nullptr,
                   uint64_t fixed_cap = 0) {
    (void)jobs;
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = off;
        off = (off + bytes + 255) & ~uint64_t(255);
        return o;
    };
    auto bytes_of = [](const auto &v) { return uint64_t(v.size()) * sizeof(v[0]); };
    const uint64_t o_arena = take(P.arena_len);
    const uint64_t o_tables = off;
    const uint64_t o_lane = take(bytes_of(P.cb_lane)), o_plain = take(bytes_of(P.cb_lane_plain));
    const uint64_t o_wave = take(bytes_of(P.cb_wave)), o_ref = take(bytes_of(P.cb_refine));
    const uint64_t o_lv0 = take(bytes_of(P.lv[0])), o_lv1 = take(bytes_of(P.lv[1]));
    const uint64_t o_fin0 = take(bytes_of(P.fin[0])), o_fin1 = take(bytes_of(P.fin[1]));
    const uint64_t o_flv = take(bytes_of(P.final_lv));
    const uint64_t o_lbj0 = take(bytes_of(P.lv_block_job[0])), o_lbj1 = take(bytes_of(P.lv_block_job[1]));
    const uint64_t o_fbj0 = take(bytes_of(P.fin_block_job[0])), o_fbj1 = take(bytes_of(P.fin_block_job[1]));
    const uint64_t o_cls = take(bytes_of(P.classic)), o_cseg = take(bytes_of(P.cseg));
    const uint64_t o_sto = take(bytes_of(P.store)), o_sbj = take(bytes_of(P.store_block_job));
    const uint64_t o_pkt = take(bytes_of(P.pk_tiles)), o_pkp = take(bytes_of(P.pk_parts));
    const uint64_t o_pkk = take(bytes_of(P.pk_pkts)), o_pkb = take(bytes_of(P.pk_bands));
    const uint64_t o_pkc = take(bytes_of(P.pk_cbs));
    const uint64_t host_bytes = off;
    const uint64_t o_pkn = take(P.pk_nodes * 2 * sizeof(uint32_t));
    const uint64_t o_pkl = take(P.pk_cbs.size() * sizeof(uint32_t)), o_pks = take(P.pk_tiles.size());
    const uint64_t o_stl = take(P.cb_lane.size()), o_stw = take(P.cb_wave.size());
    const uint64_t o_ff = take(P.fin_block_job[1].size() * sizeof(uint32_t));
    uint64_t crows_bytes = 0;
    // active lanes per classic wave: the most that still give every SIMD two
    // waves, else the fewest (SGN_CLASSIC_LANES=64, 32 or 16 forces a count)
    const char *lanes_env = std::getenv("SGN_CLASSIC_LANES");
    const uint32_t forced = lanes_env ? uint32_t(std::atoi(lanes_env)) : 0u;
    for (int v = 0; v < 4; ++v) {
        const uint32_t n = P.classic_begin[v + 1] - P.classic_begin[v];
        int opt = kClassicLaneOpts - 1;
        for (int o = 0; o < kClassicLaneOpts; ++o) {
            const bool fits = forced ? kClassicLaneCount[o] == forced
                                     : uint64_t(n) >= uint64_t(kClassicLaneCount[o]) * 8u *
                                                          uint64_t(session_cus(s));
            if (fits) {
                opt = o;
                break;
            }
        }
        P.classic_opt[v] = opt;
        const uint32_t lanes = classic_lanes(v, opt);
        P.classic_grid[v] = n ? std::min((n + lanes - 1) / lanes, classic_resident(s, v, opt)) : 0;
        crows_bytes = std::max<uint64_t>(crows_bytes, uint64_t(P.classic_grid[v]) * lanes *
                                                          kClassicStateDwords * sizeof(uint32_t));
    }
    const uint64_t o_crows = take(crows_bytes);
    const uint64_t o_queue = take(16), o_coef = take(P.coef_elems * sizeof(uint32_t));
    P.device_bytes = off;
    if (fixed) {
        if (off > fixed_cap) {
            set_last_error("the batch needs " + std::to_string(off) +
                           " bytes of device memory; the no-allocation scratch holds " +
                           std::to_string(fixed_cap));
            return kErrArg;
        }
        P.d_block = fixed;
        P.borrowed = true;
        uint8_t *d = static_cast<uint8_t *>(fixed);
        for (const sgn_j2k_plan::CopyRun &run : P.runs)
            HIP_TRY(hipMemcpyAsync(d + o_arena + run.arena_off, run.src, run.len,
                                   hipMemcpyHostToDevice, copy));
        auto put_dev = [&](uint64_t o, const auto &v) -> hipError_t {
            if (v.empty()) return hipSuccess;
            return hipMemcpyAsync(d + o, v.data(), bytes_of(v), hipMemcpyHostToDevice, copy);
        };
        HIP_TRY(put_dev(o_lane, P.cb_lane));
        HIP_TRY(put_dev(o_plain, P.cb_lane_plain));
        HIP_TRY(put_dev(o_wave, P.cb_wave));
        HIP_TRY(put_dev(o_ref, P.cb_refine));
        HIP_TRY(put_dev(o_lv0, P.lv[0]));
        HIP_TRY(put_dev(o_lv1, P.lv[1]));
        HIP_TRY(put_dev(o_fin0, P.fin[0]));
        HIP_TRY(put_dev(o_fin1, P.fin[1]));
        HIP_TRY(put_dev(o_flv, P.final_lv));
        HIP_TRY(put_dev(o_lbj0, P.lv_block_job[0]));
        HIP_TRY(put_dev(o_lbj1, P.lv_block_job[1]));
        HIP_TRY(put_dev(o_fbj0, P.fin_block_job[0]));
        HIP_TRY(put_dev(o_fbj1, P.fin_block_job[1]));
        HIP_TRY(put_dev(o_cls, P.classic));
        HIP_TRY(put_dev(o_cseg, P.cseg));
        HIP_TRY(put_dev(o_sto, P.store));
        HIP_TRY(put_dev(o_sbj, P.store_block_job));
        HIP_TRY(put_dev(o_pkt, P.pk_tiles));
        HIP_TRY(put_dev(o_pkp, P.pk_parts));
        HIP_TRY(put_dev(o_pkk, P.pk_pkts));
        HIP_TRY(put_dev(o_pkb, P.pk_bands));
        HIP_TRY(put_dev(o_pkc, P.pk_cbs));
        HIP_TRY(hipStreamSynchronize(copy));
        P.t_stage = 0;
        P.t_copy = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                              t0).count();
    } else if ((P.blocks ? P.blocks->get(off, &P.d_block) : hipMalloc(&P.d_block, off)) !=
               hipSuccess) {
        P.d_block = nullptr;
        set_last_error("hipMalloc failed for the plan (" + std::to_string(off) + " bytes)");
        return kErrDevice;
    }
    if (!fixed) {
        // which runs go up straight from the caller's memory
        std::vector<uint8_t> direct(P.runs.size(), 0);
        for (size_t r = 0; r < P.runs.size(); ++r)
            direct[r] = P.runs[r].len >= (uint64_t(64) << 10) && host_pinned(P.runs[r].src);
        // the staging slot mirrors [0, host_bytes) of the device block; arena
        // bytes of direct runs are left unfilled there and never copied
        uint8_t *h = nullptr;
        {
            const Status st = slot_reserve(s, slot, host_bytes, &h);
            if (st) return st;
        }
        // staged codestream bytes, copied by the worker pool in 1 MB pieces
        struct Piece {
            const uint8_t *src;
            uint64_t off, len;
        };
        std::vector<Piece> pieces;
        constexpr uint64_t kPiece = uint64_t(1) << 20;
        for (size_t r = 0; r < P.runs.size(); ++r) {
            if (direct[r]) continue;
            const sgn_j2k_plan::CopyRun &run = P.runs[r];
            for (uint64_t q = 0; q < run.len; q += kPiece)
                pieces.push_back({run.src + q, o_arena + run.arena_off + q,
                                  std::min(kPiece, run.len - q)});
        }
        WorkerPool::get().run(pieces.size(), [&](size_t i) {
            std::memcpy(h + pieces[i].off, pieces[i].src, pieces[i].len);
        });
        auto put = [&](uint64_t o, const auto &v) {
            if (!v.empty()) std::memcpy(h + o, v.data(), bytes_of(v));
        };
        put(o_lane, P.cb_lane);
        put(o_plain, P.cb_lane_plain);
        put(o_wave, P.cb_wave);
        put(o_ref, P.cb_refine);
        put(o_lv0, P.lv[0]);
        put(o_lv1, P.lv[1]);
        put(o_fin0, P.fin[0]);
        put(o_fin1, P.fin[1]);
        put(o_flv, P.final_lv);
        put(o_lbj0, P.lv_block_job[0]);
        put(o_lbj1, P.lv_block_job[1]);
        put(o_fbj0, P.fin_block_job[0]);
        put(o_fbj1, P.fin_block_job[1]);
        put(o_cls, P.classic);
        put(o_cseg, P.cseg);
        put(o_sto, P.store);
        put(o_sbj, P.store_block_job);
        put(o_pkt, P.pk_tiles);
        put(o_pkp, P.pk_parts);
        put(o_pkk, P.pk_pkts);
        put(o_pkb, P.pk_bands);
        put(o_pkc, P.pk_cbs);
        const auto t1 = std::chrono::steady_clock::now();
        uint8_t *d = static_cast<uint8_t *>(P.d_block);
        // direct runs first, then the staged range: consecutive staged runs and
        // the tables in one copy each
        for (size_t r = 0; r < P.runs.size(); ++r)
            if (direct[r])
                HIP_TRY(hipMemcpyAsync(d + o_arena + P.runs[r].arena_off, P.runs[r].src,
                                       P.runs[r].len, hipMemcpyHostToDevice, copy));
        {
            uint64_t a = UINT64_MAX, e = 0;
            auto flush = [&]() -> hipError_t {
                if (a == UINT64_MAX) return hipSuccess;
                const hipError_t err =
                    hipMemcpyAsync(d + a, h + a, e - a, hipMemcpyHostToDevice, copy);
                a = UINT64_MAX;
                return err;
            };
            for (size_t r = 0; r < P.runs.size(); ++r) {
                if (direct[r]) {
                    HIP_TRY(flush());
                    continue;
                }
                if (a == UINT64_MAX) a = o_arena + P.runs[r].arena_off;
                e = o_arena + P.runs[r].arena_off + P.runs[r].len;
            }
            HIP_TRY(flush());
            if (host_bytes > o_tables)
                HIP_TRY(hipMemcpyAsync(d + o_tables, h + o_tables, host_bytes - o_tables,
                                       hipMemcpyHostToDevice, copy));
        }
        HIP_TRY(hipEventRecord(s->slot[slot].done, copy));
        s->slot[slot].pending = true;
        if (sync) {
            HIP_TRY(hipEventSynchronize(s->slot[slot].done));
            s->slot[slot].pending = false;
        }
        const auto t2 = std::chrono::steady_clock::now();
        P.t_stage = std::chrono::duration<double, std::milli>(t1 - t0).count();
        P.t_copy = std::chrono::duration<double, std::milli>(t2 - t1).count();
    }
    uint8_t *const d = static_cast<uint8_t *>(P.d_block);
    auto at = [&](uint64_t o, auto *&ptr, bool present) {
        ptr = present ? reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(d + o) : nullptr;
    };
    at(o_arena, P.d_arena, P.arena_len != 0);
    at(o_lane, P.d_cb_lane, !P.cb_lane.empty());
    at(o_plain, P.d_cb_lane_plain, !P.cb_lane_plain.empty());
    at(o_wave, P.d_cb_wave, !P.cb_wave.empty());
    at(o_ref, P.d_cb_refine, !P.cb_refine.empty());
    at(o_lv0, P.d_lv[0], !P.lv[0].empty());
    at(o_lv1, P.d_lv[1], !P.lv[1].empty());
    at(o_fin0, P.d_fin[0], !P.fin[0].empty());
    at(o_fin1, P.d_fin[1], !P.fin[1].empty());
    at(o_flv, P.d_final_lv, !P.final_lv.empty());
    at(o_lbj0, P.d_lv_block_job[0], !P.lv_block_job[0].empty());
    at(o_lbj1, P.d_lv_block_job[1], !P.lv_block_job[1].empty());
    at(o_fbj0, P.d_fin_block_job[0], !P.fin_block_job[0].empty());
    at(o_fbj1, P.d_fin_block_job[1], !P.fin_block_job[1].empty());
    at(o_cls, P.d_classic, !P.classic.empty());
    at(o_cseg, P.d_cseg, !P.cseg.empty());
    at(o_sto, P.d_store, !P.store.empty());
    at(o_sbj, P.d_store_block_job, !P.store_block_job.empty());
    at(o_pkt, P.d_pk_tiles, !P.pk_tiles.empty());
    at(o_pkp, P.d_pk_parts, !P.pk_parts.empty());
    at(o_pkk, P.d_pk_pkts, !P.pk_pkts.empty());
    at(o_pkb, P.d_pk_bands, !P.pk_bands.empty());
    at(o_pkc, P.d_pk_cbs, !P.pk_cbs.empty());
    at(o_pkn, P.d_pk_nodes, P.pk_nodes != 0);
    at(o_pkl, P.d_pk_lens, !P.pk_cbs.empty());
    at(o_pks, P.d_pk_status, !P.pk_tiles.empty());
    at(o_stl, P.d_status_lane, !P.cb_lane.empty());
    at(o_stw, P.d_status_wave, !P.cb_wave.empty());
    at(o_ff, P.d_fin_redo, !P.fin_block_job[1].empty());
    at(o_crows, P.d_crows, crows_bytes != 0);
    at(o_queue, P.d_queue, true);
    at(o_coef, P.d_coef, P.coef_elems != 0);
    return kOk;
}

uint32_t group_blocks(const std::vector<LevelJob> &v, uint32_t a, uint32_t b) {
    if (a >= b) return 0;
    return v[b - 1].block_base + uint32_t(v[b - 1].nbx) * v[b - 1].nby - v[a].block_base;
}

uint32_t final_blocks(const std::vector<FinalJob> &v, uint32_t a, uint32_t b) {
    if (a >= b) return 0;
    return v[b - 1].block_base + uint32_t(v[b - 1].nbx) * v[b - 1].nby - v[a].block_base;
}

// Enqueue helpers; each returns the number of kernel launches issued.
// Work-groups of variant v that the device holds at once.
uint32_t resident_groups(sgn_hip_session *s, int v, bool narrow) {
    if (s->resident.empty()) s->resident.assign(2 * kRowsVariantCount, 0);
    const int slot = 2 * v + (narrow ? 1 : 0);
    if (!s->resident[slot]) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, kRowsVariants[v].kernel[narrow ? 1 : 0], kRowsVariants[v].waves * 64,
                0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        if (!s->cus && (hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount,
                                              s->device) != hipSuccess ||
                        s->cus < 1))
            s->cus = 1;
        s->resident[slot] = uint32_t(per_cu) * uint32_t(s->cus);
    }
    return s->resident[slot];
}

uint32_t session_cus(sgn_hip_session *s) {
    if (!s->cus && (hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount,
                                          s->device) != hipSuccess ||
                    s->cus < 1))
        s->cus = 1;
    return uint32_t(s->cus);
}

uint32_t classic_resident(sgn_hip_session *s, int v, int opt) {
    if (!s->classic_resident[v][opt]) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void *>(kClassicKernels[v][opt]), 64, 0) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        s->classic_resident[v][opt] = uint32_t(per_cu) * session_cus(s);
    }
    return s->classic_resident[v][opt];
}

// Clears the plan's device counters (word 0: lane work counter, word 1:
// final redo count) with a kernel, never a memset node (graph capture).
uint32_t clear_counters(sgn_hip_session *s, sgn_j2k_plan &P, uint32_t words) {
    hipLaunchKernelGGL(sgn_fill_bytes, dim3(1), dim3(64), 0, s->stream,
                       reinterpret_cast<uint8_t *>(P.d_queue), size_t(4) * words, size_t(words), 0u);
    return 1;
}

// clear: reset the work counter first (false: the caller already did)
uint32_t launch_lane_variant(sgn_hip_session *s, sgn_j2k_plan &P, int v, const CbJob *d_jobs,
                             uint32_t count, uint8_t *status, bool clear = true) {
    if (!count) return 0;
    const RowsVariant &rv = kRowsVariants[v];
    const uint32_t per_group = rv.coop ? 1u : 64u * uint32_t(rv.waves);
    uint32_t grid = (count + per_group - 1) / per_group;
    uint32_t k = 1;
    if (rv.persist) {
        if (clear) k += clear_counters(s, P, 1);
        grid = std::min(grid, resident_groups(s, v, P.narrow));
    }
    rv.launch[P.narrow ? 1 : 0](s->stream, grid, d_jobs, count, P.d_arena, P.d_coef, s->d_tables,
                                status, P.d_queue, P.lane_refined);
    if (!rv.coop && d_jobs == P.d_cb_lane && (!P.cb_refine.empty() || P.device_packets)) {
        // refinement passes of the batch's lane blocks, after their cleanup
        // (with device packets the pass counts are known on the device only:
        // the kernel walks the whole lane list and skips the other blocks)
        const CbJob *rj = P.device_packets ? P.d_cb_lane : P.d_cb_refine;
        const uint32_t nr = static_cast<uint32_t>(P.device_packets ? P.cb_lane.size()
                                                                   : P.cb_refine.size());
        if (P.narrow)
            hipLaunchKernelGGL((sgn_ht_refine<int16_t>), dim3((nr + 255) / 256), dim3(256), 0,
                               s->stream, rj, nr, P.d_arena, reinterpret_cast<int16_t *>(P.d_coef));
        else
            hipLaunchKernelGGL((sgn_ht_refine<int32_t>), dim3((nr + 255) / 256), dim3(256), 0,
                               s->stream, rj, nr, P.d_arena, reinterpret_cast<int32_t *>(P.d_coef));
        ++k;
    }
    return k;
}

uint32_t launch_lane(sgn_hip_session *s, sgn_j2k_plan &P, const CbJob *d_jobs, uint32_t count,
                     uint8_t *status, bool clear = true) {
    return launch_lane_variant(s, P, P.lane_variant, d_jobs, count, status, clear);
}

uint32_t launch_wave(sgn_hip_session *s, sgn_j2k_plan &P, const CbJob *d_jobs, uint32_t count,
                     uint8_t *status) {
    if (!count) return 0;
    if (P.narrow)
        hipLaunchKernelGGL((sgn_ht_coop<1024, true, int16_t>), dim3(count), dim3(64), 0,
                           s->stream, d_jobs, count, P.d_arena,
                           reinterpret_cast<int16_t *>(P.d_coef), s->d_tables, status, 0u);
    else
        hipLaunchKernelGGL((sgn_ht_coop<1024, true, int32_t>), dim3(count), dim3(64), 0,
                           s->stream, d_jobs, count, P.d_arena,
                           reinterpret_cast<int32_t *>(P.d_coef), s->d_tables, status, 0u);
    return 1;
}

// Every classic code-block of the plan: one launch per kernel shape (one
// block per lane; the grid at most the resident work-groups, which then
// walk the block groups).
uint32_t launch_classic(sgn_hip_session *s, sgn_j2k_plan &P) {
    uint32_t k = 0;
    for (int v = 0; v < 4; ++v) {
        const uint32_t a = P.classic_begin[v], b = P.classic_begin[v + 1];
        if (a >= b || !P.classic_grid[v]) continue;
        const uint32_t n = b - a;
        const ClassicJob *jobs = P.d_classic + a;
        int32_t *coef = reinterpret_cast<int32_t *>(P.d_coef);
        hipLaunchKernelGGL(kClassicKernels[v][P.classic_opt[v]], dim3(P.classic_grid[v]), dim3(64),
                           0, s->stream,
                           jobs, n, P.d_arena, P.d_cseg, coef, P.d_crows);
        ++k;
    }
    return k;
}

// True when every level of [a, b) is at most 32x32 (one Small block).
bool small_levels(const std::vector<LevelJob> &v, uint32_t a, uint32_t b) {
    for (uint32_t j = a; j < b; ++j)
        if (v[j].x1 - v[j].x0 > uint32_t(Small::kBlk) || v[j].y1 - v[j].y0 > uint32_t(Small::kBlk))
            return false;
    return true;
}

uint32_t launch_levels(sgn_hip_session *s, sgn_j2k_plan &P, int rev, uint32_t a, uint32_t b) {
    const uint32_t blocks = group_blocks(P.lv[rev], a, b);
    if (!blocks) return 0;
    const uint32_t off = P.lv[rev][a].block_base;
    const DwtShape &g = kDwtShapes[P.dwt_shape];
    // levels of at most 32x32 take the one-wave kernel where the shape says
    // so (same block count: each is one block for either kernel)
    const bool small = g.small && small_levels(P.lv[rev], a, b);
    const auto kernel = small ? (rev ? sgn_idwt_level_rev_small : sgn_idwt_level_irr_small)
                              : g.kernel[rev];
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(small ? Small::kThreads : g.threads), 0,
                       s->stream, P.d_lv[rev], uint32_t(P.lv[rev].size()), off,
                       P.d_lv_block_job[rev], P.d_coef);
    return 1;
}

uint32_t launch_finals(sgn_hip_session *s, sgn_j2k_plan &P, int rev, uint32_t a, uint32_t b,
                       bool clear = true) {
    const uint32_t blocks = final_blocks(P.fin[rev], a, b);
    if (!blocks) return 0;
    const uint32_t off = P.fin[rev][a].block_base;
    if (rev) {
        // int16-plane pass over every block, then the int32-plane pass over
        // the blocks it listed (word 1 of the work-counter area counts them;
        // cleared by a kernel, graph-safe)
        uint32_t *redo_n = P.d_queue + 1;
        uint32_t k = 2;
        if (clear) {
            hipLaunchKernelGGL(sgn_fill_bytes, dim3(1), dim3(64), 0, s->stream,
                               reinterpret_cast<uint8_t *>(redo_n), size_t(4), size_t(1), 0u);
            ++k;
        }
        hipLaunchKernelGGL(sgn_final_rev16, dim3(blocks), dim3(kIdwtThreads), 0, s->stream,
                           P.d_fin[1], uint32_t(P.fin[1].size()), off, P.d_fin_block_job[1],
                           P.d_final_lv, P.d_coef, P.d_fin_redo, redo_n);
        const uint32_t cus = s->cus ? uint32_t(s->cus) : 256u;
        hipLaunchKernelGGL(sgn_final_rev_redo, dim3(std::min(blocks, 2 * cus)), dim3(kIdwtThreads),
                           0, s->stream, P.d_fin[1], uint32_t(P.fin[1].size()),
                           P.d_fin_block_job[1], P.d_final_lv, P.d_coef,
                           static_cast<const uint32_t *>(P.d_fin_redo),
                           static_cast<const uint32_t *>(redo_n));
        return k;
    }
    hipLaunchKernelGGL(sgn_final_irr, dim3(blocks), dim3(kIdwtThreads), 0, s->stream,
                       P.d_fin[0], uint32_t(P.fin[0].size()), off, P.d_fin_block_job[0],
                       P.d_final_lv, P.d_coef);
    return 1;
}

// Store jobs [a, b) (sub-sampled components): one work-group per 64x64
// output block.
uint32_t launch_stores(sgn_hip_session *s, sgn_j2k_plan &P, uint32_t a, uint32_t b) {
    if (a >= b) return 0;
    const StoreJob &first = P.store[a], &last = P.store[b - 1];
    const uint32_t blocks = last.block_base + uint32_t(last.nbx) * last.nby - first.block_base;
    if (!blocks) return 0;
    hipLaunchKernelGGL(sgn_store_sampled, dim3(blocks), dim3(kIdwtThreads), 0, s->stream,
                       P.d_store, uint32_t(P.store.size()), first.block_base, P.d_store_block_job,
                       P.d_coef);
    return 1;
}

// Packet headers of the tiles planned with device packets (fills their
// code-block jobs); the first launch of every decode of such a plan.
uint32_t launch_packets(sgn_hip_session *s, sgn_j2k_plan &P) {
    if (P.pk_tiles.empty()) return 0;
    const uint32_t n = static_cast<uint32_t>(P.pk_tiles.size());
    hipLaunchKernelGGL(sgn_packet_headers, dim3(n), dim3(64), 0, s->stream,
                       P.d_pk_tiles, n, P.d_pk_parts, P.d_pk_pkts, P.d_pk_bands, P.d_pk_cbs,
                       P.d_pk_nodes, P.d_pk_lens, P.d_arena, P.d_cb_lane, P.d_cb_lane_plain,
                       P.d_cb_wave, P.d_pk_status);
    return 1;
}

uint32_t enqueue_batched(sgn_hip_session *s, sgn_j2k_plan &P) {
    uint32_t k = clear_counters(s, P, 2);  // one node for both counters
    k += launch_packets(s, P);
    k += launch_classic(s, P);
    k += launch_lane(s, P, P.d_cb_lane, static_cast<uint32_t>(P.cb_lane.size()), P.d_status_lane,
                     false);
    k += launch_wave(s, P, P.d_cb_wave, static_cast<uint32_t>(P.cb_wave.size()), P.d_status_wave);
    for (uint32_t d = 1; d <= P.depth_max; ++d)
        for (int rev = 0; rev < 2; ++rev)
            k += launch_levels(s, P, rev, P.lv_group[rev][d - 1], P.lv_group[rev][d]);
    for (int rev = 0; rev < 2; ++rev)
        k += launch_finals(s, P, rev, 0, static_cast<uint32_t>(P.fin[rev].size()), false);
    k += launch_stores(s, P, 0, static_cast<uint32_t>(P.store.size()));
    return k;
}

uint32_t enqueue_plain(sgn_hip_session *s, sgn_j2k_plan &P) {
    // packet headers, then the classic code-blocks of every tile (one
    // launch per kernel shape)
    uint32_t k = launch_packets(s, P);
    k += launch_classic(s, P);
    for (size_t t = 0; t < P.ntiles; ++t) {
        const uint32_t a0 = P.lane_plain_begin[t], a1 = P.lane_plain_begin[t + 1];
        // one tile's code-blocks: the cooperative kernel (a tile has far
        // fewer blocks than the lane kernel needs to fill the device)
        k += launch_lane_variant(s, P, kCoopVariant, P.d_cb_lane_plain + a0, a1 - a0,
                                 P.d_status_lane + a0);
        const uint32_t b0 = P.wave_begin[t], b1 = P.wave_begin[t + 1];
        k += launch_wave(s, P, P.d_cb_wave + b0, b1 - b0, P.d_status_wave + b0);
        for (const LevelRange &rg : P.tile_levels[t])
            k += launch_levels(s, P, rg.rev, rg.start, rg.start + rg.count);
        if (P.tile_final_count[t]) {
            const auto ff = P.tile_final_first[t];
            k += launch_finals(s, P, ff.first, ff.second, ff.second + P.tile_final_count[t]);
        }
        k += launch_stores(s, P, P.tile_store[t].first,
                           P.tile_store[t].first + P.tile_store[t].second);
    }
    return k;
}

Status decode_plan(sgn_hip_session *s, sgn_j2k_plan &P, int mode) {
    if (mode == SGN_MODE_PLAIN) {
        s->dispatches += enqueue_plain(s, P);
    } else if (mode == SGN_MODE_BATCHED) {
        P.launches_batched = enqueue_batched(s, P);
        s->dispatches += P.launches_batched;
    } else if (mode == SGN_MODE_GRAPH) {
        if (!P.graph_exec) {
            HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
            P.launches_batched = enqueue_batched(s, P);
            HIP_TRY(hipStreamEndCapture(s->stream, &P.graph));
            HIP_TRY(hipGraphInstantiate(&P.graph_exec, P.graph, nullptr, nullptr, 0));
        }
        HIP_TRY(hipGraphLaunch(P.graph_exec, s->stream));
        s->dispatches += P.launches_batched;
    } else {
        set_last_error("unknown decode mode");
        return kErrArg;
    }
    HIP_TRY(hipGetLastError());
    return kOk;
}

// First failing tile from the per-block status words (after a sync).
Status check_status(sgn_j2k_plan &P, int mode, size_t *first_fail) {
    size_t bad = SIZE_MAX;
    if (!P.pk_tiles.empty()) {
        std::vector<uint8_t> h(P.pk_tiles.size());
        HIP_TRY(hipMemcpy(h.data(), P.d_pk_status, h.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < h.size(); ++k)
            if (h[k]) {
                // the tile of this packet table: its first job's tile
                const PktTile &t = P.pk_tiles[k];
                for (uint32_t c = 0; c < t.ncbs; ++c) {
                    const uint32_t job = P.pk_cbs[t.cb0 + c].job;
                    if (job == kPktNoJob) continue;
                    const CbJob &j = (job & kPktWave) ? P.cb_wave[job & ~kPktWave]
                                                      : P.cb_lane_plain[job];
                    bad = std::min<size_t>(bad, j.tile);
                    break;
                }
            }
    }
    if (!P.cb_lane.empty()) {
        std::vector<uint8_t> h(P.cb_lane.size());
        HIP_TRY(hipMemcpy(h.data(), P.d_status_lane, h.size(), hipMemcpyDeviceToHost));
        const std::vector<CbJob> &jobs = mode == SGN_MODE_PLAIN ? P.cb_lane_plain : P.cb_lane;
        for (size_t k = 0; k < h.size(); ++k)
            if (h[k]) bad = std::min<size_t>(bad, jobs[k].tile);
    }
    if (!P.cb_wave.empty()) {
        std::vector<uint8_t> h(P.cb_wave.size());
        HIP_TRY(hipMemcpy(h.data(), P.d_status_wave, h.size(), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < h.size(); ++k)
            if (h[k]) bad = std::min<size_t>(bad, P.cb_wave[k].tile);
    }
    if (bad != SIZE_MAX) {
        if (first_fail) *first_fail = bad;
        set_last_error("code-block decode failure in tile " + std::to_string(bad));
        return kErrBlock;
    }
    return kOk;
}

void drop_graph(sgn_j2k_plan &P) {
    if (P.graph_exec) (void)hipGraphExecDestroy(P.graph_exec);
    if (P.graph) (void)hipGraphDestroy(P.graph);
    P.graph_exec = nullptr;
    P.graph = nullptr;
}

Status make_plan(sgn_hip_session *s, const sgn_tile_req *jobs, size_t n,
                 std::unique_ptr<sgn_j2k_plan> &out, int32_t *tile_status,
                 DeviceBlocks *blocks = nullptr) {
    if (!s || (!jobs && n)) {
        set_last_error("null session or jobs");
        return kErrArg;
    }
    HIP_TRY(hipSetDevice(s->device));
    auto P = std::make_unique<sgn_j2k_plan>();
    P->blocks = blocks;
    Status st = build_plan(jobs, n, *P, tile_status);
    if (st) return st;
    st = upload_plan(s, jobs, *P, s->stream, 0, true);
    if (st) return st;
    out = std::move(P);
    return kOk;
}

// Per-device address of `scratch`, resolved once per device with that
// device current (sessions do it at creation), and a lock that serialises
// the no-allocation calls of this process on that device: they share its
// scratch.
struct NoAllocSlot {
    std::mutex lock;
    void *ptr = nullptr;
};

constexpr int kMaxDevices = 64;

NoAllocSlot &noalloc_slot(int device) {
    static NoAllocSlot slots[kMaxDevices];
    return slots[device];
}

Status noalloc_scratch(int device, void **out) {
    static std::mutex resolve;
    if (device < 0 || device >= kMaxDevices) {
        set_last_error("device index beyond the no-allocation scratch table");
        return kErrArg;
    }
    std::lock_guard<std::mutex> g(resolve);
    NoAllocSlot &sl = noalloc_slot(device);
    if (!sl.ptr) {
        HIP_TRY(hipSetDevice(device));
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_last_error("hipGetSymbolAddress returned no address for the scratch array");
            return kErrDevice;
        }
        sl.ptr = p;
    }
    *out = sl.ptr;
    return kOk;
}

// Launch helpers enqueue on s->stream; this points it at another stream of
// the session for one scope.
class OnStream {
  public:
    OnStream(sgn_hip_session *s, hipStream_t st) : s_(s), keep_(s->stream) { s->stream = st; }
    ~OnStream() { s_->stream = keep_; }

  private:
    sgn_hip_session *s_;
    hipStream_t keep_;
};

// Tiles per chunk of a pipelined batch decode (SGN_PIPELINE_TILES
// overrides; 0 or a value >= n decodes the batch as one plan).  768 tiles
// of 256x256 RGB: chunks large enough for the lane kernel (57600
// code-blocks), small enough that the first parse and the last decode,
// which nothing overlaps, stay short.
size_t pipeline_tiles() {
    const char *e = std::getenv("SGN_PIPELINE_TILES");
    if (e && *e) return size_t(std::strtoull(e, nullptr, 10));
    return 768;
}

// Batch decode from host codestreams, pipelined in chunks: while chunk k's
// codestreams and tables travel on the copy stream and chunk k-1 decodes
// (chunks alternate between two decode streams, so neighbouring chunks
// also overlap on the device), the host parses chunk k+1.  Each chunk is
// its own plan in a device block of the session's cache.  Synchronous;
// reports the first failing tile index like TileBatchError.
Status decode_pipelined(sgn_hip_session *s, const sgn_tile_req *reqs, size_t n, size_t chunk,
                        int32_t *tile_status, size_t *first_fail) {
    HIP_TRY(hipSetDevice(s->device));
    std::vector<std::unique_ptr<sgn_j2k_plan>> plans;
    std::vector<size_t> first;
    Status parse_st = kOk;
    size_t parse_bad = 0;
    std::string parse_msg;
    hipEvent_t join = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    std::unique_ptr<std::remove_pointer_t<hipEvent_t>, hipError_t (*)(hipEvent_t)> join_guard(
        join, hipEventDestroy);
    auto chunk_len = [&](size_t a) { return std::min(chunk, n - a); };
    auto status_at = [&](size_t a) { return tile_status ? tile_status + a : nullptr; };
    // chunk k + 1 is parsed (on a helper thread and the worker pool) while
    // chunk k is merged, uploaded and enqueued on this thread
    const bool dev_packets = device_packets_enabled();
    auto parse_into = [&](size_t a, ParsedBatch &B) {
        try {
            parse_batch(reqs + a, chunk_len(a), B, status_at(a), dev_packets);
        } catch (const std::bad_alloc &) {
            B.st = kErrNoMem;
            B.msg = "out of host memory";
        }
    };
    ParsedBatch cur, nxt;
    if (n) parse_into(0, cur);
    for (size_t a = 0, c = 0; a < n; a += chunk, ++c) {
        const size_t m = chunk_len(a);
        std::thread helper;
        if (a + chunk < n) helper = std::thread(parse_into, a + chunk, std::ref(nxt));
        struct Join {
            std::thread &t;
            ~Join() {
                if (t.joinable()) t.join();
            }
        } join_helper{helper};
        auto P = std::make_unique<sgn_j2k_plan>();
        P->blocks = s->blocks;
        Status st = cur.st;
        std::string msg = cur.msg;
        if (!st) {
            st = merge_batch(cur, *P, false);
            if (st) msg = last_error();
        }
        if (st) {
            parse_st = st;
            parse_msg = msg;
            for (size_t i = 0; i < m; ++i)
                if (!tile_status || tile_status[a + i]) {
                    parse_bad = a + i;
                    break;
                }
            break;
        }
        const int k = int(c & 1);
        hipStream_t cs = k ? s->aux_stream : s->stream;
        st = upload_plan(s, reqs + a, *P, s->copy_stream, k, false);
        if (!st && hipStreamWaitEvent(cs, s->slot[k].done, 0) != hipSuccess) {
            set_last_error("hipStreamWaitEvent failed");
            st = kErrDevice;
        }
        if (!st) {
            OnStream on(s, cs);
            st = decode_plan(s, *P, SGN_MODE_BATCHED);
        }
        if (st) {
            // let the chunks already enqueued finish before their plans go
            for (hipStream_t q : {s->copy_stream, s->stream, s->aux_stream})
                (void)hipStreamSynchronize(q);
            return st;
        }
        plans.push_back(std::move(P));
        first.push_back(a);
        if (helper.joinable()) helper.join();
        std::swap(cur, nxt);
    }
    HIP_TRY(hipEventRecord(join, s->aux_stream));
    HIP_TRY(hipStreamWaitEvent(s->stream, join, 0));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (size_t c = 0; c < plans.size(); ++c) {
        size_t bad = 0;
        const Status st = check_status(*plans[c], SGN_MODE_BATCHED, &bad);
        if (st) {
            if (first_fail) *first_fail = first[c] + bad;
            set_last_error("code-block decode failure in tile " + std::to_string(first[c] + bad));
            return st;
        }
    }
    if (parse_st) {
        if (first_fail) *first_fail = parse_bad;
        set_last_error(parse_msg);
        return parse_st;
    }
    return kOk;
}

// Uniform-format batch as per-tile requests.
std::vector<sgn_tile_req> as_requests(const sgn_tile_job *jobs, size_t n, int fmt, uint32_t den) {
    std::vector<sgn_tile_req> r(n);
    for (size_t i = 0; i < n; ++i)
        r[i] = sgn_tile_req{jobs[i].bytes, jobs[i].len, jobs[i].dst, jobs[i].pitch,
                            jobs[i].capacity, fmt, den};
    return r;
}

}  // namespace

#define API_GUARD_BEGIN try {
#define API_GUARD_END                                \
    }                                                \
    catch (const std::bad_alloc &) {                 \
        set_last_error("out of host memory");        \
        return SGN_E_NOMEM;                          \
    }                                                \
    catch (const std::exception &e) {                \
        set_last_error(e.what());                    \
        return SGN_E_ARG;                            \
    }

extern "C" {

int sgn_hip_session_create(int device, sgn_hip_session **out) {
    API_GUARD_BEGIN
    if (!out) return SGN_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_last_error("no HIP device available");
        return SGN_E_DEVICE;
    }
    if (device < 0 || device >= count) {
        set_last_error("device index out of range");
        return SGN_E_ARG;
    }
    auto s = std::make_unique<sgn_hip_session>();
    s->device = device;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&s->aux_stream, hipStreamNonBlocking));
    s->blocks = new DeviceBlocks(uint64_t(16) << 30);
    const HtTables &T = ht_tables();
    std::vector<uint16_t> tab(2624);
    std::memcpy(tab.data(), T.vlc0, 2048);
    std::memcpy(tab.data() + 1024, T.vlc1, 2048);
    std::memcpy(tab.data() + 2048, T.uvlc0, 640);
    std::memcpy(tab.data() + 2368, T.uvlc1, 512);
    HIP_TRY(hipMalloc(&s->d_tables, tab.size() * 2));
    HIP_TRY(hipMemcpy(s->d_tables, tab.data(), tab.size() * 2, hipMemcpyHostToDevice));
    {
        void *scr = nullptr;
        const Status st = noalloc_scratch(device, &scr);
        if (st) return st;
    }
    *out = s.release();
    return SGN_OK;
    API_GUARD_END
}

void sgn_hip_session_destroy(sgn_hip_session *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    for (hipStream_t st : {s->stream, s->copy_stream, s->aux_stream})
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    for (PinnedSlot &sl : s->slot) {
        if (sl.p) (void)hipHostFree(sl.p);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    delete s->blocks;
    if (s->d_tables) (void)hipFree(s->d_tables);
    delete s;
}

void *sgn_hip_session_stream(sgn_hip_session *s) { return s ? s->stream : nullptr; }

int sgn_hip_session_synchronize(sgn_hip_session *s) {
    if (!s) return SGN_E_ARG;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SGN_OK;
}

uint64_t sgn_hip_session_dispatches(sgn_hip_session *s) { return s ? s->dispatches : 0; }

int sgn_device_alloc(sgn_hip_session *s, size_t bytes, void **out) {
    if (!s || !out) return SGN_E_ARG;
    *out = nullptr;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMalloc(out, bytes ? bytes : 1));
    return SGN_OK;
}

void sgn_device_free(sgn_hip_session *s, void *p) {
    if (!p) return;
    if (s) {
        (void)hipSetDevice(s->device);
        (void)hipStreamSynchronize(s->stream);
    }
    (void)hipFree(p);
}

int sgn_device_fill(sgn_hip_session *s, void *p, size_t bytes, uint8_t value) {
    if (!s || (!p && bytes)) return SGN_E_ARG;
    if (!bytes) return SGN_OK;
    HIP_TRY(hipSetDevice(s->device));
    const size_t words = bytes / 4;
    const uint32_t v4 = 0x01010101u * value;
    const uint32_t grid = static_cast<uint32_t>(std::min<size_t>((bytes + 1023) / 1024, 65535));
    hipLaunchKernelGGL(sgn_fill_bytes, dim3(grid), dim3(256), 0, s->stream,
                       static_cast<uint8_t *>(p), bytes, words, v4);
    s->dispatches++;
    HIP_TRY(hipGetLastError());
    return SGN_OK;
}

int sgn_device_copy(sgn_hip_session *s, void *dst, const void *src, size_t bytes) {
    if (!s || ((!dst || !src) && bytes)) return SGN_E_ARG;
    if (!bytes) return SGN_OK;
    HIP_TRY(hipSetDevice(s->device));
    const uint32_t grid = static_cast<uint32_t>(std::min<size_t>((bytes + 4095) / 4096, 65535));
    hipLaunchKernelGGL(sgn_copy_bytes, dim3(grid), dim3(256), 0, s->stream,
                       static_cast<uint8_t *>(dst), static_cast<const uint8_t *>(src), bytes);
    s->dispatches++;
    HIP_TRY(hipGetLastError());
    return SGN_OK;
}

int sgn_device_download(sgn_hip_session *s, void *host, const void *dev, size_t bytes) {
    if (!s || (!bytes)) return bytes ? SGN_E_ARG : SGN_OK;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SGN_OK;
}

int sgn_device_upload(sgn_hip_session *s, void *dev, const void *host, size_t bytes) {
    if (!s || (!bytes)) return bytes ? SGN_E_ARG : SGN_OK;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SGN_OK;
}

int sgn_device_download_2d(sgn_hip_session *s, void *host, size_t host_stride, const void *dev,
                           size_t dev_pitch, size_t row_bytes, size_t rows) {
    if (!s || !host || !dev) return SGN_E_ARG;
    if (!rows || !row_bytes) return SGN_OK;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipMemcpy2DAsync(host, host_stride, dev, dev_pitch, row_bytes, rows,
                             hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SGN_OK;
}

int sgn_host_alloc(size_t bytes, void **out) {
    if (!out) return SGN_E_ARG;
    *out = nullptr;
    HIP_TRY(hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault));
    return SGN_OK;
}

void sgn_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

int sgn_hip_event_create(void **ev) {
    if (!ev) return SGN_E_ARG;
    hipEvent_t e = nullptr;
    HIP_TRY(hipEventCreate(&e));
    *ev = e;
    return SGN_OK;
}

void sgn_hip_event_destroy(void *ev) {
    if (ev) (void)hipEventDestroy(static_cast<hipEvent_t>(ev));
}

int sgn_hip_event_record(sgn_hip_session *s, void *ev) {
    if (!s || !ev) return SGN_E_ARG;
    HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(ev), s->stream));
    return SGN_OK;
}

int sgn_hip_event_elapsed_ms(void *a, void *b, float *ms) {
    if (!a || !b || !ms) return SGN_E_ARG;
    HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(b)));
    HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
    return SGN_OK;
}

int sgn_j2k_plan_create(sgn_hip_session *s, const sgn_tile_job *jobs, size_t n, int fmt,
                        uint32_t den, sgn_j2k_plan **out, int32_t *tile_status) {
    API_GUARD_BEGIN
    if (!out) return SGN_E_ARG;
    *out = nullptr;
    if (!jobs && n) return SGN_E_ARG;
    std::unique_ptr<sgn_j2k_plan> P;
    const std::vector<sgn_tile_req> reqs = as_requests(jobs, n, fmt, den);
    const Status st = make_plan(s, reqs.data(), n, P, tile_status);
    if (st) return st;
    *out = P.release();
    return SGN_OK;
    API_GUARD_END
}

int sgn_j2k_plan_create_mixed(sgn_hip_session *s, const sgn_tile_req *reqs, size_t n,
                              sgn_j2k_plan **out, int32_t *tile_status) {
    API_GUARD_BEGIN
    if (!out) return SGN_E_ARG;
    *out = nullptr;
    std::unique_ptr<sgn_j2k_plan> P;
    const Status st = make_plan(s, reqs, n, P, tile_status);
    if (st) return st;
    *out = P.release();
    return SGN_OK;
    API_GUARD_END
}

void sgn_j2k_plan_destroy(sgn_j2k_plan *p) { delete p; }

int sgn_j2k_plan_set_output(sgn_j2k_plan *P, size_t i, void *dst, size_t pitch, size_t capacity) {
    API_GUARD_BEGIN
    if (!P || i >= P->ntiles) return SGN_E_ARG;
    if (pitch < P->tile_row_bytes[i] ||
        (P->tile_height[i] &&
         capacity < pitch * (P->tile_height[i] - 1) + P->tile_row_bytes[i])) {
        set_last_error("output surface too small for the decoded image");
        return SGN_E_BUFFER;
    }
    const auto ff = P->tile_final_first[i];
    std::vector<FinalJob> &v = P->fin[ff.first];
    const uint64_t bpp = P->tile_width[i] ? P->tile_row_bytes[i] / P->tile_width[i] : 1;
    for (uint32_t k = 0; k < P->tile_final_count[i]; ++k) {
        FinalJob &fj = v[ff.second + k];
        fj.pitch = pitch;
        fj.dst = reinterpret_cast<uint64_t>(dst) + uint64_t(fj.out_y) * pitch + fj.out_x * bpp;
    }
    HIP_TRY(hipMemcpy(P->d_fin[ff.first] + ff.second, v.data() + ff.second,
                      P->tile_final_count[i] * sizeof(FinalJob), hipMemcpyHostToDevice));
    drop_graph(*P);
    return SGN_OK;
    API_GUARD_END
}

int sgn_ht_lane_variant_count(void) { return kRowsVariantCount; }

const char *sgn_ht_lane_variant_name(int v) {
    return (v >= 0 && v < kRowsVariantCount) ? kRowsVariants[v].name : nullptr;
}

int sgn_j2k_plan_set_lane_variant(sgn_j2k_plan *P, int v) {
    if (!P || v < 0 || v >= kRowsVariantCount) return SGN_E_ARG;
    if (P->lane_variant != v) drop_graph(*P);
    P->lane_variant = v;
    return SGN_OK;
}

int sgn_j2k_plan_get_lane_variant(const sgn_j2k_plan *P) { return P ? P->lane_variant : -1; }

int sgn_j2k_plan_autotune(sgn_hip_session *s, sgn_j2k_plan *P, int reps, float *ms, size_t cap,
                          int *chosen) {
    API_GUARD_BEGIN
    if (!s || !P || reps < 1) return SGN_E_ARG;
    HIP_TRY(hipSetDevice(s->device));
    const uint32_t n = static_cast<uint32_t>(P->cb_lane.size());
    s->dispatches += launch_packets(s, *P);  // the jobs the variants decode
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_last_error("hipEventCreate failed");
        return SGN_E_DEVICE;
    }
    int best = P->lane_variant;
    float best_ms = 0.0f;
    Status st = kOk;
    for (int v = 0; v < kRowsVariantCount && !st; ++v) {
        float t = 0.0f;
        if (n) {
            s->dispatches += launch_lane_variant(s, *P, v, P->d_cb_lane, n, P->d_status_lane);
            bool ok = hipEventRecord(e0, s->stream) == hipSuccess;
            for (int r = 0; r < reps && ok; ++r)
                s->dispatches += launch_lane_variant(s, *P, v, P->d_cb_lane, n, P->d_status_lane);
            ok = ok && hipEventRecord(e1, s->stream) == hipSuccess &&
                 hipEventSynchronize(e1) == hipSuccess &&
                 hipEventElapsedTime(&t, e0, e1) == hipSuccess && hipGetLastError() == hipSuccess;
            if (!ok) {
                set_last_error("lane-kernel variant failed to run");
                st = kErrDevice;
                break;
            }
            t /= float(reps);
        }
        if (ms && size_t(v) < cap) ms[v] = t;
        if (v == 0 || t < best_ms) {
            best = v;
            best_ms = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    if (P->lane_variant != best) drop_graph(*P);
    P->lane_variant = best;
    if (chosen) *chosen = best;
    return SGN_OK;
    API_GUARD_END
}

int sgn_dwt_shape_count(void) { return kDwtShapeCount; }

const char *sgn_dwt_shape_name(int v) {
    return (v >= 0 && v < kDwtShapeCount) ? kDwtShapes[v].name : nullptr;
}

int sgn_j2k_plan_set_dwt_shape(sgn_j2k_plan *P, int v) {
    if (!P || v < 0 || v >= kDwtShapeCount) return SGN_E_ARG;
    if (P->dwt_shape != v) drop_graph(*P);
    P->dwt_shape = v;
    return SGN_OK;
}

int sgn_j2k_plan_get_dwt_shape(const sgn_j2k_plan *P) { return P ? P->dwt_shape : -1; }

int sgn_j2k_plan_autotune_dwt(sgn_hip_session *s, sgn_j2k_plan *P, int reps, float *ms, size_t cap,
                              int *chosen) {
    API_GUARD_BEGIN
    if (!s || !P || reps < 1) return SGN_E_ARG;
    HIP_TRY(hipSetDevice(s->device));
    const int keep = P->dwt_shape;
    int rc = sgn_j2k_plan_decode_stage(s, P, SGN_STAGE_BLOCKS);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_last_error("hipEventCreate failed");
        return SGN_E_DEVICE;
    }
    int best = keep;
    float best_ms = 0.0f;
    for (int v = 0; v < kDwtShapeCount && !rc; ++v) {
        P->dwt_shape = v;
        float t = 0.0f;
        rc = sgn_j2k_plan_decode_stage(s, P, SGN_STAGE_LEVELS);  // warm-up
        bool ok = !rc && hipEventRecord(e0, s->stream) == hipSuccess;
        for (int r = 0; r < reps && ok; ++r) ok = !sgn_j2k_plan_decode_stage(s, P, SGN_STAGE_LEVELS);
        ok = ok && hipEventRecord(e1, s->stream) == hipSuccess &&
             hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&t, e0, e1) == hipSuccess;
        if (!ok) {
            if (!rc) {
                set_last_error("synthesis-level shape failed to run");
                rc = SGN_E_DEVICE;
            }
            break;
        }
        t /= float(reps);
        if (ms && size_t(v) < cap) ms[v] = t;
        if (v == 0 || t < best_ms) {
            best = v;
            best_ms = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    P->dwt_shape = keep;
    if (rc) return rc;
    if (P->dwt_shape != best) drop_graph(*P);
    P->dwt_shape = best;
    if (chosen) *chosen = best;
    return SGN_OK;
    API_GUARD_END
}

int sgn_j2k_plan_decode(sgn_hip_session *s, sgn_j2k_plan *P, int mode) {
    API_GUARD_BEGIN
    if (!s || !P) return SGN_E_ARG;
    HIP_TRY(hipSetDevice(s->device));
    return decode_plan(s, *P, mode);
    API_GUARD_END
}

int sgn_j2k_plan_decode_stage(sgn_hip_session *s, sgn_j2k_plan *P, int stage) {
    API_GUARD_BEGIN
    if (!s || !P) return SGN_E_ARG;
    HIP_TRY(hipSetDevice(s->device));
    uint32_t k = 0;
    if (stage == SGN_STAGE_BLOCKS) {
        k += launch_packets(s, *P);
        k += launch_classic(s, *P);
        k += launch_lane(s, *P, P->d_cb_lane, static_cast<uint32_t>(P->cb_lane.size()),
                         P->d_status_lane);
        k += launch_wave(s, *P, P->d_cb_wave, static_cast<uint32_t>(P->cb_wave.size()),
                         P->d_status_wave);
    } else if (stage == SGN_STAGE_LEVELS) {
        for (uint32_t d = 1; d <= P->depth_max; ++d)
            for (int rev = 0; rev < 2; ++rev)
                k += launch_levels(s, *P, rev, P->lv_group[rev][d - 1], P->lv_group[rev][d]);
    } else if (stage == SGN_STAGE_FINAL) {
        for (int rev = 0; rev < 2; ++rev)
            k += launch_finals(s, *P, rev, 0, static_cast<uint32_t>(P->fin[rev].size()));
        k += launch_stores(s, *P, 0, static_cast<uint32_t>(P->store.size()));
    } else {
        set_last_error("unknown stage");
        return SGN_E_ARG;
    }
    s->dispatches += k;
    HIP_TRY(hipGetLastError());
    return SGN_OK;
    API_GUARD_END
}

int sgn_j2k_plan_get_stats(const sgn_j2k_plan *P, sgn_j2k_plan_stats *st) {
    if (!P || !st) return SGN_E_ARG;
    st->tiles = P->ntiles;
    st->code_blocks = P->cb_lane.size() + P->cb_wave.size() + P->classic.size();
    st->coded_bytes = P->coded_bytes;
    st->output_bytes = P->out_bytes;
    st->output_pixels = P->out_pixels;
    st->coefficient_bytes = P->coef_elems * 4;
    st->launches_batched = P->launches_batched;
    st->narrow_samples = P->narrow ? 1u : 0u;
    st->host_parse_ms = float(P->t_parse);
    st->host_merge_ms = float(P->t_merge);
    st->host_stage_ms = float(P->t_stage);
    st->host_copy_ms = float(P->t_copy);
    st->pruned_code_blocks = P->pruned_cbs;
    return SGN_OK;
}

int sgn_j2k_plan_tile_dims(const sgn_j2k_plan *P, size_t i, uint32_t *w, uint32_t *h,
                           size_t *row_bytes) {
    if (!P || i >= P->ntiles) return SGN_E_ARG;
    if (w) *w = P->tile_width[i];
    if (h) *h = P->tile_height[i];
    if (row_bytes) *row_bytes = P->tile_row_bytes[i];
    return SGN_OK;
}

int sgn_j2k_decode_tiles_device(sgn_hip_session *s, const sgn_tile_job *jobs, size_t n, int fmt,
                                uint32_t den, int mode, size_t *first_fail) {
    API_GUARD_BEGIN
    if (!s || (!jobs && n)) return SGN_E_ARG;
    std::vector<int32_t> ts(n, 0);
    std::unique_ptr<sgn_j2k_plan> P;
    const std::vector<sgn_tile_req> reqs = as_requests(jobs, n, fmt, den);
    const size_t chunk = pipeline_tiles();
    if (mode == SGN_MODE_BATCHED && chunk && n > chunk)
        return decode_pipelined(s, reqs.data(), n, chunk, ts.data(), first_fail);
    Status st = make_plan(s, reqs.data(), n, P, ts.data());
    if (st) {
        if (first_fail)
            for (size_t i = 0; i < n; ++i)
                if (ts[i]) {
                    *first_fail = i;
                    break;
                }
        return st;
    }
    st = decode_plan(s, *P, mode);
    if (st) return st;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return check_status(*P, mode, first_fail);
    API_GUARD_END
}

int sgn_j2k_decode_tiles_device_noalloc(sgn_hip_session *s, const sgn_tile_job *jobs, size_t n,
                                        int fmt, uint32_t den, size_t *first_fail) {
    API_GUARD_BEGIN
    if (!s || (!jobs && n)) return SGN_E_ARG;
    void *scr = nullptr;
    Status st = noalloc_scratch(s->device, &scr);
    if (st) return st;
    std::lock_guard<std::mutex> g(noalloc_slot(s->device).lock);
    HIP_TRY(hipSetDevice(s->device));
    const std::vector<sgn_tile_req> reqs = as_requests(jobs, n, fmt, den);
    std::vector<int32_t> ts(n, 0);
    sgn_j2k_plan P;
    st = build_plan(reqs.data(), n, P, ts.data());
    if (st) {
        if (first_fail)
            for (size_t i = 0; i < n; ++i)
                if (ts[i]) {
                    *first_fail = i;
                    break;
                }
        return st;
    }
    st = upload_plan(s, reqs.data(), P, s->stream, 0, true, scr, kNoAllocScratchBytes);
    if (st) {
        (void)hipStreamSynchronize(s->stream);
        return st;
    }
    st = decode_plan(s, P, SGN_MODE_BATCHED);
    if (st) {
        (void)hipStreamSynchronize(s->stream);
        return st;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return check_status(P, SGN_MODE_BATCHED, first_fail);
    API_GUARD_END
}

uint64_t sgn_j2k_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

int sgn_j2k_decode_tiles_region_device(sgn_hip_session *s, const sgn_tile_req *reqs, size_t n,
                                       size_t *first_fail) {
    API_GUARD_BEGIN
    if (!s || (!reqs && n)) return SGN_E_ARG;
    std::vector<int32_t> ts(n, 0);
    std::unique_ptr<sgn_j2k_plan> P;
    Status st = make_plan(s, reqs, n, P, ts.data());
    if (st) {
        if (first_fail)
            for (size_t i = 0; i < n; ++i)
                if (ts[i]) {
                    *first_fail = i;
                    break;
                }
        return st;
    }
    st = decode_plan(s, *P, SGN_MODE_BATCHED);
    if (st) return st;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return check_status(*P, SGN_MODE_BATCHED, first_fail);
    API_GUARD_END
}

int sgn_j2k_decode_region_device(sgn_hip_session *s, const uint8_t *bytes, size_t len, int fmt,
                                 uint32_t den, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                                 void *dst, size_t pitch, size_t cap) {
    sgn_tile_req r{bytes, len, dst, pitch, cap, fmt, den, 1u, x, y, w, h};
    size_t ff = 0;
    return sgn_j2k_decode_tiles_region_device(s, &r, 1, &ff);
}

int sgn_j2k_decode_region_into(sgn_hip_session *s, const uint8_t *bytes, size_t len, int fmt,
                               uint32_t den, uint32_t x, uint32_t y, uint32_t w, uint32_t h,
                               uint8_t *out, size_t out_len, size_t stride) {
    API_GUARD_BEGIN
    if (!s || !bytes || !out) return SGN_E_ARG;
    uint32_t rx, ry, rw, rh;
    const int rc0 = sgn_j2k_region_dims(bytes, len, den, x, y, w, h, &rx, &ry, &rw, &rh);
    if (rc0) return rc0;
    Container c;
    Status st = locate_codestream(bytes, len, c);
    if (st) return st;
    MainHeader hd;
    if ((st = read_main_header(c.cs, c.cs_len, hd))) return st;
    OutputGeom og;
    if ((st = output_geometry(hd, fmt, den, og))) return st;
    const size_t row = size_t(rw) * og.channels * og.bytes_per_sample;
    if (stride < row || (rh && out_len < stride * (rh - 1) + row)) {
        set_last_error("output buffer too small");
        return SGN_E_BUFFER;
    }
    if (!rw || !rh) return SGN_OK;
    HIP_TRY(hipSetDevice(s->device));
    void *d = nullptr;
    HIP_TRY(hipMalloc(&d, row * rh));
    int rc = sgn_j2k_decode_region_device(s, bytes, len, fmt, den, x, y, w, h, d, row, row * rh);
    if (rc == SGN_OK &&
        hipMemcpy2D(out, stride, d, row, row, rh, hipMemcpyDeviceToHost) != hipSuccess) {
        set_last_error("download failed");
        rc = SGN_E_DEVICE;
    }
    (void)hipFree(d);
    return rc;
    API_GUARD_END
}

int sgn_j2k_decode_device(sgn_hip_session *s, const uint8_t *bytes, size_t len, int fmt,
                          uint32_t den, void *dst, size_t pitch, size_t cap) {
    sgn_tile_job j{bytes, len, dst, pitch, cap};
    size_t ff = 0;
    return sgn_j2k_decode_tiles_device(s, &j, 1, fmt, den, SGN_MODE_BATCHED, &ff);
}

int sgn_j2k_decode_into(sgn_hip_session *s, const uint8_t *bytes, size_t len, int fmt,
                        uint32_t den, uint8_t *out, size_t out_len, size_t stride) {
    API_GUARD_BEGIN
    if (!s || !bytes || !out) return SGN_E_ARG;
    Container c;
    Status st = locate_codestream(bytes, len, c);
    if (st) return st;
    MainHeader h;
    st = read_main_header(c.cs, c.cs_len, h);
    if (st) return st;
    OutputGeom og;
    st = output_geometry(h, fmt, den, og);
    if (st) return st;
    if (stride < og.row_bytes || (og.height && out_len < stride * (og.height - 1) + og.row_bytes)) {
        set_last_error("output buffer too small");
        return SGN_E_BUFFER;
    }
    HIP_TRY(hipSetDevice(s->device));
    void *d = nullptr;
    const size_t bytes_dev = std::max<size_t>(1, og.row_bytes * og.height);
    HIP_TRY(hipMalloc(&d, bytes_dev));
    int rc = sgn_j2k_decode_device(s, bytes, len, fmt, den, d, og.row_bytes, bytes_dev);
    if (rc == SGN_OK && og.height) {
        if (hipMemcpy2D(out, stride, d, og.row_bytes, og.row_bytes, og.height,
                        hipMemcpyDeviceToHost) != hipSuccess) {
            set_last_error("download failed");
            rc = SGN_E_DEVICE;
        }
    }
    (void)hipFree(d);
    return rc;
    API_GUARD_END
}

// ---- stage hooks ----------------------------------------------------------

static int hook_blocks(sgn_hip_session *s, const uint8_t *data, size_t data_len,
                       const sgn_ht_code_block_job *jobs, size_t n, int32_t *out, size_t out_len,
                       bool raw) {
    if (!s || (!data && data_len) || (!jobs && n) || (!out && out_len)) return SGN_E_ARG;
    HIP_TRY(hipSetDevice(s->device));
    std::vector<CbJob> lane, wave;
    bool refined = false;
    for (size_t i = 0; i < n; ++i) {
        const sgn_ht_code_block_job &q = jobs[i];
        CbJob j{};
        if (q.width == 0 || q.height == 0) continue;
        if (q.data_offset + q.cleanup_length + q.refinement_length > data_len ||
            q.output_offset + uint64_t(q.output_stride) * (q.height - 1) + q.width > out_len ||
            q.output_stride < q.width || q.width * q.height > 4096 || q.cleanup_length > 65535 ||
            q.refinement_length > 65535 || data_len >= (1ull << 32)) {
            set_last_error("code-block job outside its buffers");
            return SGN_E_ARG;
        }
        j.src = static_cast<uint32_t>(q.data_offset);
        j.dst = q.output_offset;
        j.stride = q.output_stride;
        j.lcup = static_cast<uint16_t>(q.cleanup_length);
        j.lref = static_cast<uint16_t>(q.refinement_length);
        j.w = static_cast<uint16_t>(q.width);
        j.h = static_cast<uint16_t>(q.height);
        j.zblk = q.missing_bit_planes;
        j.passes = q.number_of_coding_passes;
        j.kmax = q.num_bitplanes;
        j.flags = q.stripe_causal ? kCbStripeCausal : 0;
        const uint32_t actual = q.num_bitplanes > q.missing_bit_planes
                                    ? q.num_bitplanes - q.missing_bit_planes : 0;
        if (q.num_bitplanes == 0 || q.number_of_coding_passes == 0 || actual == 0)
            j.flags |= kCbZero;
        j.tile = static_cast<uint32_t>(i);
        if (!raw && lane_eligible(j)) {
            lane.push_back(j);
            refined |= lane_refines(j);
        } else {
            wave.push_back(j);
        }
    }
    uint8_t *d_data = nullptr, *d_st = nullptr;
    uint32_t *d_out = nullptr;
    CbJob *d_lane = nullptr, *d_wave = nullptr;
    int rc = SGN_OK;
    std::vector<uint8_t> stl(lane.size()), stw(wave.size());
    auto cleanup = [&]() {
        void *p[] = {d_data, d_st, d_out, d_lane, d_wave};
        for (void *q : p)
            if (q) (void)hipFree(q);
    };
    // chunked stream readers load whole aligned 8-byte words: pad to 16
    if (hipMalloc(&d_data, (std::max<size_t>(16, data_len) + 15) & ~size_t(15)) != hipSuccess ||
        hipMalloc(&d_out, std::max<size_t>(4, out_len * 4)) != hipSuccess ||
        hipMalloc(&d_st, std::max<size_t>(1, n)) != hipSuccess) {
        cleanup();
        set_last_error("hipMalloc failed");
        return SGN_E_DEVICE;
    }
    if (data_len) (void)hipMemcpy(d_data, data, data_len, hipMemcpyHostToDevice);
    (void)hipMemset(d_out, 0, out_len * 4);
    if (dev_upload(d_lane, lane) || dev_upload(d_wave, wave)) {
        cleanup();
        return SGN_E_DEVICE;
    }
    // a hook call is one sub-band or a few: the cooperative kernel
    if (!lane.empty())
        kRowsVariants[kCoopVariant].launch[0](s->stream, uint32_t(lane.size()), d_lane,
                                              uint32_t(lane.size()), d_data, d_out, s->d_tables,
                                              d_st, nullptr, refined);
    if (!wave.empty())
        hipLaunchKernelGGL((sgn_ht_coop<1024, true, int32_t>), dim3(wave.size()), dim3(64), 0,
                           s->stream, d_wave, uint32_t(wave.size()), d_data,
                           reinterpret_cast<int32_t *>(d_out), s->d_tables, d_st + lane.size(),
                           raw ? 1u : 0u);
    s->dispatches += (lane.empty() ? 0 : 1) + (wave.empty() ? 0 : 1);
    if (hipStreamSynchronize(s->stream) != hipSuccess ||
        hipMemcpy(out, d_out, out_len * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        cleanup();
        set_last_error("block decode launch failed");
        return SGN_E_DEVICE;
    }
    std::vector<uint8_t> sts(n);
    (void)hipMemcpy(sts.data(), d_st, lane.size() + wave.size(), hipMemcpyDeviceToHost);
    for (size_t k = 0; k < lane.size() + wave.size(); ++k)
        if (sts[k]) {
            set_last_error("code-block decode failure");
            rc = SGN_E_BLOCK;
            break;
        }
    cleanup();
    return rc;
}

int sgn_ht_decode_code_blocks(sgn_hip_session *s, const uint8_t *data, size_t data_len,
                              const sgn_ht_code_block_job *jobs, size_t n, float *out,
                              size_t out_len) {
    API_GUARD_BEGIN
    std::vector<int32_t> tmp(out_len);
    const int rc = hook_blocks(s, data, data_len, jobs, n, tmp.data(), out_len, false);
    if (rc) return rc;
    std::vector<float> step(out_len, 1.0f);
    for (size_t i = 0; i < n; ++i) {
        const sgn_ht_code_block_job &q = jobs[i];
        for (uint32_t y = 0; y < q.height; ++y)
            for (uint32_t x = 0; x < q.width; ++x)
                step[q.output_offset + size_t(y) * q.output_stride + x] = q.dequantization_step;
    }
    for (size_t i = 0; i < out_len; ++i) out[i] = static_cast<float>(tmp[i]) * step[i];
    return SGN_OK;
    API_GUARD_END
}

int sgn_ht_decode_code_blocks_raw(sgn_hip_session *s, const uint8_t *data, size_t data_len,
                                  const sgn_ht_code_block_job *jobs, size_t n, uint32_t *out,
                                  size_t out_len) {
    API_GUARD_BEGIN
    return hook_blocks(s, data, data_len, jobs, n, reinterpret_cast<int32_t *>(out), out_len, true);
    API_GUARD_END
}

int sgn_idwt_level(sgn_hip_session *s, int reversible, uint32_t x0, uint32_t y0, uint32_t x1,
                   uint32_t y1, const float *ll, const float *hl, const float *lh,
                   const float *hh, float *out) {
    API_GUARD_BEGIN
    if (!s || x1 < x0 || y1 < y0) return SGN_E_ARG;
    // a 5/3 level one sample wide or high at an odd origin halves that
    // sample (idwt.rs:726-731): such a level is lifted in f32 with the 5/3
    // steps, on the f32 inputs as given
    const bool f53 = reversible && ((x1 - x0 == 1 && (x0 & 1)) || (y1 - y0 == 1 && (y0 & 1)));
    if (f53) reversible = 0;
    const uint32_t w = x1 - x0, h = y1 - y0;
    if (!w || !h) return SGN_OK;
    HIP_TRY(hipSetDevice(s->device));
    const uint32_t wl = (x1 + 1) / 2 - (x0 + 1) / 2, wh = x1 / 2 - x0 / 2;
    const uint32_t hl_ = (y1 + 1) / 2 - (y0 + 1) / 2, hh_ = y1 / 2 - y0 / 2;
    const size_t n[4] = {size_t(wl) * hl_, size_t(wh) * hl_, size_t(wl) * hh_, size_t(wh) * hh_};
    const float *src[4] = {ll, hl, lh, hh};
    std::vector<uint32_t> host;
    LevelJob lj{};
    // 5/3 levels take integers: each f32 input is truncated toward zero,
    // NaN to 0, saturating at the int32 range; 9/7 levels take the f32 bits
    auto to_int = [](float f) -> int32_t {
        if (!(f == f)) return 0;
        if (f >= 2147483647.0f) return 2147483647;
        if (f <= -2147483648.0f) return -2147483647 - 1;
        return static_cast<int32_t>(f);
    };
    for (int b = 0; b < 4; ++b) {
        lj.band[b] = static_cast<uint32_t>(host.size());
        lj.bw[b] = (b & 1) ? wh : wl;
        lj.delta[b] = 1.0f;
        for (size_t i = 0; i < n[b]; ++i) {
            uint32_t bits;
            if (reversible)
                bits = static_cast<uint32_t>(to_int(src[b][i]));
            else
                std::memcpy(&bits, &src[b][i], 4);
            host.push_back(bits);
        }
    }
    lj.float_in = reversible ? 0 : 0xF;
    lj.out = static_cast<uint32_t>(host.size());
    const uint32_t opitch = coef_pitch(w);  // the kernel writes rows at this pitch
    host.resize(host.size() + size_t(opitch) * h, 0);
    lj.x0 = x0;
    lj.y0 = y0;
    lj.x1 = x1;
    lj.y1 = y1;
    lj.nbx = static_cast<uint16_t>(ceil_div(w, 64));
    lj.nby = static_cast<uint16_t>(ceil_div(h, 64));
    lj.reversible = reversible ? 1 : 0;
    lj.lift53 = f53 ? 1 : 0;
    lj.levels = 1;
    uint32_t *d_coef = nullptr;
    LevelJob *d_job = nullptr;
    HIP_TRY(hipMalloc(&d_coef, host.size() * 4));
    HIP_TRY(hipMemcpy(d_coef, host.data(), host.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&d_job, sizeof(LevelJob)));
    HIP_TRY(hipMemcpy(d_job, &lj, sizeof(LevelJob), hipMemcpyHostToDevice));
    const uint32_t blocks = uint32_t(lj.nbx) * lj.nby;
    if (reversible)
        hipLaunchKernelGGL(sgn_idwt_level_rev, dim3(blocks), dim3(kIdwtThreads), 0, s->stream,
                           d_job, 1u, 0u, static_cast<const uint32_t *>(nullptr), d_coef);
    else
        hipLaunchKernelGGL(sgn_idwt_level_irr, dim3(blocks), dim3(kIdwtThreads), 0, s->stream,
                           d_job, 1u, 0u, static_cast<const uint32_t *>(nullptr), d_coef);
    s->dispatches++;
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<uint32_t> res(size_t(w) * h);
    HIP_TRY(hipMemcpy2D(res.data(), size_t(w) * 4, d_coef + lj.out, size_t(opitch) * 4,
                        size_t(w) * 4, h, hipMemcpyDeviceToHost));
    (void)hipFree(d_coef);
    (void)hipFree(d_job);
    for (size_t i = 0; i < res.size(); ++i) {
        if (reversible)
            out[i] = static_cast<float>(static_cast<int32_t>(res[i]));
        else
            std::memcpy(&out[i], &res[i], 4);
    }
    return SGN_OK;
    API_GUARD_END
}

int sgn_ht_decode_sub_band(sgn_hip_session *s, const uint8_t *data, size_t data_len,
                           uint32_t width, uint32_t height, const sgn_ht_code_block_batch_job *jobs,
                           size_t n, float *out, size_t out_len) {
    API_GUARD_BEGIN
    if (!s || (!jobs && n) || (!out && out_len)) return SGN_E_ARG;
    const uint64_t band = uint64_t(width) * height;
    if (out_len < band) {
        set_last_error("sub-band output smaller than width * height");
        return SGN_E_BUFFER;
    }
    // each block at (output_x, output_y) of the band, rows output_stride
    // apart (the band width, decode.rs:1622); blocks must lie in the band
    std::vector<sgn_ht_code_block_job> cbs(n);
    for (size_t i = 0; i < n; ++i) {
        const sgn_ht_code_block_batch_job &b = jobs[i];
        sgn_ht_code_block_job j = b.code_block;
        if (uint64_t(b.output_x) + j.width > width || uint64_t(b.output_y) + j.height > height ||
            j.output_stride < j.width) {
            set_last_error("code-block " + std::to_string(i) + " outside its sub-band");
            return SGN_E_ARG;
        }
        const uint64_t at = uint64_t(b.output_y) * j.output_stride + b.output_x;
        if (j.height && at + uint64_t(j.output_stride) * (j.height - 1) + j.width > out_len) {
            set_last_error("code-block " + std::to_string(i) + " outside the output");
            return SGN_E_BUFFER;
        }
        j.output_offset = static_cast<uint32_t>(at);
        cbs[i] = j;
    }
    std::vector<float> tmp(out_len, 0.0f);
    const int rc = sgn_ht_decode_code_blocks(s, data, data_len, cbs.data(), n, tmp.data(), out_len);
    if (rc) return rc;
    // samples of the blocks replace the caller's; the rest is left alone
    for (const sgn_ht_code_block_job &j : cbs)
        for (uint32_t y = 0; y < j.height; ++y)
            std::memcpy(out + j.output_offset + size_t(y) * j.output_stride,
                        tmp.data() + j.output_offset + size_t(y) * j.output_stride,
                        size_t(j.width) * sizeof(float));
    return SGN_OK;
    API_GUARD_END
}

int sgn_store_component(sgn_hip_session *s, const float *input, size_t input_len,
                        uint32_t input_width, uint32_t source_x, uint32_t source_y,
                        uint32_t copy_width, uint32_t copy_height, float *output,
                        size_t output_len, uint32_t output_width, uint32_t output_x,
                        uint32_t output_y, float addend) {
    API_GUARD_BEGIN
    if (!s || (!input && input_len) || (!output && output_len)) return SGN_E_ARG;
    if (!copy_width || !copy_height) return SGN_OK;
    const uint64_t in_end = (uint64_t(source_y) + copy_height - 1) * input_width + source_x +
                            copy_width;
    const uint64_t out_end = (uint64_t(output_y) + copy_height - 1) * output_width + output_x +
                             copy_width;
    if (uint64_t(source_x) + copy_width > input_width || in_end > input_len) {
        set_last_error("store source rectangle outside the input");
        return SGN_E_ARG;
    }
    if (uint64_t(output_x) + copy_width > output_width || out_end > output_len) {
        set_last_error("store destination rectangle outside the output plane");
        return SGN_E_BUFFER;
    }
    HIP_TRY(hipSetDevice(s->device));
    float *d_in = nullptr, *d_out = nullptr;
    auto cleanup = [&]() {
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
    };
    if (hipMalloc(&d_in, input_len * 4) != hipSuccess ||
        hipMalloc(&d_out, output_len * 4) != hipSuccess) {
        cleanup();
        set_last_error("hipMalloc failed");
        return SGN_E_DEVICE;
    }
    hipError_t e = hipMemcpy(d_in, input, input_len * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_out, output, output_len * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const dim3 grid((copy_width + 255) / 256, std::min<uint32_t>(copy_height, 65535));
        hipLaunchKernelGGL(sgn_store_plane, grid, dim3(256), 0, s->stream, d_in, input_width,
                           source_x, source_y, copy_width, copy_height, d_out, output_width,
                           output_x, output_y, addend);
        s->dispatches++;
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(output, d_out, output_len * 4, hipMemcpyDeviceToHost);
    cleanup();
    if (e != hipSuccess) {
        set_last_error(std::string("HIP: ") + hipGetErrorString(e) + " in the store hook");
        return SGN_E_DEVICE;
    }
    return SGN_OK;
    API_GUARD_END
}

int sgn_inverse_mct(sgn_hip_session *s, int reversible, float *p0, float *p1, float *p2, size_t n,
                    float a0, float a1, float a2) {
    API_GUARD_BEGIN
    if (!s || (n && (!p0 || !p1 || !p2))) return SGN_E_ARG;
    if (!n) return SGN_OK;
    HIP_TRY(hipSetDevice(s->device));
    float *d = nullptr;
    HIP_TRY(hipMalloc(&d, n * 12));
    HIP_TRY(hipMemcpy(d, p0, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d + n, p1, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d + 2 * n, p2, n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(sgn_inverse_mct_planes, dim3((n + 255) / 256), dim3(256), 0, s->stream, d,
                       d + n, d + 2 * n, uint32_t(n), reversible, a0, a1, a2);
    s->dispatches++;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(p0, d, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(p1, d + n, n * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(p2, d + 2 * n, n * 4, hipMemcpyDeviceToHost));
    (void)hipFree(d);
    return SGN_OK;
    API_GUARD_END
}

}  // extern "C"

// ---- decode-cost model ---------------------------------------------------

namespace {

// He-uniform initial weights, zero biases, from a splitmix64 stream.
std::vector<float> mlp_init(uint64_t seed) {
    std::vector<float> p(kMlpParams, 0.0f);
    uint64_t x = seed;
    auto next = [&]() {
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return double(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int count, int fan_in) {
        const double a = std::sqrt(6.0 / fan_in);
        for (int k = 0; k < count; ++k) p[off + k] = float((2.0 * next() - 1.0) * a);
    };
    fill(kW1, kMlpH * kMlpIn, kMlpIn);
    fill(kW2, kMlpH * kMlpH, kMlpH);
    fill(kW3, kMlpH, kMlpH);
    return p;
}

// Device buffer that frees itself.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    Status alloc(size_t n) {
        if (hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) {
            p = nullptr;
            set_last_error("hipMalloc failed");
            return kErrDevice;
        }
        return kOk;
    }
    Status upload(const T *h, size_t n) {
        Status st = alloc(n);
        if (st) return st;
        if (n && hipMemcpy(p, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
            set_last_error("hipMemcpy failed");
            return kErrDevice;
        }
        return kOk;
    }
};

// Enqueues the mean-loss gradient of the model on (X, Y) into d_grad.
Status mlp_gradient(sgn_hip_session *s, const float *d_params, const float *d_x, const float *d_y,
                    uint32_t n, float *d_partial, float *d_grad) {
    const uint32_t groups = (n + kGradWg - 1) / kGradWg;
    hipLaunchKernelGGL(sgn_mlp_grad_partial, dim3(groups), dim3(kGradWg), 0, s->stream, d_params,
                       d_x, d_y, n, d_partial);
    hipLaunchKernelGGL(sgn_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, s->stream,
                       d_partial, groups, 1.0f / float(n), d_grad);
    s->dispatches += 2;
    HIP_TRY(hipGetLastError());
    return kOk;
}

}  // namespace

int sgn_cost_model_create(sgn_hip_session *s, uint64_t seed, sgn_cost_model **out) {
    API_GUARD_BEGIN
    if (!s || !out) return SGN_E_ARG;
    *out = nullptr;
    HIP_TRY(hipSetDevice(s->device));
    auto m = std::make_unique<sgn_cost_model>();
    m->device = s->device;
    const std::vector<float> p = mlp_init(seed);
    const std::vector<float> z(kMlpParams, 0.0f);
    HIP_TRY(hipMalloc(&m->d_params, kMlpParams * sizeof(float)));
    HIP_TRY(hipMalloc(&m->d_m, kMlpParams * sizeof(float)));
    HIP_TRY(hipMalloc(&m->d_v, kMlpParams * sizeof(float)));
    HIP_TRY(hipMemcpy(m->d_params, p.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->d_m, z.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m->d_v, z.data(), kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    *out = m.release();
    return SGN_OK;
    API_GUARD_END
}

void sgn_cost_model_destroy(sgn_cost_model *m) { delete m; }

int sgn_cost_model_param_count(void) { return kMlpParams; }

int sgn_cost_model_get_params(const sgn_cost_model *m, float *out, size_t n) {
    if (!m || !out || n < size_t(kMlpParams)) return SGN_E_ARG;
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipMemcpy(out, m->d_params, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return SGN_OK;
}

int sgn_cost_model_set_params(sgn_cost_model *m, const float *p, size_t n) {
    if (!m || !p || n < size_t(kMlpParams)) return SGN_E_ARG;
    HIP_TRY(hipSetDevice(m->device));
    HIP_TRY(hipMemcpy(m->d_params, p, kMlpParams * sizeof(float), hipMemcpyHostToDevice));
    return SGN_OK;
}

int sgn_cost_model_gradient(sgn_hip_session *s, const sgn_cost_model *m, const float *X,
                            const float *Y, size_t n, float *grad) {
    API_GUARD_BEGIN
    if (!s || !m || !X || !Y || !grad || n == 0 || n > (size_t(1) << 24)) return SGN_E_ARG;
    HIP_TRY(hipSetDevice(s->device));
    DevBuf<float> dx, dy, dpart, dgrad;
    Status st;
    if ((st = dx.upload(X, n * kMlpIn)) || (st = dy.upload(Y, n)) ||
        (st = dpart.alloc(((n + kGradWg - 1) / kGradWg) * kMlpParams)) ||
        (st = dgrad.alloc(kMlpParams)))
        return st;
    if ((st = mlp_gradient(s, m->d_params, dx.p, dy.p, uint32_t(n), dpart.p, dgrad.p))) return st;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(grad, dgrad.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return SGN_OK;
    API_GUARD_END
}

int sgn_cost_model_train(sgn_hip_session *s, sgn_cost_model *m, const float *X, const float *Y,
                         size_t n, uint32_t epochs, float lr, float *loss) {
    API_GUARD_BEGIN
    if (!s || !m || !X || !Y || n == 0 || n > (size_t(1) << 24) || !(lr > 0.0f))
        return SGN_E_ARG;
    HIP_TRY(hipSetDevice(s->device));
    DevBuf<float> dx, dy, dpart, dgrad, dpred;
    Status st;
    if ((st = dx.upload(X, n * kMlpIn)) || (st = dy.upload(Y, n)) ||
        (st = dpart.alloc(((n + kGradWg - 1) / kGradWg) * kMlpParams)) ||
        (st = dgrad.alloc(kMlpParams)) || (st = dpred.alloc(n)))
        return st;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t e = 0; e < epochs; ++e) {
        if ((st = mlp_gradient(s, m->d_params, dx.p, dy.p, uint32_t(n), dpart.p, dgrad.p)))
            return st;
        ++m->steps;
        const float c1 = 1.0f - std::pow(b1, float(m->steps));
        const float c2 = 1.0f - std::pow(b2, float(m->steps));
        hipLaunchKernelGGL(sgn_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s->stream,
                           m->d_params, dgrad.p, m->d_m, m->d_v, lr, b1, b2, eps, c1, c2);
        s->dispatches++;
    }
    HIP_TRY(hipGetLastError());
    if (loss) {
        hipLaunchKernelGGL(sgn_mlp_predict, dim3(uint32_t((n + 255) / 256)), dim3(256), 0,
                           s->stream, m->d_params, dx.p, uint32_t(n), dpred.p);
        s->dispatches++;
        std::vector<float> pred(n);
        HIP_TRY(hipStreamSynchronize(s->stream));
        HIP_TRY(hipMemcpy(pred.data(), dpred.p, n * sizeof(float), hipMemcpyDeviceToHost));
        double acc = 0;
        for (size_t i = 0; i < n; ++i) acc += 0.5 * double(pred[i] - Y[i]) * (pred[i] - Y[i]);
        *loss = float(acc / double(n));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return SGN_OK;
    API_GUARD_END
}

int sgn_cost_model_predict(sgn_hip_session *s, const sgn_cost_model *m, const float *X, size_t n,
                           float *out) {
    API_GUARD_BEGIN
    if (!s || !m || (!X && n) || (!out && n) || n > (size_t(1) << 24)) return SGN_E_ARG;
    if (!n) return SGN_OK;
    HIP_TRY(hipSetDevice(s->device));
    DevBuf<float> dx, dpred;
    Status st;
    if ((st = dx.upload(X, n * kMlpIn)) || (st = dpred.alloc(n))) return st;
    hipLaunchKernelGGL(sgn_mlp_predict, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, s->stream,
                       m->d_params, dx.p, uint32_t(n), dpred.p);
    s->dispatches++;
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, dpred.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return SGN_OK;
    API_GUARD_END
}

int sgn_j2k_plan_time_code_blocks(sgn_hip_session *s, sgn_j2k_plan *P, size_t stride,
                                  size_t max_samples, float *features, float *cycles,
                                  size_t *count) {
    API_GUARD_BEGIN
    if (!s || !P || !stride || !count) return SGN_E_ARG;
    if (P->device_packets) {
        set_last_error("block timing needs the host's code-block lengths (no device packets)");
        return SGN_E_UNSUPPORTED;
    }
    HIP_TRY(hipSetDevice(s->device));
    std::vector<size_t> pick;
    for (size_t i = 0; i < P->cb_lane.size() && pick.size() < max_samples; i += stride)
        if (!(P->cb_lane[i].flags & kCbZero)) pick.push_back(i);
    *count = pick.size();
    if (pick.empty()) return SGN_OK;
    // wave k: lane 0 holds sample k, lanes 1..63 an empty block
    CbJob empty{};
    empty.flags = kCbZero;
    std::vector<CbJob> jobs(pick.size() * 64, empty);
    for (size_t k = 0; k < pick.size(); ++k) jobs[k * 64] = P->cb_lane[pick[k]];
    DevBuf<CbJob> djobs;
    DevBuf<uint8_t> dst;
    DevBuf<uint64_t> dcyc;
    Status st;
    if ((st = djobs.upload(jobs.data(), jobs.size())) || (st = dst.alloc(jobs.size())) ||
        (st = dcyc.alloc(pick.size())))
        return st;
    const uint32_t waves = uint32_t(pick.size());
    if (P->narrow)
        hipLaunchKernelGGL(sgn_ht_rows_timed<int16_t>, dim3((waves + 3) / 4), dim3(256), 0,
                           s->stream, djobs.p, uint32_t(jobs.size()), P->d_arena,
                           reinterpret_cast<int16_t *>(P->d_coef), s->d_tables, dst.p, dcyc.p);
    else
        hipLaunchKernelGGL(sgn_ht_rows_timed<int32_t>, dim3((waves + 3) / 4), dim3(256), 0,
                           s->stream, djobs.p, uint32_t(jobs.size()), P->d_arena,
                           reinterpret_cast<int32_t *>(P->d_coef), s->d_tables, dst.p, dcyc.p);
    s->dispatches++;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<uint64_t> cyc(pick.size());
    HIP_TRY(hipMemcpy(cyc.data(), dcyc.p, cyc.size() * 8, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < pick.size(); ++k) {
        if (cycles) cycles[k] = float(cyc[k]);
        if (features) {
            const CbFeatures f = cb_features(P->cb_lane[pick[k]]);
            for (int q = 0; q < kMlpIn; ++q) features[k * kMlpIn + q] = f.f[q];
        }
    }
    return SGN_OK;
    API_GUARD_END
}

int sgn_j2k_plan_order_code_blocks(sgn_hip_session *s, sgn_j2k_plan *P, int order,
                                   const sgn_cost_model *m) {
    API_GUARD_BEGIN
    const bool model = order == SGN_ORDER_MODEL || order == SGN_ORDER_GEOMETRY_MODEL;
    if (!s || !P || order < 0 || order > SGN_ORDER_GEOMETRY_MODEL || (model && !m))
        return SGN_E_ARG;
    if (P->device_packets) {
        set_last_error("the code-block order of a plan with device packet parsing is fixed");
        return SGN_E_UNSUPPORTED;
    }
    HIP_TRY(hipSetDevice(s->device));
    std::vector<CbJob> v = P->cb_lane_plain;
    // per-block cost key: the model's prediction or the length heuristic
    std::vector<float> key(v.size());
    if (model && !v.empty()) {
        DevBuf<float> dpred;
        Status st;
        if ((st = dpred.alloc(v.size()))) return st;
        hipLaunchKernelGGL(sgn_mlp_predict_blocks, dim3(uint32_t((v.size() + 255) / 256)),
                           dim3(256), 0, s->stream, m->d_params, P->d_cb_lane_plain,
                           uint32_t(v.size()), dpred.p);
        s->dispatches++;
        HIP_TRY(hipStreamSynchronize(s->stream));
        HIP_TRY(hipMemcpy(key.data(), dpred.p, key.size() * 4, hipMemcpyDeviceToHost));
    } else {
        for (size_t i = 0; i < v.size(); ++i) key[i] = float(v[i].cost);
    }
    std::vector<uint32_t> idx(v.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = uint32_t(i);
    auto area = [&](uint32_t i) { return uint32_t(v[i].w) * v[i].h; };
    if (order == SGN_ORDER_LENGTH || order == SGN_ORDER_MODEL) {
        std::stable_sort(idx.begin(), idx.end(),
                         [&](uint32_t a, uint32_t b) { return key[a] > key[b]; });
    } else if (order == SGN_ORDER_GEOMETRY) {
        // equal block shapes together, tile order kept inside each shape:
        // a wave's lanes get equal-sized blocks of neighbouring tiles
        std::stable_sort(idx.begin(), idx.end(),
                         [&](uint32_t a, uint32_t b) { return area(a) > area(b); });
    } else if (order != SGN_ORDER_INPUT) {
        // shapes as above; inside a shape, tiles by their summed block cost
        // (a tile's blocks stay together, so neighbouring lanes still read
        // and write neighbouring memory)
        std::vector<double> tile_cost(P->ntiles, 0.0);
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i].tile < tile_cost.size()) tile_cost[v[i].tile] += key[i];
        auto tc = [&](uint32_t i) { return v[i].tile < tile_cost.size() ? tile_cost[v[i].tile] : 0.0; };
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
            if (area(a) != area(b)) return area(a) > area(b);
            return tc(a) > tc(b);
        });
    }
    std::vector<CbJob> w(v.size());
    for (size_t i = 0; i < idx.size(); ++i) w[i] = v[idx[i]];
    v.swap(w);
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (!v.empty())
        HIP_TRY(hipMemcpy(P->d_cb_lane, v.data(), v.size() * sizeof(CbJob), hipMemcpyHostToDevice));
    P->cb_lane.swap(v);
    drop_graph(*P);
    return SGN_OK;
    API_GUARD_END
}

int sgn_j2k_plan_autotune_order(sgn_hip_session *s, sgn_j2k_plan *P, int reps, float *ms,
                                int *chosen) {
    API_GUARD_BEGIN
    if (!s || !P || reps < 1) return SGN_E_ARG;
    static const int kOrders[4] = {SGN_ORDER_LENGTH, SGN_ORDER_INPUT, SGN_ORDER_GEOMETRY,
                                   SGN_ORDER_GEOMETRY_COST};
    int best_order = -1, best_variant = 0;
    float best_ms = 0.0f;
    for (int k = 0; k < 4; ++k) {
        int rc = sgn_j2k_plan_order_code_blocks(s, P, kOrders[k], nullptr);
        if (rc) return rc;
        std::vector<float> t(kRowsVariantCount, 0.0f);
        int v = 0;
        rc = sgn_j2k_plan_autotune(s, P, reps, t.data(), t.size(), &v);
        if (rc) return rc;
        if (ms) ms[k] = t[v];
        if (best_order < 0 || t[v] < best_ms) {
            best_order = kOrders[k];
            best_variant = v;
            best_ms = t[v];
        }
    }
    int rc = sgn_j2k_plan_order_code_blocks(s, P, best_order, nullptr);
    if (rc) return rc;
    if (P->lane_variant != best_variant) drop_graph(*P);
    P->lane_variant = best_variant;
    if (chosen) *chosen = best_order;
    return SGN_OK;
    API_GUARD_END
}

#include "server.hip"
