// HT packet headers parsed on the device: one tile per wavefront.
//
// A tile's packet headers are one serial bit stream per tile-part (each
// header's tag-tree and length fields depend on everything read before
// it), so the parallelism is across tiles: one work-group (one wave) per
// tile, the wave clearing the tile's trees and jobs, lane 0 walking the
// headers (64 tiles on the lanes of one wave diverged on every branch).  The walk follows the host
// parser (host/parse.cpp parse_packets, which restates segment.rs:36-133
// and 508-581): per packet an optional SOP marker, the non-empty bit, per
// precinct-band and code-block the inclusion tag tree (first layer), the
// zero-bit-plane tag tree, the HT pass-count codeword with placeholder
// passes, Lblock increments and the cleanup / refinement lengths, byte
// alignment and an optional EPH marker; then the body assigns each
// included block its bytes in header order.  A malformed packet ends its
// tile-part and the next tile-part continues (non-strict, segment.rs:
// 36-50).  The results land in the plan's code-block jobs, which the host
// created as zero blocks with their geometry.
//
// The including file provides the HIP definitions (__global__, __device__,
// blockIdx, threadIdx, min).
#pragma once
#include <stdint.h>

#include "../host/plan_types.h"

namespace sgn {

// Packet-header bits of one tile-part with the B.10.1 stuffing rule.
struct PktBits {
    const uint8_t *d;
    uint32_t len;
    uint32_t bit;
    __device__ bool get(uint32_t &out) {
        const uint32_t byte = bit >> 3;
        if (byte >= len) return false;
        const bool stuff = ((bit & 7u) == 7u) && d[byte] == 0xFF;
        out = (d[byte] >> (7u - (bit & 7u))) & 1u;
        ++bit;
        if (stuff) {
            const uint32_t b2 = bit >> 3;
            if (b2 >= len) return false;
            const uint32_t sb = (d[b2] >> (7u - (bit & 7u))) & 1u;
            ++bit;
            if (sb) return false;
        }
        return true;
    }
    __device__ bool bits(uint32_t n, uint32_t &v) {
        v = 0;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t b;
            if (!get(b)) return false;
            v = (v << 1) | b;
        }
        return true;
    }
};

// Tag-tree decode against `threshold` (B.10.2) over the nodes of one tree:
// (value, low) pairs, levels from the w x h leaves up to the root.
// Returns 1 when the leaf's value is below the threshold, 0 when not, -1
// when the stream ends.
__device__ int pkt_tag_decode(uint32_t *nodes, uint32_t w, uint32_t h, uint32_t x, uint32_t y,
                              uint32_t threshold, PktBits &r, uint32_t &value) {
    uint32_t lw[20], off[20], levels = 0;
    {
        uint32_t cw = w, ch = h, total = 0;
        for (;;) {
            lw[levels] = cw;
            off[levels] = total;
            total += cw * ch;
            ++levels;
            if ((cw == 1 && ch == 1) || levels == 20) break;
            cw = (cw + 1) / 2;
            ch = (ch + 1) / 2;
        }
    }
    uint32_t low = 0;
    for (int l = int(levels) - 1; l >= 0; --l) {
        const uint32_t n = off[l] + (y >> l) * lw[l] + (x >> l);
        uint32_t &nv = nodes[2 * n], &nl = nodes[2 * n + 1];
        if (low > nl)
            nl = low;
        else
            low = nl;
        while (low < threshold && low < nv) {
            uint32_t b;
            if (!r.get(b)) return -1;
            if (b)
                nv = low;
            else
                ++low;
        }
        nl = low;
    }
    value = nodes[0 + 2 * (off[0] + y * lw[0] + x)];
    return value < threshold ? 1 : 0;
}

__device__ __forceinline__ CbJob *pkt_job(const PktCb &cb, CbJob *lanes, CbJob *wave) {
    if (cb.job == kPktNoJob) return nullptr;
    return (cb.job & kPktWave) ? wave + (cb.job & ~kPktWave) : lanes + cb.lane_pos;
}

// Tiles tiles[0 .. n), one per 64-thread work-group.  lanes / wave: the
// plan's code-block lists (a block's place in `lanes` is PktCb::lane_pos);
// lane2: the per-tile lane list of the plain mode (at PktCb::job) or null.  nodes: tag-tree scratch (two words per node);
// lens: one word per PktCb (Lcup | Lref << 16 of an included block, 0
// otherwise: blocks without a job still have bytes in the body).
// status[t] = 1 for a tile whose header asks for more than 31 bit-planes.
__global__ void __launch_bounds__(64)
sgn_packet_headers(const PktTile *__restrict__ tiles, uint32_t n,
                   const PktPart *__restrict__ parts, const PktPacket *__restrict__ pkts,
                   const PktBand *__restrict__ bands, const PktCb *__restrict__ cbs,
                   uint32_t *__restrict__ nodes, uint32_t *__restrict__ lens,
                   const uint8_t *__restrict__ arena,
                   CbJob *__restrict__ lanes, CbJob *__restrict__ lane2, CbJob *__restrict__ wave,
                   uint8_t *__restrict__ status) {
    const uint32_t i = blockIdx.x, lane = threadIdx.x;
    if (i >= n) return;
    const PktTile T = tiles[i];
    uint8_t bad = 0;
    // fresh trees and zero blocks on every run (plans are replayed)
    for (uint32_t k = lane; k < T.nnodes; k += 64) {
        nodes[2 * (T.node0 + k)] = 0xFFFFFFFFu;
        nodes[2 * (T.node0 + k) + 1] = 0;
    }
    for (uint32_t k = lane; k < T.ncbs; k += 64) {
        lens[T.cb0 + k] = 0;
        CbJob *j = pkt_job(cbs[T.cb0 + k], lanes, wave);
        if (!j) continue;
        j->src = 0;
        j->lcup = j->lref = 0;
        j->zblk = 0;
        j->passes = 0;
        j->flags |= kCbZero;
    }
    __syncthreads();
    uint32_t next = 0;  // next packet of the progression
    for (uint32_t pi = 0; pi < T.nparts && lane == 0; ++pi) {
        const PktPart part = parts[T.part0 + pi];
        const uint8_t *body = arena + part.off;
        const uint32_t blen = part.len;
        PktBits r{body, blen, 0};
        for (;;) {
            if ((r.bit >> 3) >= r.len || next >= T.npkts) break;
            const PktPacket P = pkts[T.pkt0 + next++];
            bool ok = true;
            {
                const uint32_t at = r.bit >> 3;
                if (P.sop && at + 1 < blen && body[at] == 0xFF && body[at + 1] == 0x91) {
                    if (at + 6 > blen) break;
                    r.bit += 48;
                }
            }
            uint32_t nonzero;
            if (!r.get(nonzero)) break;
            if (nonzero) {
                for (uint32_t bi = 0; bi < P.nbands && ok; ++bi) {
                    const PktBand B = bands[P.band0 + bi];
                    for (uint32_t k = 0; k < uint32_t(B.ncbx) * B.ncby && ok; ++k) {
                        const uint32_t x = k % B.ncbx, y = k / B.ncbx;
                        const PktCb cb = cbs[B.cb0 + k];
                        uint32_t v;
                        int rc = pkt_tag_decode(nodes + 2 * B.incl, B.ncbx, B.ncby, x, y, 1u, r, v);
                        if (rc < 0) {
                            ok = false;
                            break;
                        }
                        if (v > 0) continue;  // not included in the first layer
                        uint32_t thr = 1;
                        for (;;) {
                            rc = pkt_tag_decode(nodes + 2 * B.zbp, B.ncbx, B.ncby, x, y, thr, r, v);
                            if (rc < 0 || (!rc && ++thr > 255)) {
                                ok = false;
                                break;
                            }
                            if (rc) break;
                        }
                        if (!ok) break;
                        uint32_t zblk = v;
                        // HT pass count codeword (segment.rs:508-530)
                        uint32_t np = 1, b1;
                        if (!r.get(b1)) {
                            ok = false;
                            break;
                        }
                        if (b1) {
                            np = 2;
                            if (!r.get(b1)) {
                                ok = false;
                                break;
                            }
                            if (b1) {
                                uint32_t e;
                                if (!r.bits(2, e)) {
                                    ok = false;
                                    break;
                                }
                                np = 3 + e;
                                if (e == 3) {
                                    if (!r.bits(5, e)) {
                                        ok = false;
                                        break;
                                    }
                                    np = 6 + e;
                                    if (e == 31) {
                                        if (!r.bits(7, e)) {
                                            ok = false;
                                            break;
                                        }
                                        np = 37 + e;
                                    }
                                }
                            }
                        }
                        if (np > 94) {
                            ok = false;
                            break;
                        }
                        // placeholder passes and lengths (segment.rs:542-581)
                        const uint32_t groups = (np - 1) / 3;
                        zblk += groups;
                        const uint32_t ph = groups * 3, actual = np - ph;
                        uint32_t lblock = 3;
                        for (;;) {
                            uint32_t bb;
                            if (!r.get(bb)) {
                                ok = false;
                                break;
                            }
                            if (!bb) break;
                            ++lblock;
                        }
                        if (!ok) break;
                        const uint32_t nb = lblock + (31u - __builtin_clz(ph + 1));
                        uint32_t lc = 0, lr = 0;
                        if (nb > 32 || !r.bits(nb, lc) || lc < 2 || lc >= 65535) {
                            ok = false;
                            break;
                        }
                        if (actual > 1) {
                            const uint32_t nb2 = lblock + (actual > 2 ? 1u : 0u);
                            if (nb2 > 32 || !r.bits(nb2, lr) || lr >= 2047) {
                                ok = false;
                                break;
                            }
                        }
                        CbJob *j = pkt_job(cb, lanes, wave);
                        if (cb.kmax > 31) bad = 1;
                        lens[B.cb0 + k] = lc | (lr << 16);
                        if (j) {
                            // lengths now, the data offset with the body below
                            j->lcup = (uint16_t)lc;
                            j->lref = (uint16_t)lr;
                            j->zblk = (uint8_t)min(zblk, 255u);
                            j->passes = (uint8_t)actual;
                        }
                    }
                }
            }
            if (!ok) break;
            if (r.bit & 7u) r.bit += 8 - (r.bit & 7u);
            if (P.eph) {
                const uint32_t bb = r.bit >> 3;
                if (bb + 2 > r.len || body[bb] != 0xFF || body[bb + 1] != 0x92) break;
                r.bit += 16;
            }
            uint32_t bpos = r.bit >> 3;
            if (nonzero) {
                for (uint32_t bi = 0; bi < P.nbands && ok; ++bi) {
                    const PktBand B = bands[P.band0 + bi];
                    for (uint32_t k = 0; k < uint32_t(B.ncbx) * B.ncby; ++k) {
                        const uint32_t L = lens[B.cb0 + k];
                        if (!L) continue;
                        const uint32_t need = (L & 0xFFFFu) + (L >> 16);
                        if (bpos + need > blen) {
                            ok = false;
                            break;
                        }
                        CbJob *j = pkt_job(cbs[B.cb0 + k], lanes, wave);
                        if (!j) {
                            bpos += need;  // a block of a skipped resolution or outside the region
                            continue;
                        }
                        const uint32_t kmax = cbs[B.cb0 + k].kmax;
                        const uint32_t act = kmax > j->zblk ? kmax - j->zblk : 0;
                        j->src = part.off + bpos;
                        if (kmax != 0 && act != 0) j->flags &= (uint8_t)~kCbZero;
                        bpos += need;
                    }
                }
            }
            if (!ok) break;
            r.bit = bpos * 8;
        }
    }
    __syncthreads();
    // the per-tile copy of the lane list for the plain mode
    if (lane2)
        for (uint32_t k = lane; k < T.ncbs; k += 64) {
            const PktCb cb = cbs[T.cb0 + k];
            if (cb.job != kPktNoJob && !(cb.job & kPktWave)) lane2[cb.job] = lanes[cb.lane_pos];
        }
    if (lane == 0) status[i] = bad;
}

}  // namespace sgn
