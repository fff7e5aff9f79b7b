// HTJ2K encoder: the input generator of the decode path.
//
// Lossless mode restates the reference's HT cleanup-only encoder
// (crates/signinum-j2k-native/src/j2c/ht_block_encode.rs:85-317,651-1010,
// 1409-1612): one cleanup pass per code-block with missing MSBs = K - 1,
// MEL / VLC / MagSgn streams merged into one cleanup segment.  Codestream
// layout follows the reference's writer (codestream_write.rs:60-200):
// SOC, SIZ, CAP, COD, QCD, then one tile-part per tile and EOC.
//
// The irreversible mode (9/7 + ICT + scalar-expounded quantisation) has no
// counterpart in the reference; it produces the lossy test inputs.  Step
// sizes are base * 2^P / ||g_b|| with ||g_b|| the L2 norm of the band's 9/7
// synthesis basis function, and two guard bits.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "j2k_host.hpp"

namespace sgn {

namespace {

constexpr int kMelExp[13] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 4, 5};

// MEL: adaptive run-length coder, bits MSB-first with 0xFF stuffing.
struct MelWriter {
    std::vector<uint8_t> buf;
    uint32_t acc = 0;
    int nbits = 0, maxbits = 8;
    int k = 0;
    uint32_t run = 0;
    void bit(uint32_t b) {
        acc = (acc << 1) | b;
        if (++nbits == maxbits) {
            buf.push_back(static_cast<uint8_t>(acc));
            maxbits = acc == 0xFF ? 7 : 8;
            acc = 0;
            nbits = 0;
        }
    }
    void encode(bool sym) {
        if (!sym) {
            if (++run >= (1u << kMelExp[k])) {
                bit(1);
                run = 0;
                k = std::min(k + 1, 12);
            }
        } else {
            bit(0);
            for (int t = kMelExp[k] - 1; t >= 0; --t) bit((run >> t) & 1);
            run = 0;
            k = std::max(k - 1, 0);
        }
    }
};

// VLC: written backward from the end of the segment, bits LSB-first; a
// byte that follows (in backward order) a byte above 0x8F holds only 7 bits
// when those 7 bits are all ones.  The first byte reserves its low nibble
// for the suffix-length field.
struct VlcWriter {
    std::vector<uint8_t> rev;  // bytes in backward order; rev[0] is the last byte
    uint32_t acc = 0x0F;
    int used = 4;
    bool prev_gt_8f = true;
    VlcWriter() { rev.push_back(0xFF); }
    void put(uint32_t cwd, int len) {
        while (len > 0) {
            int avail = 8 - (prev_gt_8f ? 1 : 0) - used;
            const int take = std::min(avail, len);
            acc |= (cwd & ((1u << take) - 1)) << used;
            used += take;
            avail -= take;
            len -= take;
            cwd >>= take;
            if (avail == 0) {
                if (prev_gt_8f && acc != 0x7F) {
                    prev_gt_8f = false;  // the 8th bit is usable after all
                    continue;
                }
                rev.push_back(static_cast<uint8_t>(acc));
                prev_gt_8f = acc > 0x8F;
                acc = 0;
                used = 0;
            }
        }
    }
};

// MagSgn: forward, LSB-first, 7 bits after a 0xFF byte.
struct MagSgnWriter {
    std::vector<uint8_t> buf;
    uint64_t acc = 0;
    int used = 0, maxbits = 8;
    void put(uint32_t v, int len) {
        while (len > 0) {
            const int take = std::min(maxbits - used, len);
            acc |= static_cast<uint64_t>(v & ((1u << take) - 1)) << used;
            used += take;
            v = take == 32 ? 0 : v >> take;
            len -= take;
            if (used >= maxbits) {
                buf.push_back(static_cast<uint8_t>(acc));
                maxbits = (acc & 0xFF) == 0xFF ? 7 : 8;
                acc = 0;
                used = 0;
            }
        }
    }
    void terminate() {
        if (used > 0) {
            const int unused = maxbits - used;
            acc |= static_cast<uint64_t>((1u << unused) - 1) << used;
            if ((acc & 0xFF) != 0xFF) buf.push_back(static_cast<uint8_t>(acc));
        } else if (maxbits == 7 && !buf.empty()) {
            buf.pop_back();
        }
    }
};

struct UvlcCode {
    uint32_t pre, pre_len, suf, suf_len;
};
UvlcCode uvlc_code(int u) {
    if (u <= 0) return {0, 0, 0, 0};
    if (u == 1) return {1, 1, 0, 0};
    if (u == 2) return {2, 2, 0, 0};
    if (u <= 4) return {4, 3, static_cast<uint32_t>(u - 3), 1};
    return {0, 3, static_cast<uint32_t>(u - 5), 5};
}

struct Quad {
    uint32_t rho = 0;
    int e[4] = {0, 0, 0, 0};
    uint32_t s[4] = {0, 0, 0, 0};
    int emax = 0;
};

inline int bitlen(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }

}  // namespace

// Cleanup-only HT encode of one code-block of signed integer coefficients.
// Returns the cleanup segment (empty when the block is all zero).
Status ht_encode_block(const int32_t *coef, uint32_t w, uint32_t h, size_t stride,
                       uint32_t k_max, std::vector<uint8_t> &out) {
    out.clear();
    const HtTables &T = ht_tables();
    uint32_t maxmag = 0;
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const int32_t v = coef[y * stride + x];
            const uint32_t m = static_cast<uint32_t>(v < 0 ? -static_cast<int64_t>(v) : v);
            maxmag = std::max(maxmag, m);
        }
    if (maxmag == 0) return kOk;
    if (k_max == 0 || k_max > 30 || bitlen(maxmag) > static_cast<int>(k_max)) {
        set_last_error("HT encode: magnitude exceeds the band's bit-plane count");
        return kErrArg;
    }
    MelWriter mel;
    VlcWriter vlc;
    MagSgnWriter ms;
    const uint32_t nq = (w + 1) / 2;
    std::vector<uint32_t> prev_rho(nq + 2, 0), cur_rho(nq + 2, 0);
    std::vector<int> prev_e(w + 4, 0), cur_e(w + 4, 0);  // E of the lower row, index col + 1

    auto load_quad = [&](uint32_t qx, uint32_t y, Quad &q) {
        q = Quad();
        for (int i = 0; i < 4; ++i) {
            const uint32_t x = 2 * qx + (i >> 1), yy = y + (i & 1);
            if (x >= w || yy >= h) continue;
            const int32_t v = coef[yy * stride + x];
            if (v == 0) continue;
            const uint32_t mu = static_cast<uint32_t>(v < 0 ? -static_cast<int64_t>(v) : v);
            q.rho |= 1u << i;
            q.e[i] = bitlen(2 * mu - 1);
            q.s[i] = 2 * (mu - 1) + (v < 0 ? 1u : 0u);
            q.emax = std::max(q.emax, q.e[i]);
        }
    };
    // Emits VLC codeword, MEL symbol and MagSgn bits of one quad; returns u.
    auto code_quad = [&](int table, uint32_t c_q, const Quad &q, int kappa, int &u_out) -> bool {
        const int uq = std::max(q.emax, kappa);
        const int u = uq - kappa;
        uint32_t emb = 0;
        if (u > 0)
            for (int i = 0; i < 4; ++i)
                if ((q.rho >> i & 1) && q.e[i] == uq) emb |= 1u << i;
        uint32_t ek = 0;
        if (!(c_q == 0 && q.rho == 0)) {
            const uint32_t t = T.enc[table][(c_q << 8) | (q.rho << 4) | emb];
            if (!t) return false;
            vlc.put(t & 0x7F, (t >> 8) & 7);
            ek = (t >> 12) & 0xF;
        }
        if (c_q == 0) mel.encode(q.rho != 0);
        for (int i = 0; i < 4; ++i) {
            if (!(q.rho >> i & 1)) continue;
            const int m = uq - static_cast<int>((ek >> i) & 1);
            if (m > 0) ms.put(q.s[i] & (m >= 32 ? 0xFFFFFFFFu : ((1u << m) - 1)), m);
        }
        u_out = u;
        return true;
    };

    for (uint32_t y = 0; y < h; y += 2) {
        const bool initial = (y == 0);
        std::fill(cur_rho.begin(), cur_rho.end(), 0);
        std::fill(cur_e.begin(), cur_e.end(), 0);
        uint32_t rho_w = 0;  // rho of the quad to the left
        for (uint32_t qx = 0; qx < nq; qx += 2) {
            Quad q[2];
            const bool has1 = qx + 1 < nq;
            load_quad(qx, y, q[0]);
            if (has1) load_quad(qx + 1, y, q[1]);
            int u[2] = {0, 0};
            for (int k = 0; k < (has1 ? 2 : 1); ++k) {
                const uint32_t j = qx + static_cast<uint32_t>(k);
                uint32_t c_q;
                int kappa = 1;
                if (initial) {
                    c_q = ((rho_w & 1) | ((rho_w >> 1) & 1)) | (((rho_w >> 2) & 1) << 1) |
                          (((rho_w >> 3) & 1) << 2);
                } else {
                    const uint32_t nw = j > 0 ? prev_rho[j - 1] : 0, n = prev_rho[j];
                    const uint32_t ne = prev_rho[j + 1];
                    c_q = (((nw >> 3) & 1) | ((n >> 1) & 1)) |
                          ((((rho_w >> 2) & 1) | ((rho_w >> 3) & 1)) << 1) |
                          ((((n >> 3) & 1) | ((ne >> 1) & 1)) << 2);
                    if (q[k].rho & (q[k].rho - 1)) {
                        // columns 2j-1 .. 2j+2 of the row above, stored at +1
                        int m = 0;
                        for (uint32_t cc = 2 * j; cc < 2 * j + 4; ++cc) m = std::max(m, prev_e[cc]);
                        kappa = std::max(1, m - 1);
                    }
                }
                if (!code_quad(initial ? 0 : 1, c_q, q[k], kappa, u[k])) {
                    set_last_error("HT encode: no VLC codeword for quad context");
                    return kErrBlock;
                }
                rho_w = q[k].rho;
                cur_rho[j] = q[k].rho;
                cur_e[2 * j + 1] = q[k].e[1];
                cur_e[2 * j + 2] = q[k].e[3];
            }
            // U-VLC (and, on the first row, the MEL pair event)
            if (initial && has1 && u[0] > 0 && u[1] > 0) {
                mel.encode(std::min(u[0], u[1]) > 2);
                if (u[0] > 2 && u[1] > 2) {
                    const UvlcCode a = uvlc_code(u[0] - 2), b = uvlc_code(u[1] - 2);
                    vlc.put(a.pre, a.pre_len);
                    vlc.put(b.pre, b.pre_len);
                    vlc.put(a.suf, a.suf_len);
                    vlc.put(b.suf, b.suf_len);
                } else if (u[0] > 2) {
                    const UvlcCode a = uvlc_code(u[0]);
                    vlc.put(a.pre, a.pre_len);
                    vlc.put(static_cast<uint32_t>(u[1] - 1), 1);
                    vlc.put(a.suf, a.suf_len);
                } else {
                    const UvlcCode a = uvlc_code(u[0]), b = uvlc_code(u[1]);
                    vlc.put(a.pre, a.pre_len);
                    vlc.put(b.pre, b.pre_len);
                    vlc.put(a.suf, a.suf_len);
                    vlc.put(b.suf, b.suf_len);
                }
            } else {
                const UvlcCode a = uvlc_code(u[0]), b = uvlc_code(u[1]);
                vlc.put(a.pre, a.pre_len);
                vlc.put(b.pre, b.pre_len);
                vlc.put(a.suf, a.suf_len);
                vlc.put(b.suf, b.suf_len);
            }
            if (!has1) rho_w = 0;
        }
        prev_rho.swap(cur_rho);
        prev_e.swap(cur_e);
    }

    // Terminate MEL and VLC, fusing their last partial bytes when possible.
    // After a 0xFF MEL byte the stuffed MSB counts as used.
    if (mel.run > 0) mel.bit(1);
    const int mel_rem = mel.maxbits - mel.nbits;
    const uint32_t mel_tmp = (mel.acc << mel_rem) & 0xFF;
    const uint32_t mel_mask = (0xFFu << mel_rem) & 0xFF;
    const uint32_t vlc_mask = vlc.used == 0 ? 0 : ((1u << vlc.used) - 1);
    if ((mel_mask | vlc_mask) != 0) {
        const uint32_t fused = mel_tmp | vlc.acc;
        const bool ok = ((((fused ^ mel_tmp) & mel_mask) | ((fused ^ vlc.acc) & vlc_mask)) == 0) &&
                        fused != 0xFF;
        if (ok && vlc.rev.size() > 1) {
            mel.buf.push_back(static_cast<uint8_t>(fused));
        } else {
            mel.buf.push_back(static_cast<uint8_t>(mel_tmp));
            vlc.rev.push_back(static_cast<uint8_t>(vlc.acc));
        }
    }
    ms.terminate();
    const size_t scup = mel.buf.size() + vlc.rev.size();
    out.reserve(ms.buf.size() + scup);
    out.insert(out.end(), ms.buf.begin(), ms.buf.end());
    out.insert(out.end(), mel.buf.begin(), mel.buf.end());
    for (size_t i = vlc.rev.size(); i-- > 0;) out.push_back(vlc.rev[i]);
    if (out.size() < 2 || scup > 4079 || out.size() >= 65535) {
        set_last_error("HT encode: cleanup segment out of range");
        return kErrBlock;
    }
    out[out.size() - 1] = static_cast<uint8_t>(scup >> 4);
    out[out.size() - 2] = static_cast<uint8_t>((out[out.size() - 2] & 0xF0) | (scup & 0x0F));
    return kOk;
}

}  // namespace sgn
