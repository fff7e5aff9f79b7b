// Containers around a JPEG 2000 codestream: JP2 files (T.800 Annex I) and
// DICOM encapsulated Pixel Data (PS3.5 A.4).
//
// JP2: box order and the required boxes follow the reference facade's
// parser (crates/signinum-j2k/src/parse/boxes.rs:14-160): signature box
// first and once, ftyp once before jp2h and jp2c, jp2h once before jp2c and
// holding an ihdr, then exactly one jp2c.  The colour specification maps
// to the facade's Colorspace as in parse_colr (boxes.rs:130-147).  The
// jp2h children that change the decoded samples after the codestream
// decode (pclr/cmap palette, cdef channel reordering, sYCC and CIELab
// conversion) are recorded so the device path can refuse what it does not
// apply, as the native decoder's direct device plane reuse does
// (crates/signinum-j2k-native/src/lib.rs:1493-1509).
//
// DICOM: frame extraction restates extract_dicom_encapsulated_frames_with_limit
// (crates/signinum-j2k-metal/src/dicom.rs:22-427): the Pixel Data element
// (7FE0,0010) of undefined length, its Basic Offset Table item, the Extended
// Offset Table (7FE0,0001) and its lengths (7FE0,0002), fragments grouped
// into frames by item offset.
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "parse.hpp"
#include "sgn_api.hpp"

namespace sgn {

namespace {

uint32_t be32(const uint8_t *p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3];
}

uint64_t be64(const uint8_t *p) { return (uint64_t(be32(p)) << 32) | be32(p + 4); }

struct Box {
    uint32_t type = 0;
    size_t payload = 0, end = 0;
};

constexpr uint32_t tag(const char (&s)[5]) {
    return (uint32_t(uint8_t(s[0])) << 24) | (uint32_t(uint8_t(s[1])) << 16) |
           (uint32_t(uint8_t(s[2])) << 8) | uint32_t(uint8_t(s[3]));
}

// Box header at `off` of [0, len): LBox 0 runs to the end, LBox 1 takes the
// 64-bit XLBox (boxes.rs:161-210).
Status read_box(const uint8_t *d, size_t len, size_t off, Box &b) {
    if (len < 8 || off > len - 8) return fail_status(kErrMalformed, "truncated JP2 box header");
    const uint32_t lbox = be32(d + off);
    b.type = be32(d + off + 4);
    if (lbox == 0) {
        b.payload = off + 8;
        b.end = len;
    } else if (lbox == 1) {
        if (len < 16 || off > len - 16)
            return fail_status(kErrMalformed, "truncated JP2 extended box header");
        const uint64_t xl = be64(d + off + 8);
        if (xl < 16 || xl > len - off) return fail_status(kErrMalformed, "JP2 box runs past the input");
        b.payload = off + 16;
        b.end = off + size_t(xl);
    } else {
        if (lbox < 8 || lbox > len - off) return fail_status(kErrMalformed, "JP2 box runs past the input");
        b.payload = off + 8;
        b.end = off + lbox;
    }
    return kOk;
}

// Channel definitions (I.5.3.6) are the identity when channel i is colour
// channel i+1 (typ 0, association i+1) for every colour channel and any
// remaining channel is an opacity channel of the whole image (association
// 0) placed last: the decoded component order is then the output order.
bool cdef_is_identity(const uint8_t *p, size_t n) {
    if (n < 2) return false;
    const uint32_t count = (uint32_t(p[0]) << 8) | p[1];
    if (n < 2 + size_t(count) * 6) return false;
    for (uint32_t i = 0; i < count; ++i) {
        const uint8_t *e = p + 2 + i * 6;
        const uint32_t cn = (uint32_t(e[0]) << 8) | e[1];
        const uint32_t typ = (uint32_t(e[2]) << 8) | e[3];
        const uint32_t asoc = (uint32_t(e[4]) << 8) | e[5];
        if (cn != i) return false;
        if (typ == 0 && asoc == i + 1) continue;
        if ((typ == 1 || typ == 2) && asoc == 0 && i + 1 == count) continue;
        return false;
    }
    return true;
}

}  // namespace

Status fail_status(Status s, const char *msg) {
    set_last_error(msg);
    return s;
}

bool looks_like_jp2(const uint8_t *d, size_t len) {
    static const uint8_t prefix[8] = {0, 0, 0, 12, 'j', 'P', ' ', ' '};
    return len >= 8 && std::memcmp(d, prefix, 8) == 0;
}

Status locate_codestream(const uint8_t *d, size_t len, Container &c) {
    c = Container();
    c.cs = d;
    c.cs_len = len;
    if (!d) return fail_status(kErrArg, "null input");
    if (!looks_like_jp2(d, len)) {
        if (len >= 2 && d[0] == 0xFF && d[1] == 0x4F) return kOk;
        return fail_status(kErrUnsupported,
                           "input is not a JP2 container or raw JPEG 2000 codestream");
    }
    c.jp2 = true;
    static const uint8_t sig[4] = {0x0D, 0x0A, 0x87, 0x0A};
    bool saw_sig = false, saw_ftyp = false, saw_jp2h = false, saw_ihdr = false, saw_cs = false;
    size_t off = 0;
    while (off < len) {
        Box b;
        Status st = read_box(d, len, off, b);
        if (st) return st;
        const uint8_t *p = d + b.payload;
        const size_t n = b.end - b.payload;
        if (b.type == tag("jP  ")) {
            if (saw_sig || off != 0 || n != 4 || std::memcmp(p, sig, 4) != 0)
                return fail_status(kErrMalformed, "invalid JP2 signature box");
            saw_sig = true;
        } else if (b.type == tag("ftyp")) {
            if (!saw_sig || saw_ftyp || saw_jp2h || saw_cs || n < 8)
                return fail_status(kErrMalformed,
                                   "file type box must appear once before jp2h and jp2c");
            saw_ftyp = true;
        } else if (b.type == tag("jp2h")) {
            if (!saw_ftyp || saw_jp2h || saw_cs)
                return fail_status(kErrMalformed, "jp2h must appear once after ftyp and before jp2c");
            saw_jp2h = true;
            size_t o = 0;
            while (o < n) {
                Box cb;
                st = read_box(p, n, o, cb);
                if (st) return st;
                const uint8_t *q = p + cb.payload;
                const size_t m = cb.end - cb.payload;
                if (cb.type == tag("ihdr")) {
                    if (m < 14) return fail_status(kErrMalformed, "ihdr payload shorter than 14 bytes");
                    saw_ihdr = true;
                } else if (cb.type == tag("colr")) {
                    // Info's colour space: every colr box that names one
                    // replaces the previous (boxes.rs:118-120, 130-147);
                    // sample conversion: the first colr box only, as a
                    // conforming JP2 reader does (native jp2/colr.rs:9-19)
                    if (m >= 3) {
                        const bool first = !c.saw_colr;
                        c.saw_colr = true;
                        if (q[0] == 1 && m >= 7) {
                            const uint32_t e = be32(q + 3);
                            if (first) c.enumcs = e;
                            c.colorspace = e == 16   ? kCsSRgb
                                           : e == 17 ? kCsSGray
                                           : e == 18 ? kCsYCbCr
                                                     : kCsIccTagged;
                        } else if (q[0] == 2) {
                            c.colorspace = kCsIccTagged;
                        }
                    }
                } else if (cb.type == tag("pclr")) {
                    c.palette = true;
                } else if (cb.type == tag("cdef")) {
                    c.cdef = true;
                    c.cdef_identity = cdef_is_identity(q, m);
                }
                o = cb.end;
            }
        } else if (b.type == tag("jp2c")) {
            if (!saw_jp2h || saw_cs) return fail_status(kErrMalformed, "jp2c must appear once after jp2h");
            saw_cs = true;
            c.cs = p;
            c.cs_len = n;
            c.cs_offset = b.payload;
        }
        off = b.end;
    }
    if (!saw_ftyp) return fail_status(kErrMalformed, "missing required JP2 box ftyp");
    if (!saw_jp2h) return fail_status(kErrMalformed, "missing required JP2 box jp2h");
    if (!saw_ihdr) return fail_status(kErrMalformed, "missing required JP2 box ihdr");
    if (!saw_cs) return fail_status(kErrMalformed, "missing required JP2 box jp2c");
    return kOk;
}

Status device_decodable(const Container &c) {
    if (!c.jp2) return kOk;
    if (c.palette)
        return fail_status(kErrUnsupported, "JP2 palette (pclr) images are not decoded on the device");
    if (c.cdef && !c.cdef_identity)
        return fail_status(kErrUnsupported,
                           "JP2 channel definitions that reorder channels are not applied on the device");
    if (c.enumcs == 18 || c.enumcs == 14)
        return fail_status(kErrUnsupported,
                           "JP2 sYCC and CIELab colour spaces are not converted on the device");
    return kOk;
}

// ---- DICOM encapsulated Pixel Data ---------------------------------------

namespace {

constexpr uint32_t kUndefined = 0xFFFFFFFFu;

uint32_t le32(const uint8_t *p) {
    return uint32_t(p[0]) | (uint32_t(p[1]) << 8) | (uint32_t(p[2]) << 16) | (uint32_t(p[3]) << 24);
}

uint64_t le64(const uint8_t *p) { return uint64_t(le32(p)) | (uint64_t(le32(p + 4)) << 32); }

size_t find_tag(const uint8_t *d, size_t len, const uint8_t (&t)[4]) {
    if (len < 4) return SIZE_MAX;
    for (size_t i = 0; i + 4 <= len; ++i)
        if (d[i] == t[0] && d[i + 1] == t[1] && d[i + 2] == t[2] && d[i + 3] == t[3]) return i;
    return SIZE_MAX;
}

bool explicit_vr(const uint8_t *vr) {
    return vr[0] >= 'A' && vr[0] <= 'Z' && vr[1] >= 'A' && vr[1] <= 'Z';
}

bool long_explicit_vr(const uint8_t *vr) {
    static const char *kLong[] = {"OB", "OD", "OF", "OL", "OV", "OW", "SQ", "UC", "UR", "UT", "UN"};
    for (const char *v : kLong)
        if (vr[0] == uint8_t(v[0]) && vr[1] == uint8_t(v[1])) return true;
    return false;
}

// Value offset and length of the element whose tag starts at `at`
// (explicit VR with a 2- or 4-byte length, or implicit VR).
Status element_header(const uint8_t *d, size_t len, size_t at, const char *what, size_t &value,
                      uint32_t &length) {
    const size_t after = at + 4;
    if (after + 2 > len) return fail_status(kErrMalformed, what);
    const uint8_t *vr = d + after;
    if (explicit_vr(vr)) {
        if (long_explicit_vr(vr)) {
            if (after + 8 > len) return fail_status(kErrMalformed, what);
            length = le32(d + after + 4);
            value = after + 8;
        } else {
            if (after + 4 > len) return fail_status(kErrMalformed, what);
            length = uint32_t(d[after + 2]) | (uint32_t(d[after + 3]) << 8);
            value = after + 4;
        }
    } else {
        if (after + 4 > len) return fail_status(kErrMalformed, what);
        length = le32(d + after);
        value = after + 4;
    }
    return kOk;
}

// An element (Extended Offset Table or its lengths) searched for before
// the Pixel Data tag; absent is not an error (dicom.rs:255-274).
Status u64_table_before(const uint8_t *d, size_t len, size_t before, const uint8_t (&t)[4],
                        const char *what, bool &present, std::vector<uint64_t> &out) {
    present = false;
    const size_t at = find_tag(d, before, t);
    if (at == SIZE_MAX) return kOk;
    size_t value;
    uint32_t length;
    Status st = element_header(d, len, at, what, value, length);
    if (st) return st;
    if (length == kUndefined) return fail_status(kErrMalformed, "Extended Offset Table element has undefined length");
    if (value > len || length > len - value) return fail_status(kErrMalformed, what);
    if (length % 8) return fail_status(kErrMalformed, what);
    out.resize(length / 8);
    for (size_t i = 0; i < out.size(); ++i) out[i] = le64(d + value + 8 * i);
    present = true;
    return kOk;
}

struct Fragment {
    uint64_t item_offset;
    size_t off, len;
};

}  // namespace

Status dicom_frames(const uint8_t *d, size_t len, size_t limit, std::vector<DicomFrame> &frames) {
    frames.clear();
    if (limit == 0) return kOk;
    static const uint8_t kPixel[4] = {0xE0, 0x7F, 0x10, 0x00};
    static const uint8_t kExtOff[4] = {0xE0, 0x7F, 0x01, 0x00};
    static const uint8_t kExtLen[4] = {0xE0, 0x7F, 0x02, 0x00};
    static const uint8_t kItem[4] = {0xFE, 0xFF, 0x00, 0xE0};
    static const uint8_t kSeqDelim[4] = {0xFE, 0xFF, 0xDD, 0xE0};
    const size_t tag_at = find_tag(d, len, kPixel);
    if (tag_at == SIZE_MAX) return fail_status(kErrMalformed, "DICOM Pixel Data element was not found");
    if (len - tag_at < 8) return fail_status(kErrMalformed, "DICOM Pixel Data is truncated while reading Pixel Data header");
    size_t value;
    uint32_t length;
    Status st = element_header(d, len, tag_at, "DICOM Pixel Data is truncated while reading Pixel Data length", value, length);
    if (st) return st;
    if (length != kUndefined) return fail_status(kErrUnsupported, "DICOM Pixel Data is not encapsulated");
    if (value > len) return fail_status(kErrMalformed, "DICOM Pixel Data is truncated while reading Pixel Data value");
    bool have_eo = false, have_el = false;
    std::vector<uint64_t> ext_off, ext_len;
    st = u64_table_before(d, len, tag_at, kExtOff, "Extended Offset Table", have_eo, ext_off);
    if (st) return st;
    st = u64_table_before(d, len, tag_at, kExtLen, "Extended Offset Table Lengths", have_el, ext_len);
    if (st) return st;
    if (have_eo && have_el && ext_off.size() != ext_len.size())
        return fail_status(kErrMalformed, "Extended Offset Table and Lengths have different entry counts");
    // Basic Offset Table: the first item
    if (value + 8 > len) return fail_status(kErrMalformed, "DICOM Pixel Data is truncated while reading Basic Offset Table item tag");
    if (std::memcmp(d + value, kItem, 4) != 0)
        return fail_status(kErrMalformed, "encapsulated Pixel Data does not start with a Basic Offset Table item");
    const uint32_t bot_len = le32(d + value + 4);
    const size_t bot_at = value + 8;
    if (bot_len > len - bot_at) return fail_status(kErrMalformed, "DICOM Pixel Data is truncated while reading Basic Offset Table payload");
    if (bot_len % 4) return fail_status(kErrMalformed, "Basic Offset Table length is not a multiple of four");
    std::vector<uint64_t> bot(bot_len / 4);
    for (size_t i = 0; i < bot.size(); ++i) bot[i] = le32(d + bot_at + 4 * i);
    const std::vector<uint64_t> &offs = !bot.empty() ? bot : (have_eo ? ext_off : bot);
    const bool by_table = !offs.empty();
    const bool use_lengths = bot.empty() && have_eo && have_el;
    uint64_t stop = UINT64_MAX;
    if (limit < offs.size()) stop = offs[limit];
    // fragments
    std::vector<Fragment> frags;
    size_t cur = bot_at + bot_len;
    bool have_first = false;
    size_t first = 0;
    for (;;) {
        if (cur > len || len - cur < 8) return fail_status(kErrMalformed, "DICOM Pixel Data is truncated while reading Pixel Data item tag");
        const uint8_t *t = d + cur;
        const uint32_t l = le32(t + 4);
        const size_t item_at = cur;
        cur += 8;
        if (std::memcmp(t, kSeqDelim, 4) == 0) {
            if (l != 0) return fail_status(kErrMalformed, "non-zero sequence delimitation length");
            break;
        }
        if (std::memcmp(t, kItem, 4) != 0) return fail_status(kErrMalformed, "expected item tag in encapsulated Pixel Data");
        if (l == kUndefined) return fail_status(kErrMalformed, "undefined-length Pixel Data item");
        if (l > len - cur) return fail_status(kErrMalformed, "DICOM Pixel Data is truncated while reading Pixel Data item payload");
        if (!have_first) {
            have_first = true;
            first = item_at;
        }
        const uint64_t io = uint64_t(item_at - first);
        if (io >= stop) break;
        frags.push_back({io, cur, l});
        if (!by_table && frags.size() >= limit) break;
        cur += l;
    }
    if (frags.empty()) return kOk;
    if (!by_table) {
        for (const Fragment &f : frags) frames.push_back({{{f.off, f.len}}, f.len});
        return kOk;
    }
    const size_t nf = std::min(limit, offs.size());
    for (size_t k = 0; k < nf; ++k) {
        const uint64_t a = offs[k];
        // the table is cut at the limit first (dicom.rs:75-88), so the
        // last frame kept runs to the end of the collected fragments
        const uint64_t b = k + 1 < nf ? offs[k + 1] : UINT64_MAX;
        if (b <= a) return fail_status(kErrMalformed, "offset table offsets are not strictly increasing");
        DicomFrame fr;
        fr.bytes = 0;
        const bool cap = use_lengths;
        const uint64_t maxlen = cap ? ext_len[k] : UINT64_MAX;
        for (const Fragment &f : frags) {
            if (f.item_offset < a || f.item_offset >= b) continue;
            size_t take = f.len;
            if (cap) {
                const uint64_t rem = maxlen > fr.bytes ? maxlen - fr.bytes : 0;
                take = size_t(std::min<uint64_t>(rem, take));
            }
            if (take) fr.pieces.push_back({f.off, take});
            fr.bytes += take;
            if (cap && fr.bytes >= maxlen) break;
        }
        if (fr.bytes == 0) return fail_status(kErrMalformed, "offset table frame has no fragments");
        frames.push_back(std::move(fr));
    }
    return kOk;
}

}  // namespace sgn

extern "C" int sgn_dicom_extract_frames(const uint8_t *bytes, size_t len, size_t limit,
                                        uint8_t *out, size_t out_capacity, size_t *frame_offsets,
                                        size_t *frame_lengths, size_t frames_capacity,
                                        size_t *frame_count, size_t *total_bytes) {
    try {
        if (!bytes || !frame_count || !total_bytes) return SGN_E_ARG;
        std::vector<sgn::DicomFrame> frames;
        const sgn::Status st = sgn::dicom_frames(bytes, len, limit, frames);
        if (st) return st;
        size_t total = 0;
        for (const sgn::DicomFrame &f : frames) total += f.bytes;
        *frame_count = frames.size();
        *total_bytes = total;
        if (!out) return SGN_OK;
        if (out_capacity < total || frames_capacity < frames.size() || !frame_offsets ||
            !frame_lengths) {
            sgn::set_last_error("frame buffer too small");
            return SGN_E_BUFFER;
        }
        size_t o = 0;
        for (size_t k = 0; k < frames.size(); ++k) {
            frame_offsets[k] = o;
            frame_lengths[k] = frames[k].bytes;
            for (const auto &p : frames[k].pieces) {
                std::memcpy(out + o, bytes + p.first, p.second);
                o += p.second;
            }
        }
        return SGN_OK;
    } catch (const std::exception &e) {
        sgn::set_last_error(e.what());
        return SGN_E_ARG;
    }
}
