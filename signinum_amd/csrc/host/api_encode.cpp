// C ABI: error reporting and the test-input encoder.
#include <cstring>
#include <exception>

#include "sgn_api.hpp"

extern "C" {

const char *sgn_last_error_message(void) { return sgn::last_error(); }

void sgn_j2k_encode_default_options(sgn_j2k_encode_options *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->levels = 5;
    o->cb_width_log2 = 6;
    o->cb_height_log2 = 6;
    o->reversible = 1;
    o->use_mct = 1;
    o->progression = 0;
}

int sgn_j2k_encode_ht(const void *pixels, uint32_t width, uint32_t height,
                      uint32_t num_components, uint32_t bit_depth, size_t row_stride,
                      const sgn_j2k_encode_options *opts, uint8_t *out, size_t out_cap,
                      size_t *out_len) {
    try {
        sgn_j2k_encode_options d;
        if (!opts) {
            sgn_j2k_encode_default_options(&d);
            opts = &d;
        }
        std::vector<uint8_t> cs;
        const sgn::Status st = sgn::encode_image(pixels, width, height, num_components,
                                                 bit_depth, row_stride, *opts, cs);
        if (st) return st;
        if (out_len) *out_len = cs.size();
        if (!out || out_cap < cs.size()) {
            sgn::set_last_error("encode: output buffer too small");
            return SGN_E_BUFFER;
        }
        std::memcpy(out, cs.data(), cs.size());
        return SGN_OK;
    } catch (const std::bad_alloc &) {
        sgn::set_last_error("encode: out of memory");
        return SGN_E_NOMEM;
    } catch (const std::exception &e) {
        sgn::set_last_error(e.what());
        return SGN_E_ARG;
    }
}

}  // extern "C"
