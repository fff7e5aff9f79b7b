// Internal bridge between the C ABI header and the C++ implementation.
#pragma once
#include "../../../include/signinum_j2k_hip.h"
#include "j2k_host.hpp"

namespace sgn {
Status encode_image(const void *pixels, uint32_t width, uint32_t height, uint32_t ncomp,
                    uint32_t bit_depth, size_t row_stride, const sgn_j2k_encode_options &opt,
                    std::vector<uint8_t> &cs);
}  // namespace sgn
