// The host side of libbicos_hip (argument validation, type and width
// rules, option defaults, the network's host initialisation) under
// AddressSanitizer and UndefinedBehaviorSanitizer. tests/test_sanitizers.py
// builds the library's sources for the host only (no device code, so no
// kernel can run) and links this driver; every call here returns before
// any device work, with the error code the C ABI documents.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bicos_hip.h"

static int g_fail = 0;
#define CHECK(c)                                                                \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::fprintf(stderr, "check failed: %s (line %d): %s\n", #c, __LINE__, \
                         bicos_hip_last_error());                               \
            g_fail = 1;                                                         \
        }                                                                       \
    } while (0)

int main() {
    bicos_config c;
    bicos_hip_config_default(&c);
    CHECK(c.has_nxcorr_threshold == 1 && c.nxcorr_threshold == 0.5f && !c.has_subpixel_step &&
          !c.has_min_variance && c.mode == BICOS_MODE_LIMITED && c.variant == BICOS_VARIANT_NO_DUPLICATES);
    CHECK(bicos_hip_disparity_type(&c) == BICOS_16S && bicos_hip_corrmap_type(&c) == BICOS_32F);
    c.has_subpixel_step = 1;
    c.subpixel_step = 0.1f;
    CHECK(bicos_hip_disparity_type(&c) == BICOS_32F);
    c.precision = BICOS_PRECISION_DOUBLE;
    CHECK(bicos_hip_corrmap_type(&c) == BICOS_64F);
    for (unsigned n = 0; n < 80; ++n)
        for (int mode = 0; mode < 2; ++mode) {
            const int bits = bicos_hip_descriptor_bits(n, mode);
            CHECK(bits == -1 || bits == 32 || bits == 64 || bits == 128 || bits == 256);
        }
    CHECK(bicos_hip_descriptor_bits(33, BICOS_MODE_LIMITED) == 128);
    CHECK(bicos_hip_descriptor_bits(17, BICOS_MODE_FULL) == -1);

    // match: each validation error before any device work
    std::vector<uint16_t> host(64 * 48);
    bicos_image fr[70];
    for (auto& f : fr) f = bicos_image{host.data(), 48, 64, BICOS_16U, 0, 128};
    bicos_image disp{host.data(), 48, 64, BICOS_16S, 0, 128};
    bicos_hip_config_default(&c);
    CHECK(bicos_hip_match(fr, fr, 5, nullptr, &disp, nullptr, nullptr) == BICOS_ERR_INVALID_ARGUMENT);
    CHECK(bicos_hip_match(fr, fr, 1, &c, &disp, nullptr, nullptr) == BICOS_ERR_EXCEPTION);
    CHECK(bicos_hip_match(fr, fr, 70, &c, &disp, nullptr, nullptr) == BICOS_ERR_EXCEPTION);
    CHECK(bicos_hip_match(nullptr, fr, 5, &c, &disp, nullptr, nullptr) == BICOS_ERR_EXCEPTION);
    bicos_image bad[5];
    for (auto& f : bad) f = bicos_image{host.data(), 48, 64, BICOS_32F, 0, 256};
    CHECK(bicos_hip_match(bad, fr, 5, &c, &disp, nullptr, nullptr) == BICOS_ERR_EXCEPTION);
    bicos_image other[5];
    for (auto& f : other) f = bicos_image{host.data(), 40, 64, BICOS_16U, 0, 128};
    CHECK(bicos_hip_match(fr, other, 5, &c, &disp, nullptr, nullptr) == BICOS_ERR_EXCEPTION);
    c.mode = 7;
    CHECK(bicos_hip_match(fr, fr, 5, &c, &disp, nullptr, nullptr) == BICOS_ERR_INVALID_ARGUMENT);
    c.mode = BICOS_MODE_FULL;
    CHECK(bicos_hip_match(fr, fr, 17, &c, &disp, nullptr, nullptr) == BICOS_ERR_INVALID_ARGUMENT);
    bicos_hip_config_default(&c);
    c.variant = 5;
    CHECK(bicos_hip_match(fr, fr, 5, &c, &disp, nullptr, nullptr) == BICOS_ERR_INVALID_ARGUMENT);
    bicos_hip_config_default(&c);
    c.precision = 9;
    CHECK(bicos_hip_match(fr, fr, 5, &c, &disp, nullptr, nullptr) == BICOS_ERR_INVALID_ARGUMENT);
    bicos_hip_config_default(&c);
    c.has_subpixel_step = 1;
    c.subpixel_step = 0.0f;
    CHECK(bicos_hip_match(fr, fr, 5, &c, &disp, nullptr, nullptr) == BICOS_ERR_INVALID_ARGUMENT);
    c.subpixel_step = 1e-9f;
    CHECK(bicos_hip_match(fr, fr, 5, &c, &disp, nullptr, nullptr) == BICOS_ERR_INVALID_ARGUMENT);
    bicos_hip_config_default(&c);
    bicos_image wrong{host.data(), 48, 64, BICOS_32F, 0, 256};
    CHECK(bicos_hip_match(fr, fr, 5, &c, &wrong, nullptr, nullptr) == BICOS_ERR_EXCEPTION);
    CHECK(bicos_hip_match(fr, fr, 5, &c, nullptr, nullptr, nullptr) == BICOS_ERR_EXCEPTION);
    CHECK(bicos_hip_match_noalloc(fr, fr, 1, &c, &disp, nullptr, nullptr) == BICOS_ERR_EXCEPTION);
    CHECK(std::strlen(bicos_hip_last_error()) > 0);

    // stage entry points and switches
    CHECK(bicos_hip_set_search_impl(9) == BICOS_ERR_INVALID_ARGUMENT);
    CHECK(bicos_hip_set_subpixel_impl(9) == BICOS_ERR_INVALID_ARGUMENT);
    CHECK(bicos_hip_set_cnn_impl(9) == BICOS_ERR_INVALID_ARGUMENT);
    CHECK(bicos_hip_search_chunk_cols(BICOS_SEARCH_MFMA, 4) == 960 && bicos_hip_search_chunk_cols(0, 3) == -1);
    CHECK(bicos_hip_noalloc_scratch_bytes() == (size_t)96 << 20);

    // learned window: options, host initialisation, validation
    bicos_window_options o;
    bicos_hip_window_options_default(&o);
    CHECK(o.hw_lo == 6 && o.hw_hi == 12 && o.fallback == BICOS_FALLBACK_ALL);
    std::vector<float> params(bicos_hip_window_param_count());
    CHECK(bicos_hip_window_init(params.data(), 42) == BICOS_OK);
    bool finite = true;
    for (float v : params) finite = finite && std::isfinite(v);
    CHECK(finite);
    CHECK(bicos_hip_window_init(nullptr, 1) == BICOS_ERR_INVALID_ARGUMENT);
    bicos_hip_config_default(&c);
    CHECK(bicos_hip_match_windowed(fr, fr, 5, &c, nullptr, &o, &disp, nullptr, nullptr, nullptr) ==
          BICOS_ERR_INVALID_ARGUMENT);
    CHECK(bicos_hip_match_windowed(fr, fr, 3, &c, params.data(), &o, &disp, nullptr, nullptr, nullptr) ==
          BICOS_ERR_INVALID_ARGUMENT);
    o.bins = 4;
    CHECK(bicos_hip_match_windowed(fr, fr, 5, &c, params.data(), &o, &disp, nullptr, nullptr, nullptr) ==
          BICOS_ERR_INVALID_ARGUMENT);
    bicos_hip_window_options_default(&o);
    o.fallback = 9;
    CHECK(bicos_hip_match_windowed(fr, fr, 5, &c, params.data(), &o, &disp, nullptr, nullptr, nullptr) ==
          BICOS_ERR_INVALID_ARGUMENT);
    bicos_hip_window_options_default(&o);
    bicos_image tiny[5];
    for (auto& f : tiny) f = bicos_image{host.data(), 3, 64, BICOS_16U, 0, 128};
    CHECK(bicos_hip_match_windowed(tiny, tiny, 5, &c, params.data(), &o, &disp, nullptr, nullptr, nullptr) ==
          BICOS_ERR_INVALID_ARGUMENT);
    CHECK(bicos_hip_search_windowed(nullptr, nullptr, 8, 8, 4, 1, 0, nullptr, nullptr, nullptr, nullptr) ==
          BICOS_ERR_INVALID_ARGUMENT);
    CHECK(bicos_hip_window_label(nullptr, 8, nullptr, nullptr) == BICOS_ERR_INVALID_ARGUMENT);
    CHECK(bicos_hip_window_adam(nullptr, nullptr, nullptr, nullptr, nullptr, 1e-3f, 0.9f, 0.999f, 1e-8f, 1,
                                nullptr) == BICOS_ERR_INVALID_ARGUMENT);
    std::printf(g_fail ? "FAILED\n" : "hip host checks ok\n");
    return g_fail;
}
