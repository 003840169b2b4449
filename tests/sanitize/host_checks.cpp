// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the CPU
// oracle (oracle/bicos_oracle.cpp), the scene generator
// (libbicos_amd/csrc/scene.cpp) and the PNG unfiltering
// (libbicos_amd/csrc/imageio.cpp), built together with this driver by
// tests/test_sanitizers.py. Every exported function runs on small inputs
// that reach its edge cases (ragged sizes, the smallest and largest stack
// depths, every descriptor width, windows clipped or empty, saturating
// subpixel parabolas, every PNG filter); a sanitizer report or a failed
// check ends the program with a non-zero status.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

extern "C" {
struct OracleConfig {
    int has_thr; float thr; int has_step; float step; int has_minvar; float minvar;
    int full; int dbl; int consistency; int max_lr_diff; int no_dupes;
};
const char* oracle_last_error(void);
void oracle_set_threads(int t);
int oracle_descriptor_words(int n, int full);
int oracle_transform(const void* const* frames, int64_t step, int depth, int n, int rows, int cols,
                     int full, int words, uint32_t* out);
int oracle_search(const uint32_t* d0, const uint32_t* d1, int rows, int cols, int words, int flags,
                  int max_lr_diff, int16_t* out);
float oracle_nxcorr_u16(const uint16_t* p0, const uint16_t* p1, int n, int has_minvar, float minvar);
double oracle_nxcorrd_u16(const uint16_t* p0, const uint16_t* p1, int n, int has_minvar, float minvar);
int oracle_subpixel_grid(float step, float* out, int cap);
int oracle_match(const void* const* f0, const void* const* f1, int64_t step, int depth, int n, int rows,
                 int cols, const OracleConfig* cfg, int16_t* disp16, float* dispf, int* is_float,
                 void* corrmap);
int oracle_interp_sample(int depth, int y0, int y1, int y2, float x);
int oracle_subpixel_curve(int depth, int n, int64_t npix, const void* left, const void* right3,
                          float step, int npts, float* out);
int oracle_search_window(const uint32_t* d0, const uint32_t* d1, int rows, int cols, int words,
                         int flags, int max_lr_diff, const float* center, const int16_t* halfwidth,
                         int16_t* out);
int oracle_match_windowed(const void* const* f0, const void* const* f1, int64_t step, int depth,
                          int n, int rows, int cols, const OracleConfig* cfg, const float* center,
                          const int16_t* halfwidth, int hw_lo, int fallback, int16_t* disp16,
                          float* dispf, int* is_float, void* corrmap, int64_t* n_fallback);
int bicos_scene_generate(int W, int H, int n, int depth, double dmin, double dmax, uint64_t seed,
                         double lo, double hi, double sigma, double p_dark, double dark_value,
                         int threads, void* left, void* right, float* disp);
int bicos_png_unfilter(const uint8_t* in, size_t rows, size_t stride, int bpp, uint8_t* out);
}

static int g_fail = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            g_fail = 1;                                                  \
        }                                                                \
    } while (0)

template<typename T>
static void run_scene_and_match(int W, int H, int n, int depth, bool full, bool subpix, bool cons) {
    const size_t plane = (size_t)W * H;
    std::vector<T> L(plane * n), R(plane * n);
    std::vector<float> D(plane);
    const double hi = depth == 1 ? 235.0 : 60000.0, lo = depth == 1 ? 20.0 : 1000.0;
    CHECK(bicos_scene_generate(W, H, n, depth, 0.0, W / 4.0, 7, lo, hi, depth == 1 ? 2.0 : 150.0,
                               0.05, depth == 1 ? 10.0 : 500.0, 2, L.data(), R.data(), D.data()) == 0);
    std::vector<const void*> f0(n), f1(n);
    for (int t = 0; t < n; ++t) {
        f0[t] = L.data() + t * plane;
        f1[t] = R.data() + t * plane;
    }
    OracleConfig c{1, 0.8f, subpix ? 1 : 0, 0.25f, 1, 1.0f, full ? 1 : 0, 0, cons ? 1 : 0, 1, 0};
    std::vector<int16_t> d16(plane);
    std::vector<float> df(plane), cm(plane);
    int is_float = -1;
    const int rc = oracle_match(f0.data(), f1.data(), (int64_t)W * sizeof(T), depth, n, H, W, &c,
                                d16.data(), df.data(), &is_float, cm.data());
    const int words = oracle_descriptor_words(n, full ? 1 : 0);
    CHECK(rc == (words ? 0 : -1));
    if (rc != 0) return;
    CHECK(is_float == (subpix ? 1 : 0));
    // windowed match: cell windows, some clipped, some empty
    const int Hq = H / 4, Wq = W / 4;
    if (Hq < 1 || Wq < 1) return;
    std::vector<float> ctr(Hq * Wq);
    std::vector<int16_t> hw(Hq * Wq);
    for (int i = 0; i < Hq * Wq; ++i) {
        ctr[i] = (i % 5 == 0) ? -1e4f : (i % 7 == 0 ? 1e4f : (float)(i % 9) + 0.5f);
        hw[i] = (int16_t)(i % 3 == 0 ? 0 : (i % 3 == 1 ? 6 : 12));
    }
    int64_t nfb = -1;
    for (int fb = 0; fb < 3; ++fb) {
        CHECK(oracle_match_windowed(f0.data(), f1.data(), (int64_t)W * sizeof(T), depth, n, H, W, &c,
                                    ctr.data(), hw.data(), 6, fb, d16.data(), df.data(), &is_float,
                                    cm.data(), &nfb) == 0);
        CHECK(nfb >= 0 && nfb <= (int64_t)plane);
    }
    // descriptors and both searches directly
    std::vector<uint32_t> d0(plane * words), d1(plane * words);
    CHECK(oracle_transform(f0.data(), (int64_t)W * sizeof(T), depth, n, H, W, full ? 1 : 0, words, d0.data()) == 0);
    CHECK(oracle_transform(f1.data(), (int64_t)W * sizeof(T), depth, n, H, W, full ? 1 : 0, words, d1.data()) == 0);
    for (int flags = 0; flags < 4; ++flags) {
        CHECK(oracle_search(d0.data(), d1.data(), H, W, words, flags, 2, d16.data()) == 0);
        CHECK(oracle_search_window(d0.data(), d1.data(), H, W, words, flags, 2, ctr.data(), hw.data(),
                                   d16.data()) == 0);
    }
}

int main() {
    oracle_set_threads(3);
    // stacks: u8 and u16, every descriptor width, ragged sizes
    run_scene_and_match<uint8_t>(37, 9, 2, 1, true, true, false);
    run_scene_and_match<uint8_t>(64, 12, 12, 1, true, false, true);
    run_scene_and_match<uint8_t>(50, 8, 16, 1, true, true, true);
    run_scene_and_match<uint8_t>(50, 8, 17, 1, true, true, false);   // too many bits: rejected
    run_scene_and_match<uint16_t>(41, 7, 3, 2, false, true, true);
    run_scene_and_match<uint16_t>(96, 10, 9, 2, false, true, false);
    run_scene_and_match<uint16_t>(96, 10, 17, 2, false, false, true);
    run_scene_and_match<uint16_t>(130, 8, 33, 2, false, true, false);
    run_scene_and_match<uint16_t>(70, 5, 65, 2, false, true, true);
    run_scene_and_match<uint16_t>(70, 5, 66, 2, false, true, false); // too many bits: rejected

    // NCC, the subpixel grid, interpolated samples (wrapping casts included)
    uint16_t a[64], b[64];
    for (int i = 0; i < 64; ++i) {
        a[i] = (uint16_t)(i * 997 % 65536);
        b[i] = (uint16_t)(65535 - a[i]);
    }
    CHECK(std::fabs(oracle_nxcorr_u16(a, b, 64, 0, 0.f) + 1.0f) < 1e-5f);
    CHECK(std::fabs(oracle_nxcorrd_u16(a, a, 64, 0, 0.f) - 1.0) < 1e-12);
    (void)oracle_nxcorr_u16(a, a, 1, 0, 0.f);  // one sample: 0 / 0
    float grid[64];
    CHECK(oracle_subpixel_grid(0.1f, grid, 64) == 20);
    CHECK(oracle_subpixel_grid(0.0f, grid, 64) < 0);
    CHECK(oracle_interp_sample(2, 0, 60000, 65535, 0.6f) == 4321);
    CHECK(oracle_interp_sample(1, 255, 20, 0, 0.6f) == 238);
    std::mt19937 rng(3);
    const int P = 300, n = 33;
    std::vector<uint16_t> l(P * n), r3(P * 3 * n);
    for (auto& v : l) v = (uint16_t)(rng() % 65536);
    for (auto& v : r3) v = (uint16_t)(rng() % 65536);
    std::vector<float> curve(P * 20);
    CHECK(oracle_subpixel_curve(2, n, P, l.data(), r3.data(), 0.1f, 20, curve.data()) == 0);

    // PNG unfiltering: every filter type, 1..8 bytes per pixel, ragged rows
    for (int bpp = 1; bpp <= 8; ++bpp)
        for (size_t stride : {(size_t)1, (size_t)bpp, (size_t)(3 * bpp + 1)}) {
            const size_t rows = 5;
            std::vector<uint8_t> in(rows * (stride + 1)), out(rows * stride);
            for (size_t rr = 0; rr < rows; ++rr) {
                in[rr * (stride + 1)] = (uint8_t)(rr % 5);
                for (size_t i = 0; i < stride; ++i) in[rr * (stride + 1) + 1 + i] = (uint8_t)(rng() & 255);
            }
            CHECK(bicos_png_unfilter(in.data(), rows, stride, bpp, out.data()) == 0);
            in[0] = 9;
            CHECK(bicos_png_unfilter(in.data(), rows, stride, bpp, out.data()) == -1);
        }
    // scene generator argument checks
    uint8_t dummy[16];
    CHECK(bicos_scene_generate(0, 4, 2, 1, 0, 1, 1, 0, 1, 0, 0, 0, 1, dummy, dummy, nullptr) == -1);
    CHECK(bicos_scene_generate(4, 4, 2, 3, 0, 1, 1, 0, 1, 0, 0, 0, 1, dummy, dummy, nullptr) == -1);
    std::printf(g_fail ? "FAILED\n" : "host checks ok\n");
    return g_fail;
}
