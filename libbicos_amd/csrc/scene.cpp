/*
 * Synthetic rectified multi-shot stereo scenes with known disparity.
 *
 * The reference benchmarks on a downloaded dataset (data/prepare.sh:20) that is
 * not available offline; this generator produces inputs of the same kind:
 * band-limited speckle sequences seen by two rectified cameras.
 *
 *   disparity  D(x,y) on the left grid: 4 sinusoids (periods in [W/4, 2W],
 *              random direction, amplitude and phase) + 2 planar steps
 *              (half-planes over random lines), affinely normalised to [dmin, dmax];
 *   speckle    per frame, U[0,1) noise on a canvas of width W+ceil(dmax)+8,
 *              Gaussian blur (sigma 1.5, radius 5, separable), min-max scaled
 *              to [lo, hi];
 *   views      L_t(x,y) = P_t(x,y); R_t(x',y) = P_t(x,y) where x - D(x,y) = x'
 *              for the closest surface point landing on x' (forward warp of
 *              the row, linear between samples, with a depth test), so that
 *              left pixel x sees the same surface point as right pixel
 *              x - D(x,y) (BICOS convention d = col0 - col1); right pixels
 *              no surface lands on (disocclusions) see canvas columns no
 *              left pixel sees;
 *   noise      i.i.d. N(0, sigma) per pixel, frame and view, rounded half to
 *              even and clamped;
 *   dark       a Bernoulli(p_dark) set of left pixels is held at a constant
 *              value in every frame without noise, and so is the right pixel
 *              at round(x - D(x,y)).
 *
 * Every random number is a SplitMix64 hash of (seed, stream, t, y, x), so the
 * output is a pure function of the parameters, independent of threading.
 */
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <limits>
#include <thread>
#include <vector>

#include "scene_field.hpp"

namespace {

using namespace bicos_scene;

template<typename F>
void parallel_for(int count, int threads, F&& f) {
    threads = std::max(1, std::min(threads, count));
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int i = 0; i < threads; ++i)
        pool.emplace_back([&]() {
            for (int k; (k = next.fetch_add(1)) < count;) f(k);
        });
    for (auto& th : pool) th.join();
}

template<typename T>
void generate(int W, int H, int n, double dmin, double dmax, uint64_t seed, double lo, double hi,
              double sigma, double p_dark, double dark_value, int threads, T* left, T* right,
              float* disp) {
    const Field f = make_field(seed, W, H);

    /* disparity field, affinely normalised to [dmin, dmax] */
    std::vector<double> raw((size_t)W * H);
    parallel_for(H, threads, [&](int y) {
        for (int x = 0; x < W; ++x) raw[(size_t)y * W + x] = field_value(f, x, y);
    });
    double fmin = *std::min_element(raw.begin(), raw.end());
    double fmax = *std::max_element(raw.begin(), raw.end());
    double scale = fmax > fmin ? (dmax - dmin) / (fmax - fmin) : 0.0;
    std::vector<float> D((size_t)W * H);
    for (size_t i = 0; i < raw.size(); ++i) D[i] = (float)(dmin + (raw[i] - fmin) * scale);
    if (disp) std::copy(D.begin(), D.end(), disp);

    /* dark set: left pixels and their right correspondences */
    std::vector<uint8_t> darkL((size_t)W * H, 0), darkR((size_t)W * H, 0);
    if (p_dark > 0.0) {
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                if (u01d(hash5(seed, S_DARK, 0, (uint64_t)y, (uint64_t)x)) >= p_dark) continue;
                darkL[(size_t)y * W + x] = 1;
                long xr = std::lrint(x - (double)D[(size_t)y * W + x]);
                if (xr >= 0 && xr < W) darkR[(size_t)y * W + xr] = 1;
            }
    }

    /* left coordinate seen by each right pixel (forward warp, depth test) */
    const int CW0 = W + (int)std::ceil(dmax) + 8; /* canvas columns left pixels can see */
    std::vector<float> src((size_t)W * H);
    parallel_for(H, threads, [&](int y) {
        std::vector<double> zbuf(W);
        warp_row(D.data() + (size_t)y * W, W, (double)CW0, src.data() + (size_t)y * W, zbuf.data());
    });

    const int CW = CW0 + W; /* canvas width: + a region only disoccluded right pixels see */
    const int R = BLUR_R;
    double g[2 * R + 1];
    blur_taps(g);
    const double vmax = (double)std::numeric_limits<T>::max();

    parallel_for(n, threads, [&](int t) {
        /* speckle canvas (H x CW), blurred separably with clamped borders */
        std::vector<float> noise((size_t)CW * H), tmp((size_t)CW * H), P((size_t)CW * H);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < CW; ++x)
                noise[(size_t)y * CW + x] = u01(hash5(seed, S_SPECKLE, (uint64_t)t, (uint64_t)y, (uint64_t)x));
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < CW; ++x) {
                double s = 0.0;
                for (int i = -R; i <= R; ++i) {
                    int xx = std::min(std::max(x + i, 0), CW - 1);
                    s += g[i + R] * noise[(size_t)y * CW + xx];
                }
                tmp[(size_t)y * CW + x] = (float)s;
            }
        float pmin = 1e30f, pmax = -1e30f;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < CW; ++x) {
                double s = 0.0;
                for (int i = -R; i <= R; ++i) {
                    int yy = std::min(std::max(y + i, 0), H - 1);
                    s += g[i + R] * tmp[(size_t)yy * CW + x];
                }
                float v = (float)s;
                P[(size_t)y * CW + x] = v;
                pmin = std::min(pmin, v);
                pmax = std::max(pmax, v);
            }
        const double k = pmax > pmin ? (hi - lo) / (double)(pmax - pmin) : 0.0;
        auto scaled = [&](int y, double x) {
            /* linear interpolation of the scaled canvas at fractional x */
            x = std::min(std::max(x, 0.0), (double)(CW - 1));
            int x0 = (int)std::floor(x);
            int x1 = std::min(x0 + 1, CW - 1);
            double w = x - x0;
            double v = (1.0 - w) * P[(size_t)y * CW + x0] + w * P[(size_t)y * CW + x1];
            return lo + (v - pmin) * k;
        };
        auto quantise = [&](double v) {
            v = std::nearbyint(v); /* default rounding mode: half to even */
            return (T)std::min(std::max(v, 0.0), vmax);
        };
        T* Lt = left + (size_t)t * W * H;
        T* Rt = right + (size_t)t * W * H;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                size_t i = (size_t)y * W + x;
                if (darkL[i])
                    Lt[i] = (T)dark_value;
                else
                    Lt[i] = quantise(scaled(y, x) + sigma * gauss_noise(hash5(seed, S_NOISE_L, (uint64_t)t, (uint64_t)y, (uint64_t)x)));
                if (darkR[i])
                    Rt[i] = (T)dark_value;
                else
                    Rt[i] = quantise(scaled(y, (double)src[i]) + sigma * gauss_noise(hash5(seed, S_NOISE_R, (uint64_t)t, (uint64_t)y, (uint64_t)x)));
            }
    });
}

} // namespace

extern "C" {

/* Generates one scene. depth 1 -> uint8 samples, 2 -> uint16. left/right:
 * n*H*W samples each, frame-planar [t][y][x]. disp (nullable): H*W float
 * ground-truth disparity of the left view. Returns 0, or -1 on bad arguments. */
int bicos_scene_generate(int W, int H, int n, int depth, double dmin, double dmax, uint64_t seed,
                         double lo, double hi, double sigma, double p_dark, double dark_value,
                         int threads, void* left, void* right, float* disp) {
    if (W <= 0 || H <= 0 || n <= 0 || !(dmax >= dmin) || !left || !right) return -1;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    if (depth == 1)
        generate<uint8_t>(W, H, n, dmin, dmax, seed, lo, hi, sigma, p_dark, dark_value, threads,
                          (uint8_t*)left, (uint8_t*)right, disp);
    else if (depth == 2)
        generate<uint16_t>(W, H, n, dmin, dmax, seed, lo, hi, sigma, p_dark, dark_value, threads,
                           (uint16_t*)left, (uint16_t*)right, disp);
    else
        return -1;
    return 0;
}

} // extern "C"
