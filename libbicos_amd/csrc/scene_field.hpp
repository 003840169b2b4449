// Random numbers and the disparity field of the synthetic scene generator,
// shared by its host (scene.cpp) and device (scene_gpu.hip) forms so both
// draw the same numbers: every random number is a SplitMix64 hash of
// (seed, stream, t, y, x).
#pragma once

#include <cmath>
#include <cstdint>

#ifdef __HIPCC__
#define SCENE_HD __host__ __device__
#else
#define SCENE_HD
#endif

namespace bicos_scene {

SCENE_HD inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

SCENE_HD inline uint64_t hash5(uint64_t seed, uint64_t stream, uint64_t t, uint64_t y, uint64_t x) {
    uint64_t h = splitmix64(seed);
    h = splitmix64(h ^ (stream * 0xD1B54A32D192ED03ull));
    h = splitmix64(h ^ (t * 0x8CB92BA72F3D8DD7ull));
    h = splitmix64(h ^ (y * 0xABC98388FB8FAC03ull));
    h = splitmix64(h ^ x);
    return h;
}

/* uniform in [0, 1) with 24 (u01) or 53 (u01d) random bits */
SCENE_HD inline float u01(uint64_t h) { return (float)(h >> 40) * (1.0f / 16777216.0f); }
SCENE_HD inline double u01d(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

enum Stream : uint64_t { S_PARAM = 1, S_SPECKLE = 2, S_NOISE_L = 3, S_NOISE_R = 4, S_DARK = 5 };

constexpr double SCENE_PI = 3.14159265358979323846;
constexpr int BLUR_R = 5;          // blur radius (taps -5..5)
constexpr double BLUR_SIGMA = 1.5;

struct Field {
    double dir[4], period[4], amp[4], phase[4];
    double px[2], py[2], nx[2], ny[2], h[2];
};

inline Field make_field(uint64_t seed, int W, int H) {
    Field f;
    int k = 0;
    auto r = [&]() { return u01d(hash5(seed, S_PARAM, 0, 0, (uint64_t)k++)); };
    for (int i = 0; i < 4; ++i) {
        f.dir[i] = r() * 2.0 * SCENE_PI;
        f.period[i] = W * (0.25 + 1.75 * r());
        f.amp[i] = 0.5 + 0.5 * r();
        f.phase[i] = r() * 2.0 * SCENE_PI;
    }
    for (int j = 0; j < 2; ++j) {
        f.px[j] = r() * W;
        f.py[j] = r() * H;
        double a = r() * 2.0 * SCENE_PI;
        f.nx[j] = std::cos(a);
        f.ny[j] = std::sin(a);
        f.h[j] = (0.4 + 0.6 * r()) * (r() < 0.5 ? -1.0 : 1.0);
    }
    return f;
}

SCENE_HD inline double field_value(const Field& f, double x, double y) {
    double v = 0.0;
    for (int i = 0; i < 4; ++i) {
        double u = x * cos(f.dir[i]) + y * sin(f.dir[i]);
        v += f.amp[i] * sin(2.0 * SCENE_PI * u / f.period[i] + f.phase[i]);
    }
    for (int j = 0; j < 2; ++j)
        if ((x - f.px[j]) * f.nx[j] + (y - f.py[j]) * f.ny[j] > 0.0) v += f.h[j];
    return v;
}

/* Box-Muller from the two 32-bit halves of one hash */
SCENE_HD inline double gauss_noise(uint64_t h) {
    double u1 = ((double)(h >> 32) + 0.5) * (1.0 / 4294967296.0);
    double u2 = ((double)(h & 0xFFFFFFFFull) + 0.5) * (1.0 / 4294967296.0);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * SCENE_PI * u2);
}

/* Right view of one row by forward warping the left row with a depth test:
 * the left row, linear between samples, lands at right coordinates
 * x - D(x); every right pixel x' gets the left coordinate of the closest
 * (largest-disparity) surface point that lands on it, and a right pixel no
 * forward-facing surface lands on (a disocclusion) gets the coordinate
 * right_only + x' of a canvas region no left pixel sees. zbuf: W scratch
 * values. Pure +, -, *, / and ceil, so host and device agree exactly. */
SCENE_HD inline void warp_row(const float* Dr, int W, double right_only, float* src, double* zbuf) {
    for (int xr = 0; xr < W; ++xr) {
        zbuf[xr] = -1e300;
        src[xr] = (float)(right_only + xr);
    }
    for (int x = 0; x + 1 < W; ++x) {
        const double a = x - (double)Dr[x], b = (x + 1) - (double)Dr[x + 1];
        if (!(b > a)) continue;  // folded away from the right camera
        int lo = (int)ceil(a), hi = (int)ceil(b) - 1;  // right pixels in [a, b)
        if (lo < 0) lo = 0;
        if (hi > W - 1) hi = W - 1;
        for (int xr = lo; xr <= hi; ++xr) {
            const double t = (xr - a) / (b - a);
            const double d = (1.0 - t) * (double)Dr[x] + t * (double)Dr[x + 1];
            if (d > zbuf[xr]) {
                zbuf[xr] = d;
                src[xr] = (float)(x + t);
            }
        }
    }
}

/* normalised Gaussian blur taps */
inline void blur_taps(double g[2 * BLUR_R + 1]) {
    double gs = 0.0;
    for (int i = -BLUR_R; i <= BLUR_R; ++i)
        gs += g[i + BLUR_R] = std::exp(-(double)(i * i) / (2.0 * BLUR_SIGMA * BLUR_SIGMA));
    for (int i = 0; i < 2 * BLUR_R + 1; ++i) g[i] /= gs;
}

}  // namespace bicos_scene
