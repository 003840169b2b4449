/*
 * BICOS CPU oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference libBICOS CPU backend, used as the
 * checker for the HIP implementation. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it. The product library never links,
 * loads or calls anything in this directory.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the libBICOS v2.2.0 source tree). The reference itself cannot be compiled in
 * this environment (it needs OpenCV and meson), so this file is the oracle.
 *
 * Build: see oracle/Makefile (g++ -O3 -ffp-contract=off, no FMA contraction,
 * matching a baseline x86-64 GCC build of the reference).
 *
 * Deviations from the reference, all deliberate and documented:
 *   - frames are passed as an array of n plane pointers sharing one row step
 *     instead of cv::merge()d n-channel Mats (cpu.cpp:112-114); the per-pixel
 *     sample order t = 0..n-1 is identical;
 *   - the float -> TInput cast of the interpolated subpixel sample
 *     (agree.hpp:164) is undefined behaviour in C++ for out-of-range values;
 *     the oracle spells out x86-64 GCC's de-facto result: convert to int32 with
 *     truncation (exact after roundevenf) and keep the low 8/16 bits;
 *   - subpixel_step <= 0 is rejected (the reference loops forever);
 *   - Precision::DOUBLE follows the CUDA backend's nxcorrd
 *     (include/impl/cuda/agree.cuh:35-65, 161-259), since the CPU backend has
 *     no double mode.
 */

#include <algorithm>
#include <atomic>
#include <bitset>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace oracle {

using uint128_t = __uint128_t;

/* include/common.hpp:34-48 */
template<typename T>
constexpr T INVALID_DISP =
    std::numeric_limits<T>::has_quiet_NaN ? std::numeric_limits<T>::quiet_NaN()
                                          : std::numeric_limits<T>::lowest();

template<typename T>
static bool is_invalid(T d) {
    if constexpr (std::is_floating_point_v<T>)
        return std::isnan(d);
    else
        return d == INVALID_DISP<T>;
}

/* include/impl/common.hpp:28-42 */
template<typename T> struct Wider;
template<> struct Wider<uint8_t> { using type = uint16_t; };
template<> struct Wider<uint16_t> { using type = uint32_t; };
template<typename T> using wider_t = typename Wider<T>::type;

/* include/impl/common.hpp:46-47 */
constexpr int FLAG_NODUPES = 1, FLAG_CONSISTENCY = 2;

static int g_threads = 0;

static int nthreads() {
    if (g_threads > 0) return g_threads;
    unsigned hc = std::thread::hardware_concurrency();
    return hc ? (int)hc : 1;
}

/* Row-parallel loop: the reference uses cv::parallel_for_ over rows
 * (descriptor_transform.hpp:130, bicos.hpp:89, agree.hpp:64,111). Rows are
 * independent, so any partition gives identical results. */
template<typename F>
static void parallel_rows(int rows, F&& f) {
    int nt = std::min(nthreads(), std::max(rows, 1));
    if (nt <= 1) {
        for (int r = 0; r < rows; ++r) f(r);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    for (int i = 0; i < nt; ++i)
        pool.emplace_back([&]() {
            for (int r; (r = next.fetch_add(1)) < rows;) f(r);
        });
    for (auto& th : pool) th.join();
}

/* ---------------------------------------------------------------- bitfield */

template<typename T> struct is_bitset : std::false_type {};
template<size_t N> struct is_bitset<std::bitset<N>> : std::true_type {};

/* include/impl/cpu/bitfield.hpp:34-58: bit i goes to position i, LSB first */
template<typename T>
struct Bitfield {
    unsigned int i = 0u;
    T v = T(0);
    void set(bool value) {
        if (value) {
            if constexpr (is_bitset<T>::value)
                v.set(i);
            else
                v |= T(1) << i;
        }
        i++;
    }
};

/* little-endian u32 word export of a descriptor (bit k -> word k/32, bit k%32) */
template<typename T> constexpr int words_of() {
    if constexpr (is_bitset<T>::value) return 8;
    else return (int)(sizeof(T) / 4);
}

template<typename T>
static void store_words(const T& v, uint32_t* w) {
    if constexpr (is_bitset<T>::value) {
        for (int k = 0; k < 8; ++k) {
            uint32_t x = 0;
            for (int b = 0; b < 32; ++b)
                if (v.test(32 * k + b)) x |= 1u << b;
            w[k] = x;
        }
    } else {
        for (int k = 0; k < words_of<T>(); ++k)
            w[k] = (uint32_t)(v >> (32 * k));
    }
}

template<typename T>
static T load_words(const uint32_t* w) {
    if constexpr (is_bitset<T>::value) {
        T v;
        for (int k = 0; k < 8; ++k)
            for (int b = 0; b < 32; ++b)
                if ((w[k] >> b) & 1u) v.set(32 * k + b);
        return v;
    } else {
        T v = 0;
        for (int k = 0; k < words_of<T>(); ++k)
            v |= T(w[k]) << (32 * k);
        return v;
    }
}

/* ------------------------------------------------------ descriptor transform */

/* include/impl/cpu/descriptor_transform.hpp:31-73 */
template<typename TInput, typename TDesc>
static TDesc transform_limited(const TInput* pix, size_t n) {
    Bitfield<TDesc> bf;

    float av = 0.0f;
    for (size_t t = 0; t < n; ++t)
        av += pix[t];
    av /= n;

    int prev_pair_sums[] = { -1, -1 };
    for (size_t t = 0; t < n - 2; ++t) {
        const TInput a = pix[t + 0], b = pix[t + 1], c = pix[t + 2];

        bf.set(a < b);
        bf.set(a < c);
        bf.set(a < av);

        int& prev_pair_sum = prev_pair_sums[t % 2];
        int current_sum = a + b;

        if (-1 != prev_pair_sum)
            bf.set(prev_pair_sum < current_sum);

        prev_pair_sum = current_sum;
    }

    const TInput a = pix[n - 2], b = pix[n - 1];

    bf.set(a < b);
    bf.set(a < av);
    bf.set(b < av);
    bf.set(prev_pair_sums[(n - 2) % 2] < (a + b));

    return bf.v;
}

/* include/impl/cpu/descriptor_transform.hpp:75-123 */
template<typename TInput, typename TDesc>
static TDesc transform_full(const TInput* pix, size_t _n) {
    Bitfield<TDesc> bf;
    const long n = (long)_n;

    float av = 0.0f;
    for (long t = 0; t < n; ++t)
        av += pix[t];
    av /= n;

    wider_t<TInput> pairsums[128];

    for (long t = 0; t < n - 2; ++t) {
        const TInput a = pix[t + 0], b = pix[t + 1], c = pix[t + 2];

        bf.set(a < b);
        bf.set(a < c);
        bf.set(a < av);

        pairsums[t] = wider_t<TInput>(pix[t]) + wider_t<TInput>(pix[t + 1]);
    }

    pairsums[n - 2] = wider_t<TInput>(pix[n - 2]) + wider_t<TInput>(pix[n - 1]);

    const TInput a = pix[n - 2], b = pix[n - 1];

    bf.set(a < b);
    bf.set(a < av);
    bf.set(b < av);

    for (long t = 0; t < n - 1; ++t) {
        for (long i = 0; i < n - 1; ++i) {
            if (i == t || i == t - 1 || i == t + 1)
                continue;
            bf.set(pairsums[t] < pairsums[i]);
        }
    }

    return bf.v;
}

/* Gathers the n samples of pixel (row, col): the reference reads them from a
 * cv::merge()d Mat, i.e. s.ptr<TInput>(row, col)[t] (descriptor_transform.hpp:134). */
template<typename TInput>
static void gather(const void* const* frames, int64_t step, int n, int row, int col, TInput* out) {
    for (int t = 0; t < n; ++t)
        out[t] = ((const TInput*)((const uint8_t*)frames[t] + row * step))[col];
}

/* include/impl/cpu/descriptor_transform.hpp:125-138 */
template<typename TInput, typename TDesc>
static void descriptor_transform(const void* const* frames, int64_t step, int n, int rows, int cols,
                                 bool full, std::vector<TDesc>& out) {
    out.assign((size_t)rows * cols, TDesc());
    parallel_rows(rows, [&](int row) {
        TInput pix[128];
        for (int col = 0; col < cols; ++col) {
            gather<TInput>(frames, step, n, row, col, pix);
            out[(size_t)row * cols + col] = full ? transform_full<TInput, TDesc>(pix, n)
                                                 : transform_limited<TInput, TDesc>(pix, n);
        }
    });
}

/* ------------------------------------------------------------------ search */

/* include/impl/cpu/bicos.hpp:29-48 */
static int ham(uint32_t a, uint32_t b) { return __builtin_popcount(a ^ b); }
static int ham(uint64_t a, uint64_t b) { return __builtin_popcountll(a ^ b); }
static int ham(uint128_t a, uint128_t b) {
    uint128_t diff = a ^ b;
    return __builtin_popcountll((uint64_t)(diff & 0xFFFFFFFFFFFFFFFFUL))
         + __builtin_popcountll((uint64_t)(diff >> 64));
}
template<size_t N>
static int ham(const std::bitset<N>& a, const std::bitset<N>& b) { return (int)(a ^ b).count(); }

/* include/impl/cpu/bicos.hpp:50-76 */
template<typename TDesc, int FLAGS>
static int bicos_search(const TDesc& d0, const TDesc* row1, size_t cols) {
    int best_col1 = INVALID_DISP<int>, min_cost = std::numeric_limits<int>::max(),
        num_duplicate_minima = 0;

    for (size_t col1 = 0; col1 < cols; ++col1) {
        const TDesc& d1 = row1[col1];
        int cost = ham(d0, d1);

        if (cost < min_cost) {
            min_cost = cost;
            best_col1 = (int)col1;
            if constexpr (FLAGS & FLAG_NODUPES)
                num_duplicate_minima = 0;
        } else if constexpr (FLAGS & FLAG_NODUPES) {
            if (cost == min_cost)
                num_duplicate_minima++;
        }
    }

    if constexpr (FLAGS & FLAG_NODUPES)
        if (0 < num_duplicate_minima)
            return INVALID_DISP<int>;

    return best_col1;
}

/* include/impl/cpu/bicos.hpp:78-113 */
template<typename TDesc, int FLAGS>
static void bicos(const TDesc* desc0, const TDesc* desc1, int max_lr_diff, int rows, int cols,
                  int16_t* out) {
    std::fill(out, out + (size_t)rows * cols, INVALID_DISP<int16_t>);

    parallel_rows(rows, [&](int row) {
        const TDesc *drow0 = desc0 + (size_t)row * cols, *drow1 = desc1 + (size_t)row * cols;
        int16_t* orow = out + (size_t)row * cols;

        for (int col0 = 0; col0 < cols; ++col0) {
            int best_col1 = bicos_search<TDesc, FLAGS>(drow0[col0], drow1, cols);

            if (is_invalid(best_col1))
                continue;

            if constexpr (FLAGS & FLAG_CONSISTENCY) {
                int reverse_col0 = bicos_search<TDesc, FLAGS>(drow1[best_col1], drow0, cols);

                if (is_invalid(reverse_col0) || std::abs(col0 - reverse_col0) > max_lr_diff)
                    continue;

                orow[col0] = (int16_t)((col0 + reverse_col0) / 2 - best_col1);
            } else
                orow[col0] = (int16_t)(col0 - best_col1);
        }
    });
}

template<typename TDesc>
static void bicos_dispatch(const TDesc* d0, const TDesc* d1, int flags, int max_lr_diff, int rows,
                           int cols, int16_t* out) {
    switch (flags) {
        case 0: bicos<TDesc, 0>(d0, d1, max_lr_diff, rows, cols, out); break;
        case 1: bicos<TDesc, 1>(d0, d1, max_lr_diff, rows, cols, out); break;
        case 2: bicos<TDesc, 2>(d0, d1, max_lr_diff, rows, cols, out); break;
        default: bicos<TDesc, 3>(d0, d1, max_lr_diff, rows, cols, out); break;
    }
}

/* ---------------------------------------------------------- windowed search
 *
 * The learned-window path's Hamming search (no reference counterpart): the
 * reference's bicos_search / bicos (include/impl/cpu/bicos.hpp:50-113) with
 * the candidate scan limited to a window. Pixel (row, col) belongs to the
 * 4 x 4 cell (min(row / 4, Hq - 1), min(col / 4, Wq - 1)), Hq = rows / 4,
 * Wq = cols / 4; the cell gives a centre disparity c (float) and a half-width
 * hw. The candidates are col1 = lo..hi with
 *   k  = col - rint(c)            (round half to even)
 *   lo = max(k - hw, 0),  hi = min(k + hw, cols - 1);
 * an empty window (lo > hi) gives INVALID. The first strict minimum and the
 * no-duplicates rule are bicos.hpp:59-73's, over the window only. With
 * CONSISTENCY the matched right column searches the left row back over
 * [max(col - hw, 0), min(col + hw, cols - 1)] (same flags), and the LR check
 * and midpoint are bicos.hpp:99-106's. */
struct CellWindows {
    const float* center;
    const int16_t* halfwidth;
    int Hq, Wq;
    int cell(int row, int col) const {
        return std::min(row / 4, Hq - 1) * Wq + std::min(col / 4, Wq - 1);
    }
};

/* include/impl/cpu/bicos.hpp:50-76, scanning col1 = lo..hi */
template<typename TDesc, int FLAGS>
static int bicos_search_window(const TDesc& d0, const TDesc* row1, int lo, int hi) {
    int best_col1 = INVALID_DISP<int>, min_cost = std::numeric_limits<int>::max(),
        num_duplicate_minima = 0;

    for (int col1 = lo; col1 <= hi; ++col1) {
        int cost = ham(d0, row1[col1]);

        if (cost < min_cost) {
            min_cost = cost;
            best_col1 = col1;
            if constexpr (FLAGS & FLAG_NODUPES)
                num_duplicate_minima = 0;
        } else if constexpr (FLAGS & FLAG_NODUPES) {
            if (cost == min_cost)
                num_duplicate_minima++;
        }
    }

    if constexpr (FLAGS & FLAG_NODUPES)
        if (0 < num_duplicate_minima)
            return INVALID_DISP<int>;

    return best_col1;
}

/* include/impl/cpu/bicos.hpp:78-113 with the windows above */
template<typename TDesc, int FLAGS>
static void bicos_window(const TDesc* desc0, const TDesc* desc1, int max_lr_diff, int rows,
                         int cols, const CellWindows& win, int16_t* out) {
    std::fill(out, out + (size_t)rows * cols, INVALID_DISP<int16_t>);

    parallel_rows(rows, [&](int row) {
        const TDesc *drow0 = desc0 + (size_t)row * cols, *drow1 = desc1 + (size_t)row * cols;
        int16_t* orow = out + (size_t)row * cols;

        for (int col0 = 0; col0 < cols; ++col0) {
            const int q = win.cell(row, col0);
            const int k = col0 - (int)std::nearbyint(win.center[q]);
            const int hw = win.halfwidth[q];
            const int lo = std::max(k - hw, 0), hi = std::min(k + hw, cols - 1);

            int best_col1 = bicos_search_window<TDesc, FLAGS>(drow0[col0], drow1, lo, hi);

            if (is_invalid(best_col1))
                continue;

            if constexpr (FLAGS & FLAG_CONSISTENCY) {
                int reverse_col0 = bicos_search_window<TDesc, FLAGS>(
                    drow1[best_col1], drow0, std::max(col0 - hw, 0), std::min(col0 + hw, cols - 1));

                if (is_invalid(reverse_col0) || std::abs(col0 - reverse_col0) > max_lr_diff)
                    continue;

                orow[col0] = (int16_t)((col0 + reverse_col0) / 2 - best_col1);
            } else
                orow[col0] = (int16_t)(col0 - best_col1);
        }
    });
}

template<typename TDesc>
static void bicos_window_dispatch(const TDesc* d0, const TDesc* d1, int flags, int max_lr_diff,
                                  int rows, int cols, const CellWindows& win, int16_t* out) {
    switch (flags) {
        case 0: bicos_window<TDesc, 0>(d0, d1, max_lr_diff, rows, cols, win, out); break;
        case 1: bicos_window<TDesc, 1>(d0, d1, max_lr_diff, rows, cols, win, out); break;
        case 2: bicos_window<TDesc, 2>(d0, d1, max_lr_diff, rows, cols, win, out); break;
        default: bicos_window<TDesc, 3>(d0, d1, max_lr_diff, rows, cols, win, out); break;
    }
}

/* ------------------------------------------------------------------- agree */

/* include/impl/cpu/agree.hpp:27-50 (FP32). Compiled without FP contraction:
 * every product and sum is rounded separately except the explicit fmaf. */
template<typename TA, typename TB>
static float nxcorr(const TA* pix0, const TB* pix1, size_t n, bool has_minvar, float minvar) {
    float mean0 = 0.f, mean1 = 0.f;
    for (size_t i = 0; i < n; ++i) {
        mean0 += pix0[i];
        mean1 += pix1[i];
    }
    mean0 /= n;
    mean1 /= n;

    float covar = 0.f, var0 = 0.f, var1 = 0.f;
    for (size_t i = 0; i < n; ++i) {
        float diff0 = pix0[i] - mean0, diff1 = pix1[i] - mean1;

        covar = std::fmaf(diff0, diff1, covar);
        var0 = std::fmaf(diff0, diff0, var0);
        var1 = std::fmaf(diff1, diff1, var1);
    }

    if (has_minvar && (var0 < minvar || var1 < minvar))
        return -1.f;

    return covar / std::sqrt(var0 * var1);
}

/* include/impl/cuda/agree.cuh:35-65 (FP64, Precision::DOUBLE) */
template<typename TA, typename TB>
static double nxcorrd(const TA* pix0, const TB* pix1, size_t n, bool has_minvar, double minvar) {
    double mean0 = 0.0, mean1 = 0.0;
    for (size_t i = 0; i < n; ++i) {
        mean0 = mean0 + (double)pix0[i];
        mean1 = mean1 + (double)pix1[i];
    }
    mean0 = mean0 / (double)n;
    mean1 = mean1 / (double)n;

    double covar = 0.0, var0 = 0.0, var1 = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double diff0 = pix0[i] - mean0, diff1 = pix1[i] - mean1;

        covar = std::fma(diff0, diff1, covar);
        var0 = std::fma(diff0, diff0, var0);
        var1 = std::fma(diff1, diff1, var1);
    }

    if (has_minvar && (var0 < minvar || var1 < minvar))
        return -1.0;

    return covar / std::sqrt(var0 * var1);
}

template<typename TPrec, typename TA, typename TB>
static TPrec nxc_any(const TA* p0, const TB* p1, size_t n, bool has_minvar, float minvar) {
    if constexpr (std::is_same_v<TPrec, float>)
        return nxcorr(p0, p1, n, has_minvar, minvar);
    else
        return nxcorrd(p0, p1, n, has_minvar, (double)minvar);
}

/* include/impl/cpu/agree.hpp:52-92 (FP32) and include/impl/cuda/agree.cuh:110-159 (FP64).
 * `minvar` is already multiplied by n (cpu.cpp:123). */
template<typename TInput, typename TPrec>
static void agree(int16_t* disp, const void* const* f0, const void* const* f1, int64_t step,
                  int n, int rows, int cols, float thr, bool has_minvar, float minvar,
                  TPrec* corrmap) {
    parallel_rows(rows, [&](int row) {
        TInput p0[128], p1[128];
        for (int col = 0; col < cols; ++col) {
            int16_t& d = disp[(size_t)row * cols + col];

            if (is_invalid(d))
                continue;

            const int idx1 = col - d;

            if (idx1 < 0 || cols <= idx1) {
                d = INVALID_DISP<int16_t>;
                continue;
            }

            gather<TInput>(f0, step, n, row, col, p0);
            gather<TInput>(f1, step, n, row, idx1, p1);
            TPrec nxc = nxc_any<TPrec>(p0, p1, n, has_minvar, minvar);

            if (corrmap)
                corrmap[(size_t)row * cols + col] = nxc;

            if (nxc < (TPrec)thr)
                d = INVALID_DISP<int16_t>;
        }
    });
}

/* The float -> TInput cast of agree.hpp:164 / agree.cuh:235: round half to
 * even, truncate to int32 (exact after rounding), keep the low bits. */
template<typename TInput>
static TInput cast_interp(float v) {
    return (TInput)(int32_t)roundevenf(v);
}

/* include/impl/cpu/agree.hpp:94-185 (FP32) and include/impl/cuda/agree.cuh:161-259 (FP64). */
template<typename TInput, typename TPrec>
static void agree_subpixel(const int16_t* raw_disp, const void* const* f0, const void* const* f1,
                           int64_t step, int n, int rows, int cols, float min_nxc,
                           float subpixel_step, bool has_minvar, float minvar, float* ret,
                           TPrec* corrmap) {
    std::fill(ret, ret + (size_t)rows * cols, INVALID_DISP<float>);

    parallel_rows(rows, [&](int row) {
        TInput p0[128], p1[128], interp[128], y0[128], y1[128], y2[128];
        float a[128], b[128], c[128];

        for (int col = 0; col < cols; ++col) {
            const int16_t d = raw_disp[(size_t)row * cols + col];

            if (is_invalid(d))
                continue;

            const int col1 = col - d;

            if (col1 < 0 || cols <= col1)
                continue;

            gather<TInput>(f0, step, n, row, col, p0);

            if (col1 == 0 || col1 == cols - 1) {
                gather<TInput>(f1, step, n, row, col1, p1);
                TPrec nxc = nxc_any<TPrec>(p0, p1, n, has_minvar, minvar);

                if (corrmap)
                    corrmap[(size_t)row * cols + col] = nxc;

                if (nxc < (TPrec)min_nxc)
                    continue;

                ret[(size_t)row * cols + col] = d;
            } else {
                gather<TInput>(f1, step, n, row, col1 - 1, y0);
                gather<TInput>(f1, step, n, row, col1, y1);
                gather<TInput>(f1, step, n, row, col1 + 1, y2);

                for (int t = 0; t < n; ++t) {
                    a[t] = 0.5f * (y0[t] - 2.0f * y1[t] + y2[t]);
                    b[t] = 0.5f * (-y0[t] + y2[t]);
                    c[t] = y1[t];
                }

                float best_x = 0.f;
                TPrec best_nxc = -1.f;

                for (float x = -1.f; x <= 1.f; x += subpixel_step) {
                    for (int t = 0; t < n; ++t)
                        interp[t] = cast_interp<TInput>(a[t] * x * x + b[t] * x + c[t]);

                    TPrec nxc = nxc_any<TPrec>(p0, interp, n, has_minvar, minvar);

                    if (best_nxc < nxc) {
                        best_x = x;
                        best_nxc = nxc;
                    }
                }

                if (corrmap)
                    corrmap[(size_t)row * cols + col] = best_nxc;

                if (best_nxc < (TPrec)min_nxc)
                    continue;

                ret[(size_t)row * cols + col] = d - best_x;
            }
        }
    });
}

/* --------------------------------------------------------------- dispatch */

static thread_local std::string g_err;

struct Config {
    int has_thr; float thr;
    int has_step; float step;
    int has_minvar; float minvar;
    int full;        /* TransformMode::FULL */
    int dbl;         /* Precision::DOUBLE */
    int consistency; /* Variant::Consistency */
    int max_lr_diff;
    int no_dupes;
};

/* src/impl/cpu.cpp:118-120 */
static int required_bits(int n, bool full) { return full ? n * n - 2 * n + 3 : 4 * n - 7; }

/* src/impl/cpu.cpp:125-152 -> descriptor width in u32 words, 0 if too large */
static int words_for_bits(int bits) {
    if (bits <= 32) return 1;
    if (bits <= 64) return 2;
    if (bits <= 128) return 4;
    if (bits <= 256) return 8;
    return 0;
}

template<typename TInput, typename TDesc>
static void transform_words(const void* const* frames, int64_t step, int n, int rows, int cols,
                            bool full, uint32_t* out) {
    std::vector<TDesc> d;
    descriptor_transform<TInput, TDesc>(frames, step, n, rows, cols, full, d);
    const int w = words_of<TDesc>();
    for (size_t i = 0; i < d.size(); ++i)
        store_words(d[i], out + i * w);
}

template<typename TDesc>
static void search_words(const uint32_t* w0, const uint32_t* w1, int rows, int cols, int flags,
                         int max_lr_diff, int16_t* out) {
    const int w = words_of<TDesc>();
    std::vector<TDesc> d0((size_t)rows * cols), d1((size_t)rows * cols);
    for (size_t i = 0; i < d0.size(); ++i) {
        d0[i] = load_words<TDesc>(w0 + i * w);
        d1[i] = load_words<TDesc>(w1 + i * w);
    }
    bicos_dispatch<TDesc>(d0.data(), d1.data(), flags, max_lr_diff, rows, cols, out);
}

/* src/impl/cpu.cpp:37-94 */
template<typename TInput, typename TDesc>
static void match_impl(const void* const* f0, const void* const* f1, int64_t step, int n, int rows,
                       int cols, const Config& cfg, float minvar_n, int16_t* disp16,
                       float* dispf, void* corrmap) {
    std::vector<TDesc> desc0, desc1;
    descriptor_transform<TInput, TDesc>(f0, step, n, rows, cols, cfg.full, desc0);
    descriptor_transform<TInput, TDesc>(f1, step, n, rows, cols, cfg.full, desc1);

    int flags = FLAG_NODUPES, lr = -1;
    if (cfg.consistency) {
        flags = FLAG_CONSISTENCY | (cfg.no_dupes ? FLAG_NODUPES : 0);
        lr = cfg.max_lr_diff;
    }
    bicos_dispatch<TDesc>(desc0.data(), desc1.data(), flags, lr, rows, cols, disp16);

    if (!cfg.has_thr)
        return;

    const size_t np = (size_t)rows * cols;
    if (corrmap) {
        if (cfg.dbl)
            std::fill((double*)corrmap, (double*)corrmap + np, std::numeric_limits<double>::quiet_NaN());
        else
            std::fill((float*)corrmap, (float*)corrmap + np, std::numeric_limits<float>::quiet_NaN());
    }

    if (cfg.has_step) {
        if (cfg.dbl)
            agree_subpixel<TInput, double>(disp16, f0, f1, step, n, rows, cols, cfg.thr, cfg.step,
                                           cfg.has_minvar, minvar_n, dispf, (double*)corrmap);
        else
            agree_subpixel<TInput, float>(disp16, f0, f1, step, n, rows, cols, cfg.thr, cfg.step,
                                          cfg.has_minvar, minvar_n, dispf, (float*)corrmap);
    } else {
        if (cfg.dbl)
            agree<TInput, double>(disp16, f0, f1, step, n, rows, cols, cfg.thr, cfg.has_minvar,
                                  minvar_n, (double*)corrmap);
        else
            agree<TInput, float>(disp16, f0, f1, step, n, rows, cols, cfg.thr, cfg.has_minvar,
                                 minvar_n, (float*)corrmap);
    }
}

/* The left-side variance as the NCC stage forms it: sequential mean, then
 * the fma chain of the centred samples (agree.hpp:31-44 in FP32,
 * agree.cuh:35-56 in FP64). */
template<typename TPrec, typename TInput>
static TPrec left_variance(const TInput* p0, size_t n) {
    if constexpr (std::is_same_v<TPrec, float>) {
        float mean0 = 0.f;
        for (size_t i = 0; i < n; ++i) mean0 += p0[i];
        mean0 /= n;
        float var0 = 0.f;
        for (size_t i = 0; i < n; ++i) {
            float diff0 = p0[i] - mean0;
            var0 = std::fmaf(diff0, diff0, var0);
        }
        return var0;
    } else {
        double mean0 = 0.0;
        for (size_t i = 0; i < n; ++i) mean0 = mean0 + (double)p0[i];
        mean0 = mean0 / (double)n;
        double var0 = 0.0;
        for (size_t i = 0; i < n; ++i) {
            double diff0 = p0[i] - mean0;
            var0 = std::fma(diff0, diff0, var0);
        }
        return var0;
    }
}

/* BICOS::match of the learned-window path (no reference counterpart):
 * match_impl (src/impl/cpu.cpp:37-94) with bicos() replaced by bicos_window()
 * and, when a threshold is set and `fallback` is 1 (all) or 2 (unsure), a
 * second pass for the fallback set: the pixels the NCC stage left invalid,
 * except (a) those whose left variance is below minvar when a minimum
 * variance and a threshold above -1 are set (the NCC stage rejects every
 * match of such a pixel: nxcorr returns -1, agree.hpp:46-47) and (b) with
 * fallback 2, those whose cell has the half-width hw_lo (the confident
 * cells). A fallback pixel gets the full-row search of bicos()
 * (bicos.hpp:78-113, the plain path's) and the NCC stage again; its output
 * and corrmap entry are those of that second pass. Every other pixel keeps
 * the windowed pass's. *n_fallback receives the size of the set. */
template<typename TInput, typename TDesc, typename TPrec>
static void match_windowed_impl(const void* const* f0, const void* const* f1, int64_t step, int n,
                                int rows, int cols, const Config& cfg, float minvar_n,
                                const CellWindows& win, int hw_lo, int fallback, int16_t* disp16,
                                float* dispf, TPrec* corrmap, int64_t* n_fallback) {
    std::vector<TDesc> desc0, desc1;
    descriptor_transform<TInput, TDesc>(f0, step, n, rows, cols, cfg.full, desc0);
    descriptor_transform<TInput, TDesc>(f1, step, n, rows, cols, cfg.full, desc1);

    int flags = FLAG_NODUPES, lr = -1;
    if (cfg.consistency) {
        flags = FLAG_CONSISTENCY | (cfg.no_dupes ? FLAG_NODUPES : 0);
        lr = cfg.max_lr_diff;
    }
    bicos_window_dispatch<TDesc>(desc0.data(), desc1.data(), flags, lr, rows, cols, win, disp16);
    *n_fallback = 0;
    if (!cfg.has_thr)
        return;

    const size_t np = (size_t)rows * cols;
    const TPrec nan = std::numeric_limits<TPrec>::quiet_NaN();
    if (corrmap)
        std::fill(corrmap, corrmap + np, nan);
    auto stage = [&](int16_t* d, float* out, TPrec* cm) {
        if (cfg.has_step)
            agree_subpixel<TInput, TPrec>(d, f0, f1, step, n, rows, cols, cfg.thr, cfg.step,
                                          cfg.has_minvar, minvar_n, out, cm);
        else
            agree<TInput, TPrec>(d, f0, f1, step, n, rows, cols, cfg.thr, cfg.has_minvar, minvar_n,
                                 cm);
    };
    stage(disp16, dispf, corrmap);
    if (fallback != 1 && fallback != 2)
        return;

    std::vector<uint8_t> take(np, 0);
    int64_t count = 0;
    TInput p0[128];
    for (int row = 0; row < rows; ++row)
        for (int col = 0; col < cols; ++col) {
            const size_t i = (size_t)row * cols + col;
            bool t = cfg.has_step ? std::isnan(dispf[i]) : is_invalid(disp16[i]);
            if (t && fallback == 2 && win.halfwidth[win.cell(row, col)] == hw_lo)
                t = false;
            if (t && cfg.has_minvar && cfg.thr > -1.f) {
                gather<TInput>(f0, step, n, row, col, p0);
                if (left_variance<TPrec>(p0, (size_t)n) < (TPrec)minvar_n)
                    t = false;
            }
            take[i] = t;
            count += t;
        }
    *n_fallback = count;

    // the full-row search (the plain path's) for the fallback set
    std::vector<int16_t> full(np);
    bicos_dispatch<TDesc>(desc0.data(), desc1.data(), flags, lr, rows, cols, full.data());
    for (size_t i = 0; i < np; ++i)
        if (!take[i]) full[i] = INVALID_DISP<int16_t>;
    std::vector<float> outf(cfg.has_step ? np : 0);
    std::vector<TPrec> cm2(corrmap ? np : 0, nan);
    stage(full.data(), outf.data(), corrmap ? cm2.data() : nullptr);
    for (size_t i = 0; i < np; ++i)
        if (take[i]) {
            if (cfg.has_step)
                dispf[i] = outf[i];
            else
                disp16[i] = full[i];
            if (corrmap) corrmap[i] = cm2[i];
        }
}

} // namespace oracle

using namespace oracle;

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

void oracle_set_threads(int t) { g_threads = t; }

int oracle_required_bits(int n, int full) { return required_bits(n, full != 0); }

int oracle_descriptor_words(int n, int full) { return words_for_bits(required_bits(n, full != 0)); }

/* Descriptor transform of one view. frames: n plane pointers of depth 1 (u8)
 * or 2 (u16) bytes sharing row step `step` (bytes). out: rows*cols*words u32. */
int oracle_transform(const void* const* frames, int64_t step, int depth, int n, int rows, int cols,
                     int full, int words, uint32_t* out) {
    if (n < 2 || n > 128) { g_err = "n out of range"; return -1; }
    const bool f = full != 0;
#define OT(TI, TD) transform_words<TI, TD>(frames, step, n, rows, cols, f, out)
    if (depth == 1) {
        switch (words) {
            case 1: OT(uint8_t, uint32_t); break;
            case 2: OT(uint8_t, uint64_t); break;
            case 4: OT(uint8_t, uint128_t); break;
            case 8: OT(uint8_t, std::bitset<256>); break;
            default: g_err = "bad words"; return -1;
        }
    } else if (depth == 2) {
        switch (words) {
            case 1: OT(uint16_t, uint32_t); break;
            case 2: OT(uint16_t, uint64_t); break;
            case 4: OT(uint16_t, uint128_t); break;
            case 8: OT(uint16_t, std::bitset<256>); break;
            default: g_err = "bad words"; return -1;
        }
    } else { g_err = "bad depth"; return -1; }
#undef OT
    return 0;
}

/* Row search (bicos.hpp:78-113) on descriptor words. flags: 1 NODUPES, 2 CONSISTENCY. */
int oracle_search(const uint32_t* d0, const uint32_t* d1, int rows, int cols, int words, int flags,
                  int max_lr_diff, int16_t* out) {
    switch (words) {
        case 1: search_words<uint32_t>(d0, d1, rows, cols, flags, max_lr_diff, out); break;
        case 2: search_words<uint64_t>(d0, d1, rows, cols, flags, max_lr_diff, out); break;
        case 4: search_words<uint128_t>(d0, d1, rows, cols, flags, max_lr_diff, out); break;
        case 8: search_words<std::bitset<256>>(d0, d1, rows, cols, flags, max_lr_diff, out); break;
        default: g_err = "bad words"; return -1;
    }
    return 0;
}

/* One NCC evaluation (agree.hpp:27-50), minvar already scaled by n. */
float oracle_nxcorr_u16(const uint16_t* p0, const uint16_t* p1, int n, int has_minvar, float minvar) {
    return nxcorr(p0, p1, (size_t)n, has_minvar != 0, minvar);
}
float oracle_nxcorr_u8(const uint8_t* p0, const uint8_t* p1, int n, int has_minvar, float minvar) {
    return nxcorr(p0, p1, (size_t)n, has_minvar != 0, minvar);
}
double oracle_nxcorrd_u16(const uint16_t* p0, const uint16_t* p1, int n, int has_minvar, float minvar) {
    return nxcorrd(p0, p1, (size_t)n, has_minvar != 0, (double)minvar);
}

/* In-place agree on an int16 disparity (agree.hpp:52-92); corrmap nullable,
 * float (dbl=0) or double (dbl=1). minvar already scaled by n. */
int oracle_agree(int16_t* disp, const void* const* f0, const void* const* f1, int64_t step,
                 int depth, int n, int rows, int cols, float thr, int has_minvar, float minvar,
                 int dbl, void* corrmap) {
#define OA(TI, TP) agree<TI, TP>(disp, f0, f1, step, n, rows, cols, thr, has_minvar != 0, minvar, (TP*)corrmap)
    if (depth == 1) { if (dbl) OA(uint8_t, double); else OA(uint8_t, float); }
    else if (depth == 2) { if (dbl) OA(uint16_t, double); else OA(uint16_t, float); }
    else { g_err = "bad depth"; return -1; }
#undef OA
    return 0;
}

/* Subpixel agree (agree.hpp:94-185) -> float disparity `out`. */
int oracle_agree_subpixel(const int16_t* disp, const void* const* f0, const void* const* f1,
                          int64_t step, int depth, int n, int rows, int cols, float thr,
                          float subpixel_step, int has_minvar, float minvar, int dbl, float* out,
                          void* corrmap) {
    if (!(subpixel_step > 0.f)) { g_err = "subpixel_step must be positive"; return -1; }
#define OS(TI, TP) agree_subpixel<TI, TP>(disp, f0, f1, step, n, rows, cols, thr, subpixel_step, has_minvar != 0, minvar, out, (TP*)corrmap)
    if (depth == 1) { if (dbl) OS(uint8_t, double); else OS(uint8_t, float); }
    else if (depth == 2) { if (dbl) OS(uint16_t, double); else OS(uint16_t, float); }
    else { g_err = "bad depth"; return -1; }
#undef OS
    return 0;
}

/* The subpixel x-grid of agree.hpp:162 (FP32 accumulation), for tests. */
int oracle_subpixel_grid(float step, float* out, int cap) {
    if (!(step > 0.f)) return -1;
    int k = 0;
    for (float x = -1.f; x <= 1.f; x += step) {
        if (k < cap) out[k] = x;
        ++k;
    }
    return k;
}

/* Whole BICOS::match on the CPU (cpu.cpp:96-155). Returns 0 and sets
 * *is_float (int16 disp16 valid if 0, float dispf valid if 1). Both output
 * buffers must hold rows*cols elements. corrmap nullable (float, or double
 * when cfg->dbl). Error: -1 with oracle_last_error(). */
int oracle_match(const void* const* f0, const void* const* f1, int64_t step, int depth, int n,
                 int rows, int cols, const oracle::Config* cfg, int16_t* disp16, float* dispf,
                 int* is_float, void* corrmap) {
    if (n < 2) { g_err = "need at least two images"; return -1; }
    if (depth != 1 && depth != 2) { g_err = "bad input depths, only CV_8UC1 and CV_16UC1 are supported"; return -1; }
    if (cfg->has_thr && cfg->has_step && !(cfg->step > 0.f)) { g_err = "subpixel_step must be positive"; return -1; }

    const int bits = required_bits(n, cfg->full != 0);
    const int words = words_for_bits(bits);
    if (bits < 0 || words == 0) {
        g_err = "input stacks too large, would require " + std::to_string(bits) + " bits";
        return -1;
    }
    const float minvar_n = cfg->has_minvar ? cfg->minvar * (float)n : 0.f;

#define OM(TI, TD) match_impl<TI, TD>(f0, f1, step, n, rows, cols, *cfg, minvar_n, disp16, dispf, corrmap)
    if (depth == 1) {
        switch (words) {
            case 1: OM(uint8_t, uint32_t); break;
            case 2: OM(uint8_t, uint64_t); break;
            case 4: OM(uint8_t, uint128_t); break;
            default: OM(uint8_t, std::bitset<256>); break;
        }
    } else {
        switch (words) {
            case 1: OM(uint16_t, uint32_t); break;
            case 2: OM(uint16_t, uint64_t); break;
            case 4: OM(uint16_t, uint128_t); break;
            default: OM(uint16_t, std::bitset<256>); break;
        }
    }
#undef OM
    *is_float = (cfg->has_thr && cfg->has_step) ? 1 : 0;
    return 0;
}

/* One interpolated subpixel sample as agree.hpp:152-164 forms it: the
 * parabola coefficients of (y0, y1, y2), evaluated at x, rounded and cast to
 * the input type (cast_interp). depth 1: u8, 2: u16. */
int oracle_interp_sample(int depth, int y0, int y1, int y2, float x) {
    const float fy0 = (float)y0, fy1 = (float)y1, fy2 = (float)y2;
    const float a = 0.5f * (fy0 - 2.0f * fy1 + fy2);
    const float b = 0.5f * (-fy0 + fy2);
    const float c = fy1;
    const float v = a * x * x + b * x + c;
    return depth == 1 ? (int)cast_interp<uint8_t>(v) : (int)cast_interp<uint16_t>(v);
}

/* The FP32 subpixel correlation curve of agree.hpp:143-171 for a batch of
 * pixels, without min_variance: for pixel i with left samples
 * left[i * n + t] and right columns right3[(i * 3 + j) * n + t] (j = 0, 1, 2
 * for col1 - 1, col1, col1 + 1), out[i * npts + k] = nxcorr(left,
 * interp(x_k)) at the grid points x_k of agree.hpp:162 (npts of them). */
int oracle_subpixel_curve(int depth, int n, int64_t npix, const void* left, const void* right3,
                          float step, int npts, float* out) {
    if (!(step > 0.f) || n < 2 || n > 128) { g_err = "bad arguments"; return -1; }
    auto run = [&](auto tag) {
        using TInput = decltype(tag);
        const TInput* L = (const TInput*)left;
        const TInput* R = (const TInput*)right3;
        const int64_t chunk = 4096, nchunk = (npix + chunk - 1) / chunk;
        parallel_rows((int)nchunk, [&](int ci) {
            TInput interp[128];
            float a[128], b[128], c[128];
            for (int64_t i = ci * chunk; i < std::min(npix, (ci + 1) * chunk); ++i) {
                const TInput* p0 = L + i * n;
                const TInput *y0 = R + (i * 3) * n, *y1 = y0 + n, *y2 = y1 + n;
                for (int t = 0; t < n; ++t) {
                    a[t] = 0.5f * (y0[t] - 2.0f * y1[t] + y2[t]);
                    b[t] = 0.5f * (-y0[t] + y2[t]);
                    c[t] = y1[t];
                }
                int k = 0;
                for (float x = -1.f; x <= 1.f && k < npts; x += step, ++k) {
                    for (int t = 0; t < n; ++t)
                        interp[t] = cast_interp<TInput>(a[t] * x * x + b[t] * x + c[t]);
                    out[i * npts + k] = nxcorr(p0, interp, (size_t)n, false, 0.f);
                }
            }
        });
    };
    if (depth == 1) run(uint8_t{});
    else if (depth == 2) run(uint16_t{});
    else { g_err = "bad depth"; return -1; }
    return 0;
}

/* The closed-form subpixel correlation the MOMENTS_FAST precision scores
 * (no reference counterpart; the stage it stands in for is agree.hpp:94-185):
 * for pixel i with left samples left[i * n + t] and right columns
 * right3[(i * 3 + j) * n + t] (col1 - 1 + j), the correlation of the left
 * samples with the unrounded parabola samples a x^2 + b x + c (the
 * coefficients of agree.hpp:152-156) at the grid points of agree.hpp:162,
 * in double precision straight from its definition: the interpolated
 * samples are formed and centred, and the three sums taken. out[i * npts + k]
 * (nxc), var1_out[i * npts + k] (the right variance, for the variance test). */
int oracle_subpixel_closed_form(int depth, int n, int64_t npix, const void* left, const void* right3,
                                float step, int npts, double* out, double* var1_out) {
    if (!(step > 0.f) || n < 2 || n > 128) { g_err = "bad arguments"; return -1; }
    std::vector<float> grid;
    for (float x = -1.f; x <= 1.f && (int)grid.size() < npts; x += step) grid.push_back(x);
    auto run = [&](auto tag) {
        using TInput = decltype(tag);
        const TInput* L = (const TInput*)left;
        const TInput* R = (const TInput*)right3;
        for (int64_t i = 0; i < npix; ++i) {
            const TInput* p0 = L + i * n;
            const TInput *y0 = R + (i * 3) * n, *y1 = y0 + n, *y2 = y1 + n;
            double m0 = 0;
            for (int t = 0; t < n; ++t) m0 += p0[t];
            m0 /= n;
            for (int k = 0; k < (int)grid.size(); ++k) {
                const double x = grid[k];
                std::vector<double> v(n);
                double m1 = 0;
                for (int t = 0; t < n; ++t) {
                    const double a = 0.5 * ((double)y0[t] - 2.0 * y1[t] + y2[t]);
                    const double b = 0.5 * ((double)y2[t] - (double)y0[t]);
                    v[t] = a * x * x + b * x + (double)y1[t];
                    m1 += v[t];
                }
                m1 /= n;
                double cov = 0, var0 = 0, var1 = 0;
                for (int t = 0; t < n; ++t) {
                    const double d0 = p0[t] - m0, d1 = v[t] - m1;
                    cov += d0 * d1;
                    var0 += d0 * d0;
                    var1 += d1 * d1;
                }
                out[i * npts + k] = cov / std::sqrt(var0 * var1);
                var1_out[i * npts + k] = var1;
            }
        }
    };
    if (depth == 1) run(uint8_t{});
    else if (depth == 2) run(uint16_t{});
    else { g_err = "bad depth"; return -1; }
    return 0;
}

/* Window-restricted row search (bicos_window above) on descriptor words.
 * center / halfwidth: (rows / 4) x (cols / 4) cell arrays. */
int oracle_search_window(const uint32_t* d0, const uint32_t* d1, int rows, int cols, int words,
                         int flags, int max_lr_diff, const float* center, const int16_t* halfwidth,
                         int16_t* out) {
    if (rows < 4 || cols < 4) { g_err = "windows need at least 4 rows and columns"; return -1; }
    const CellWindows win{center, halfwidth, rows / 4, cols / 4};
    const int w = words;
    auto run = [&](auto tag) {
        using TDesc = decltype(tag);
        std::vector<TDesc> a((size_t)rows * cols), b((size_t)rows * cols);
        for (size_t i = 0; i < a.size(); ++i) {
            a[i] = load_words<TDesc>(d0 + i * w);
            b[i] = load_words<TDesc>(d1 + i * w);
        }
        bicos_window_dispatch<TDesc>(a.data(), b.data(), flags, max_lr_diff, rows, cols, win, out);
    };
    switch (words) {
        case 1: run(uint32_t{}); break;
        case 2: run(uint64_t{}); break;
        case 4: run(uint128_t{}); break;
        case 8: run(std::bitset<256>{}); break;
        default: g_err = "bad words"; return -1;
    }
    return 0;
}

/* BICOS::match of the learned-window path (match_windowed_impl above) on
 * the CPU. Arguments as oracle_match, plus the cell windows (center,
 * halfwidth: (rows / 4) x (cols / 4)), hw_lo (the confident cells'
 * half-width), fallback (0 none, 1 all, 2 unsure) and *n_fallback. */
int oracle_match_windowed(const void* const* f0, const void* const* f1, int64_t step, int depth,
                          int n, int rows, int cols, const oracle::Config* cfg, const float* center,
                          const int16_t* halfwidth, int hw_lo, int fallback, int16_t* disp16,
                          float* dispf, int* is_float, void* corrmap, int64_t* n_fallback) {
    if (n < 2) { g_err = "need at least two images"; return -1; }
    if (depth != 1 && depth != 2) { g_err = "bad input depths"; return -1; }
    if (rows < 4 || cols < 4) { g_err = "windows need at least 4 rows and columns"; return -1; }
    if (cfg->has_thr && cfg->has_step && !(cfg->step > 0.f)) { g_err = "subpixel_step must be positive"; return -1; }
    const int bits = required_bits(n, cfg->full != 0);
    const int words = words_for_bits(bits);
    if (bits < 0 || words == 0) {
        g_err = "input stacks too large, would require " + std::to_string(bits) + " bits";
        return -1;
    }
    const float minvar_n = cfg->has_minvar ? cfg->minvar * (float)n : 0.f;
    const CellWindows win{center, halfwidth, rows / 4, cols / 4};

#define OW(TI, TD)                                                                                  \
    do {                                                                                            \
        if (cfg->dbl)                                                                               \
            match_windowed_impl<TI, TD, double>(f0, f1, step, n, rows, cols, *cfg, minvar_n, win,   \
                                                hw_lo, fallback, disp16, dispf, (double*)corrmap,   \
                                                n_fallback);                                        \
        else                                                                                        \
            match_windowed_impl<TI, TD, float>(f0, f1, step, n, rows, cols, *cfg, minvar_n, win,    \
                                               hw_lo, fallback, disp16, dispf, (float*)corrmap,     \
                                               n_fallback);                                         \
    } while (0)
    if (depth == 1) {
        switch (words) {
            case 1: OW(uint8_t, uint32_t); break;
            case 2: OW(uint8_t, uint64_t); break;
            case 4: OW(uint8_t, uint128_t); break;
            default: OW(uint8_t, std::bitset<256>); break;
        }
    } else {
        switch (words) {
            case 1: OW(uint16_t, uint32_t); break;
            case 2: OW(uint16_t, uint64_t); break;
            case 4: OW(uint16_t, uint128_t); break;
            default: OW(uint16_t, std::bitset<256>); break;
        }
    }
#undef OW
    *is_float = (cfg->has_thr && cfg->has_step) ? 1 : 0;
    return 0;
}

} // extern "C"
