// Host helpers of the image I/O used by the command-line front-end
// (libbicos_amd/cli.py): PNG scanline unfiltering, the step of PNG decoding
// that is sequential within a row (Sub, Average and Paeth predict each byte
// from the bytes already decoded). Filter definitions: PNG specification
// (ISO/IEC 15948), section 9.
#include <cstddef>
#include <cstdint>
#include <cstdlib>

namespace {

inline uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

}  // namespace

extern "C" {

// in: rows * (1 + stride) bytes (filter byte, then the filtered scanline);
// out: rows * stride bytes; bpp: bytes per complete pixel (at least 1).
// Returns 0, or -1 on an unknown filter type.
int bicos_png_unfilter(const uint8_t* in, size_t rows, size_t stride, int bpp, uint8_t* out) {
    for (size_t r = 0; r < rows; ++r) {
        const uint8_t ft = in[r * (stride + 1)];
        const uint8_t* src = in + r * (stride + 1) + 1;
        uint8_t* dst = out + r * stride;
        const uint8_t* up = r ? out + (r - 1) * stride : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)bpp ? dst[i - bpp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= (size_t)bpp) ? up[i - bpp] : 0;
            int v;
            switch (ft) {
                case 0: v = src[i]; break;
                case 1: v = src[i] + a; break;
                case 2: v = src[i] + b; break;
                case 3: v = src[i] + ((a + b) >> 1); break;
                case 4: v = src[i] + paeth(a, b, c); break;
                default: return -1;
            }
            dst[i] = (uint8_t)v;
        }
    }
    return 0;
}

}  // extern "C"
