// bicos_cli: a small front-end of the C++ API (include/bicos/match.hpp).
//
// Reads two raw frame-planar stacks (n x rows x cols, uint8 or uint16),
// maps its options to BICOS::Config the way the reference's bicos-cli does
// (src/cli.cpp:60-77 and 137-166: FULL mode unless --limited, threshold 0.75
// by default and disabled when <= 0, variance 1.0 by default and ignored when
// <= 0, --lr-maxdiff selects Consistency), runs BICOS::match on the GPU and
// writes the disparity (and optionally the correlation map) as raw files.
// Image codecs and the point-cloud output of the reference CLI are not part
// of this tool.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "bicos/match.hpp"

using namespace BICOS;

static std::vector<char> read_file(const std::string& path, size_t bytes) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<char> buf(bytes);
    f.read(buf.data(), (std::streamsize)bytes);
    if ((size_t)f.gcount() != bytes) throw std::runtime_error("short read from " + path);
    return buf;
}

static void write_file(const std::string& path, const void* p, size_t bytes) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f.write((const char*)p, (std::streamsize)bytes);
}

int main(int argc, char** argv) {
    std::string left, right, out = "bicosdisp.raw", corr;
    int n = 0, rows = 0, cols = 0, depth = 8;
    float thr = 0.75f, var = 1.0f;
    bool has_var = false;  // cxxopts count(): set only when -v is given
    bool has_step = false, limited = false, dbl = false, no_dupes = false, has_lr = false;
    float step = 0.0f;
    int lr = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) throw std::invalid_argument("missing value for " + a);
            return argv[++i];
        };
        if (a == "--left") left = next();
        else if (a == "--right") right = next();
        else if (a == "-o" || a == "--out") out = next();
        else if (a == "--corrmap") corr = next();
        else if (a == "-n" || a == "--stacksize") n = std::stoi(next());
        else if (a == "--rows") rows = std::stoi(next());
        else if (a == "--cols") cols = std::stoi(next());
        else if (a == "--depth") depth = std::stoi(next());
        else if (a == "-t" || a == "--threshold") thr = std::stof(next());
        else if (a == "-v" || a == "--variance") { has_var = true; var = std::stof(next()); }
        else if (a == "-s" || a == "--step") { has_step = true; step = std::stof(next()); }
        else if (a == "-m" || a == "--lr-maxdiff") { has_lr = true; lr = std::stoi(next()); }
        else if (a == "--limited") limited = true;
        else if (a == "--double") dbl = true;
        else if (a == "--no-dupes") no_dupes = true;
        else throw std::invalid_argument("unknown option " + a);
    }
    if (left.empty() || right.empty() || n < 1 || rows < 1 || cols < 1 || (depth != 8 && depth != 16))
        throw std::invalid_argument("need --left --right -n --rows --cols --depth 8|16");

    // src/cli.cpp:137-166
    Config c;
    c.nxcorr_threshold = thr;
    c.mode = TransformMode::FULL;
    if (c.nxcorr_threshold.value() <= 0.0f) c.nxcorr_threshold = std::nullopt;
    const bool need_corrmap = !corr.empty();
    if (need_corrmap && !c.nxcorr_threshold.has_value()) c.nxcorr_threshold = -1.0f;
    if (has_step) c.subpixel_step = step;
    if (limited) c.mode = TransformMode::LIMITED;
    if (has_var && var > 0.0f) c.min_variance = var;
    if (dbl) c.precision = Precision::DOUBLE;
    if (has_lr) c.variant = Variant::Consistency{lr, no_dupes};

    const int type = depth == 8 ? U8 : U16;
    const size_t es = depth / 8, plane = (size_t)rows * cols * es;
    const std::vector<char> l = read_file(left, plane * n), r = read_file(right, plane * n);

    std::vector<Image> s0(n), s1(n);
    for (int t = 0; t < n; ++t) {
        s0[t].upload(l.data() + t * plane, rows, cols, type);
        s1[t].upload(r.data() + t * plane, rows, cols, type);
    }

    Image disp, cmap;
    auto t0 = std::chrono::steady_clock::now();
    match(s0, s1, disp, c, need_corrmap ? &cmap : nullptr);
    std::vector<char> host((size_t)rows * cols * disp.elemSize());
    disp.download(host.data());
    auto t1 = std::chrono::steady_clock::now();
    write_file(out, host.data(), host.size());
    if (need_corrmap && !cmap.empty()) {
        std::vector<char> ch((size_t)rows * cols * cmap.elemSize());
        cmap.download(ch.data());
        write_file(corr, ch.data(), ch.size());
    }
    std::printf("{\"disparity_type\": %d, \"corrmap_type\": %d, \"match_ms\": %.3f}\n", disp.type(),
                cmap.empty() ? -1 : cmap.type(),
                std::chrono::duration<double, std::milli>(t1 - t0).count());
    return 0;
}
