// synth.cpp -- deterministic synthetic pictures for tests and benchmarks.
//
// Textures are uniform noise smoothed by a separable Gaussian (sigma 1.5,
// radius 4) plus flat rectangles carrying their own noise; references are
// the current picture warped by a motion field (affine zoom + pan plus
// piecewise-constant object motion) quantised to quarter pel and sampled
// bilinearly, then noised.  8-bit values are stored minus 128, the sample
// convention of the reference encoder (libdirac_common/pic_io.cpp:396-401).
// The PRNG is SplitMix64 with Box-Muller normals, so every picture is a pure
// function of its seed on any host.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/dme.h"
#include "dme_internal.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed ^ 0x9E3779B97F4A7C15ull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline int clamp255(double v) {
    long r = std::lround(v);
    return r < 0 ? 0 : (r > 255 ? 255 : (int)r);
}

}  // namespace

extern "C" int dme_synth_texture(uint64_t seed, int width, int height, int nrects, int16_t *out_y) try {
    if (width <= 0 || height <= 0 || !out_y || nrects < 0) return dme::set_error(DME_ERR_INVALID, "bad synth_texture args");
    Rng rng(seed);
    const int W = width, H = height;
    std::vector<float> a((size_t)W * H), b((size_t)W * H);
    for (auto &v : a) v = (float)(rng.uniform() * 255.0);
    // separable Gaussian, sigma 1.5, radius 4, clamp-to-edge
    float k[9];
    float ks = 0.f;
    for (int i = -4; i <= 4; ++i) ks += (k[i + 4] = (float)std::exp(-(i * i) / (2.0 * 1.5 * 1.5)));
    for (auto &v : k) v /= ks;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float s = 0.f;
            for (int i = -4; i <= 4; ++i) {
                int xx = x + i;
                xx = xx < 0 ? 0 : (xx >= W ? W - 1 : xx);
                s += k[i + 4] * a[(size_t)y * W + xx];
            }
            b[(size_t)y * W + x] = s;
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float s = 0.f;
            for (int i = -4; i <= 4; ++i) {
                int yy = y + i;
                yy = yy < 0 ? 0 : (yy >= H ? H - 1 : yy);
                s += k[i + 4] * b[(size_t)yy * W + x];
            }
            a[(size_t)y * W + x] = s;
        }
    // blurred uniform noise has a small spread around 127.5: stretch it to
    // a usable contrast (about +-48 at one sigma)
    for (auto &v : a) v = 127.5f + (v - 127.5f) * 3.5f;
    for (int r = 0; r < nrects; ++r) {
        const int rw = (int)(W * (0.08 + 0.22 * rng.uniform()));
        const int rh = (int)(H * (0.08 + 0.22 * rng.uniform()));
        const int x0 = (int)((W - rw) * rng.uniform()), y0 = (int)((H - rh) * rng.uniform());
        const double level = 30.0 + 195.0 * rng.uniform(), sig = 2.0 + 6.0 * rng.uniform();
        for (int y = y0; y < y0 + rh; ++y)
            for (int x = x0; x < x0 + rw; ++x) a[(size_t)y * W + x] = (float)(level + sig * rng.normal());
    }
    for (size_t i = 0; i < a.size(); ++i) out_y[i] = (int16_t)(clamp255(a[i]) - 128);
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_synth_warp(const int16_t *cur, int width, int height, const dme_synth_motion *m,
                              int16_t *out_ref) try {
    if (!cur || !out_ref || !m || width <= 0 || height <= 0 || (m->nobjects > 0 && !m->objects))
        return dme::set_error(DME_ERR_INVALID, "bad synth_warp args");
    const int W = width, H = height;
    const double cx = 0.5 * (W - 1), cy = 0.5 * (H - 1);
    Rng rng(m->noise_seed);
    auto S = [&](int y, int x) -> double {
        x = x < 0 ? 0 : (x >= W ? W - 1 : x);
        y = y < 0 ? 0 : (y >= H ? H - 1 : y);
        return (double)cur[(size_t)y * W + x] + 128.0;
    };
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double dx = (m->zoom - 1.0) * (x - cx) + m->pan_x;
            double dy = (m->zoom - 1.0) * (y - cy) + m->pan_y;
            for (int o = 0; o < m->nobjects; ++o) {
                const dme_synth_object &ob = m->objects[o];
                if (x >= ob.x0 && x < ob.x1 && y >= ob.y0 && y < ob.y1) {
                    dx = ob.dx + m->pan_x;
                    dy = ob.dy + m->pan_y;
                }
            }
            // quarter-pel quantisation, then bilinear sampling
            dx = std::floor(dx * 4.0 + 0.5) / 4.0;
            dy = std::floor(dy * 4.0 + 0.5) / 4.0;
            const double sx = x + dx, sy = y + dy;
            const int ix = (int)std::floor(sx), iy = (int)std::floor(sy);
            const double fx = sx - ix, fy = sy - iy;
            const double v = (1 - fx) * (1 - fy) * S(iy, ix) + fx * (1 - fy) * S(iy, ix + 1) +
                             (1 - fx) * fy * S(iy + 1, ix) + fx * fy * S(iy + 1, ix + 1);
            const double n = m->noise_sigma > 0 ? m->noise_sigma * rng.normal() : 0.0;
            out_ref[(size_t)y * W + x] = (int16_t)(clamp255(v + n) - 128);
        }
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}

extern "C" int dme_synth_chroma(const int16_t *y, int width, int height, uint64_t seed, int16_t *out_u,
                                int16_t *out_v) try {
    if (!y || !out_u || !out_v || width < 2 || height < 2) return dme::set_error(DME_ERR_INVALID, "bad synth_chroma args");
    Rng rng(seed);
    const int CW = width >> 1, CH = height >> 1;
    const double ou = 40.0 * (rng.uniform() - 0.5), ov = 40.0 * (rng.uniform() - 0.5);
    for (int j = 0; j < CH; ++j)
        for (int i = 0; i < CW; ++i) {
            const double m = 0.25 * ((double)y[(size_t)(2 * j) * width + 2 * i] + y[(size_t)(2 * j) * width + 2 * i + 1] +
                                     y[(size_t)(2 * j + 1) * width + 2 * i] + y[(size_t)(2 * j + 1) * width + 2 * i + 1]) +
                             128.0;
            out_u[(size_t)j * CW + i] = (int16_t)(clamp255(128.0 + 0.5 * (m - 128.0) + ou) - 128);
            out_v[(size_t)j * CW + i] = (int16_t)(clamp255(128.0 - 0.4 * (m - 128.0) + ov) - 128);
        }
    return DME_OK;
} catch (...) {
    return dme::exception_status();
}
