// Seeded synthetic inputs for the segmentation benchmarks and tests (host code).
//
//  * nsc_synth_mask: grayscale "mask" tile: background 220, random ellipses with intensity
//    U[60,120], semi-axes U[4,12], random orientation, about one per 400 px^2, Gaussian noise
//    sigma 8, clipped to [0,255].  The matching marker is sat(mask - 10) (nsc_synth_marker).
//  * nsc_synth_he: BGR H&E-like tile: background RGB (230,220,230) +-6 uniform noise; nuclei are
//    ellipses with semi-axes U[4,14] and colour (90,60,140) +-20 per nucleus plus +-6 per pixel,
//    about one per 1600 px^2, overlapping an earlier nucleus with probability ~3%; red-blood-cell
//    like discs (200,80,90), radius U[3,6], one per 5000 px^2.
//
// All randomness comes from splitmix64 streams derived from the seed, so a (seed, w, h) triple
// always yields the same bytes.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "nscale_hip.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double uni(double a, double b) { return a + (b - a) * uni(); }
    int irange(int a, int b) { return a + (int)(next() % (uint64_t)(b - a + 1)); }  // [a,b]
    double gauss() {
        double u1 = uni(), u2 = uni();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Ellipse {
    double cx, cy, a, b, c, s;
    int x0, x1, y0, y1;
    Ellipse(double cx_, double cy_, double a_, double b_, double th, int w, int h)
        : cx(cx_), cy(cy_), a(a_), b(b_), c(std::cos(th)), s(std::sin(th)) {
        double r = a > b ? a : b;
        x0 = (int)std::floor(cx - r); x1 = (int)std::ceil(cx + r);
        y0 = (int)std::floor(cy - r); y1 = (int)std::ceil(cy + r);
        if (x0 < 0) x0 = 0;
        if (y0 < 0) y0 = 0;
        if (x1 > w - 1) x1 = w - 1;
        if (y1 > h - 1) y1 = h - 1;
    }
    bool inside(int x, int y) const {
        double dx = x - cx, dy = y - cy;
        double u = (dx * c + dy * s) / a, v = (-dx * s + dy * c) / b;
        return u * u + v * v <= 1.0;
    }
};

inline uint8_t clip8(double v) { return v < 0 ? 0 : (v > 255 ? 255 : (uint8_t)v); }

}  // namespace

extern "C" {

int nsc_synth_mask(uint8_t* out, int w, int h, uint64_t seed) {
    if (!out || w <= 0 || h <= 0) return NSC_ERR_ARG;
    Rng rng(seed);
    std::vector<uint8_t> base((size_t)w * h, 220);
    long n_ell = std::lround((double)w * h / 400.0);
    for (long e = 0; e < n_ell; ++e) {
        double cx = rng.uni(0, w), cy = rng.uni(0, h);
        double a = rng.uni(4, 12), b = rng.uni(4, 12), th = rng.uni(0, 3.141592653589793);
        uint8_t val = (uint8_t)rng.irange(60, 120);
        Ellipse el(cx, cy, a, b, th, w, h);
        for (int y = el.y0; y <= el.y1; ++y)
            for (int x = el.x0; x <= el.x1; ++x)
                if (el.inside(x, y)) base[(size_t)y * w + x] = val;
    }
    Rng nrng(seed ^ 0x5DEECE66Dull);
    for (size_t i = 0; i < (size_t)w * h; ++i) out[i] = clip8(std::nearbyint(base[i] + 8.0 * nrng.gauss()));
    return NSC_OK;
}

int nsc_synth_marker(const uint8_t* mask, uint8_t* marker, size_t n) {
    if (!mask || !marker) return NSC_ERR_ARG;
    for (size_t i = 0; i < n; ++i) marker[i] = mask[i] > 10 ? (uint8_t)(mask[i] - 10) : 0;
    return NSC_OK;
}

int nsc_synth_he(uint8_t* bgr, int w, int h, uint64_t seed) {
    if (!bgr || w <= 0 || h <= 0) return NSC_ERR_ARG;
    Rng rng(seed);
    const size_t n = (size_t)w * h;
    // RGB working planes, then interleaved as BGR at the end
    std::vector<uint8_t> occ(n, 0);
    std::vector<int16_t> base(n * 3);
    for (size_t i = 0; i < n; ++i) { base[3 * i] = 230; base[3 * i + 1] = 220; base[3 * i + 2] = 230; }
    // red-blood-cell-like discs
    long n_rbc = std::lround((double)n / 5000.0);
    for (long e = 0; e < n_rbc; ++e) {
        double r = rng.uni(3, 6);
        Ellipse el(rng.uni(0, w), rng.uni(0, h), r, r, 0.0, w, h);
        for (int y = el.y0; y <= el.y1; ++y)
            for (int x = el.x0; x <= el.x1; ++x)
                if (el.inside(x, y)) {
                    size_t i = (size_t)y * w + x;
                    base[3 * i] = 200; base[3 * i + 1] = 80; base[3 * i + 2] = 90; occ[i] = 2;
                }
    }
    // nuclei
    long n_nuc = std::lround((double)n / 1600.0);
    for (long e = 0; e < n_nuc; ++e) {
        for (int tries = 0; tries < 16; ++tries) {
            Ellipse el(rng.uni(0, w), rng.uni(0, h), rng.uni(4, 14), rng.uni(4, 14), rng.uni(0, 3.141592653589793), w, h);
            bool overlap = false;
            for (int y = el.y0; y <= el.y1 && !overlap; ++y)
                for (int x = el.x0; x <= el.x1; ++x)
                    if (occ[(size_t)y * w + x] == 1 && el.inside(x, y)) { overlap = true; break; }
            double accept = rng.uni();
            if (overlap && accept >= 0.03) continue;
            int cr = 90 + rng.irange(-20, 20), cg = 60 + rng.irange(-20, 20), cb = 140 + rng.irange(-20, 20);
            for (int y = el.y0; y <= el.y1; ++y)
                for (int x = el.x0; x <= el.x1; ++x)
                    if (el.inside(x, y)) {
                        size_t i = (size_t)y * w + x;
                        base[3 * i] = (int16_t)cr; base[3 * i + 1] = (int16_t)cg; base[3 * i + 2] = (int16_t)cb; occ[i] = 1;
                    }
            break;
        }
    }
    Rng nrng(seed ^ 0xA5A5A5A5DEADBEEFull);
    for (size_t i = 0; i < n; ++i) {
        uint64_t r = nrng.next();
        int nr = (int)(r % 13) - 6, ng = (int)((r >> 16) % 13) - 6, nb = (int)((r >> 32) % 13) - 6;
        bgr[3 * i + 2] = clip8(base[3 * i] + nr);
        bgr[3 * i + 1] = clip8(base[3 * i + 1] + ng);
        bgr[3 * i + 0] = clip8(base[3 * i + 2] + nb);
    }
    return NSC_OK;
}

}  // extern "C"
