// Procedural whole-slide H&E image, rendered tile by tile on the device.
//
// Every pixel is a pure function of (x, y, seed), so any tile -- and the 64-pixel overlaps of
// neighbouring tiles -- can be produced independently and identically, and a 100000 x 80000 slide
// never has to exist in memory.  Content follows the single-tile generator (synth.cpp):
//  * background RGB (230, 220, 230);
//  * nuclei: one candidate per 40 x 40 cell (1 per 1600 px^2), centre uniform in the cell,
//    semi-axes U[4,14], orientation U[0,pi), colour (90,60,140) +-20 per nucleus; a candidate whose
//    bounding circle meets that of an earlier (lower-index) neighbour cell's candidate is kept
//    with probability 0.03; later nuclei cover earlier ones;
//  * red-blood-cell discs: one per 71 x 71 cell (~1 per 5000 px^2), radius U[3,6], colour
//    (200,80,90), under the nuclei;
//  * stain intensity: every channel's optical darkness (255 - c) is scaled by
//    1 + 0.15 * perlin(x / 2000, y / 2000) (2-D gradient noise, lattice spacing 2000 px);
//  * per-pixel uniform noise +-6 per channel.
// Output is BGR (OpenCV order), row-major, tile-local.
#include <climits>

#include "common.hpp"

namespace nsc {

constexpr int NCELL = 40, RCELL = 71;
constexpr float PERLIN_PX = 2000.f;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t hash3(uint64_t seed, int a, int b, int c) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull;
    z = mix64(z ^ ((uint64_t)(uint32_t)a << 32 | (uint32_t)b));
    return mix64(z + 0x9E3779B97F4A7C15ull * (uint64_t)(c + 1));
}
__device__ __forceinline__ float u01(uint64_t h, int k) {  // k-th 16-bit field -> [0,1)
    return (float)((h >> (16 * k)) & 0xFFFF) * (1.0f / 65536.0f);
}

struct Nuc {
    float cx, cy, a, b, c, s, r;
    int dr, dg, db;
    bool live;
};

__device__ Nuc nucleus_raw(uint64_t seed, int gx, int gy) {
    const uint64_t h = hash3(seed, gx, gy, 1), h2 = hash3(seed, gx, gy, 2);
    Nuc n;
    n.cx = (gx + u01(h, 0)) * NCELL;
    n.cy = (gy + u01(h, 1)) * NCELL;
    n.a = 4.f + 10.f * u01(h, 2);
    n.b = 4.f + 10.f * u01(h, 3);
    const float th = 3.14159265f * u01(h2, 0);
    n.c = cosf(th);
    n.s = sinf(th);
    n.r = fmaxf(n.a, n.b);
    n.dr = (int)(h2 >> 16 & 0xFFFF) % 41 - 20;
    n.dg = (int)(h2 >> 32 & 0xFFFF) % 41 - 20;
    n.db = (int)(h2 >> 48 & 0xFFFF) % 41 - 20;
    n.live = true;
    return n;
}

// a candidate that meets an earlier neighbour (raster order of cells) survives with p = 0.03
__device__ Nuc nucleus(uint64_t seed, int gx, int gy) {
    Nuc n = nucleus_raw(seed, gx, gy);
    bool meets = false;
    for (int dy = -1; dy <= 0; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            if (dy == 0 && dx >= 0) continue;  // earlier cells only
            const Nuc m = nucleus_raw(seed, gx + dx, gy + dy);
            const float ex = m.cx - n.cx, ey = m.cy - n.cy, rr = m.r + n.r;
            if (ex * ex + ey * ey < rr * rr) meets = true;
        }
    if (meets) n.live = u01(hash3(seed, gx, gy, 3), 0) < 0.03f;
    return n;
}

__device__ __forceinline__ float fade(float t) { return t * t * t * (t * (t * 6.f - 15.f) + 10.f); }
__device__ float perlin(uint64_t seed, float x, float y) {
    const int x0 = (int)floorf(x), y0 = (int)floorf(y);
    const float fx = x - x0, fy = y - y0;
    auto grad = [&](int ix, int iy, float dx, float dy) {
        const float ang = 6.2831853f * u01(hash3(seed ^ 0x51DEull, ix, iy, 7), 0);
        return cosf(ang) * dx + sinf(ang) * dy;
    };
    const float n00 = grad(x0, y0, fx, fy), n10 = grad(x0 + 1, y0, fx - 1.f, fy);
    const float n01 = grad(x0, y0 + 1, fx, fy - 1.f), n11 = grad(x0 + 1, y0 + 1, fx - 1.f, fy - 1.f);
    const float u = fade(fx), v = fade(fy);
    const float nx0 = n00 + u * (n10 - n00), nx1 = n01 + u * (n11 - n01);
    return fminf(1.f, fmaxf(-1.f, (nx0 + v * (nx1 - nx0)) * 1.41421356f));
}

// nucleus candidates of the cells that can reach the tile, computed once per tile
__global__ void k_slide_cells(uint64_t seed, int cx0, int cy0, int ncx, int ncy, Nuc* __restrict__ cells) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncx * ncy) return;
    cells[k] = nucleus(seed, cx0 + k % ncx, cy0 + k / ncx);
}

__global__ void k_slide_tile(uint8_t* __restrict__ bgr, int x0, int y0, int w, int h, uint64_t seed,
                             const Nuc* __restrict__ cells, int cx0, int cy0, int ncx) {
    const int lx = blockIdx.x * blockDim.x + threadIdx.x, ly = blockIdx.y;
    if (lx >= w || ly >= h) return;
    const int x = x0 + lx, y = y0 + ly;
    const float px = x + 0.5f, py = y + 0.5f;
    float r = 230.f, g = 220.f, b = 230.f;
    // red-blood-cell discs (3x3 cells around the pixel; radius < cell)
    {
        const int cx = x / RCELL, cy = y / RCELL;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const uint64_t hh = hash3(seed ^ 0xB10Dull, cx + dx, cy + dy, 4);
                const float ccx = (cx + dx + u01(hh, 0)) * RCELL, ccy = (cy + dy + u01(hh, 1)) * RCELL;
                const float rad = 3.f + 3.f * u01(hh, 2);
                const float ex = px - ccx, ey = py - ccy;
                if (ex * ex + ey * ey <= rad * rad) { r = 200.f; g = 80.f; b = 90.f; }
            }
    }
    // nuclei (3x3 cells; semi-axes < cell); later cells cover earlier ones
    {
        const int cx = x / NCELL, cy = y / NCELL;
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const Nuc n = cells[(cy + dy - cy0) * ncx + (cx + dx - cx0)];
                if (!n.live) continue;
                const float ex = px - n.cx, ey = py - n.cy;
                const float u = (ex * n.c + ey * n.s) / n.a, v = (-ex * n.s + ey * n.c) / n.b;
                if (u * u + v * v <= 1.f) { r = 90.f + n.dr; g = 60.f + n.dg; b = 140.f + n.db; }
            }
    }
    const float f = 1.f + 0.15f * perlin(seed, px / PERLIN_PX, py / PERLIN_PX);
    const uint64_t nz = hash3(seed ^ 0xA5A5ull, x, y, 5);
    auto chan = [&](float c, int k) {
        const float v = 255.f - (255.f - c) * f + (float)((int)((nz >> (16 * k)) % 13) - 6);
        return (uint8_t)fminf(255.f, fmaxf(0.f, rintf(v)));
    };
    uint8_t* o = bgr + 3 * ((size_t)ly * w + lx);
    o[0] = chan(b, 0);
    o[1] = chan(g, 1);
    o[2] = chan(r, 2);
}

}  // namespace nsc

extern "C" int nsc_synth_slide_tile(uint8_t* bgr, int x0, int y0, int w, int h, uint64_t seed, void* stream) {
    NSC_ARG(bgr && w > 0 && h > 0 && x0 >= 0 && y0 >= 0 && (long long)x0 + w < INT_MAX && (long long)y0 + h < INT_MAX,
            "synth_slide_tile");
    NSC_CHECK(nsc::ensure_init());
    hipStream_t s = nsc::as_stream(stream);
    // cells whose nucleus can touch the tile: one cell of margin on each side
    const int cx0 = x0 / nsc::NCELL - 1, cy0 = y0 / nsc::NCELL - 1;
    const int ncx = (x0 + w - 1) / nsc::NCELL + 1 - cx0 + 1, ncy = (y0 + h - 1) / nsc::NCELL + 1 - cy0 + 1;
    nsc::Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(nsc::Nuc) * (size_t)ncx * ncy, s));
    hipLaunchKernelGGL(nsc::k_slide_cells, dim3(nsc::cdiv((size_t)ncx * ncy, 256)), dim3(256), 0, s, seed, cx0, cy0, ncx, ncy,
                       sc.as<nsc::Nuc>());
    hipLaunchKernelGGL(nsc::k_slide_tile, dim3(nsc::cdiv(w, 256), h), dim3(256), 0, s, bgr, x0, y0, w, h, seed,
                       sc.as<nsc::Nuc>(), cx0, cy0, ncx);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}
