// Exact Euclidean distance transform of a u8 image: distance of every nonzero pixel to the
// nearest zero pixel (pixels outside the image are not zeros), as cv::distanceTransform(src, dst,
// CV_DIST_L2, CV_DIST_MASK_PRECISE) is called at HistologicalEntities.cpp:359.
//
// Output value = sqrtf(float(dx^2) + float(g^2)) for the minimising column offset dx, where g is
// the vertical distance of that column to its nearest zero: the same float expression OpenCV's
// precise transform evaluates (sqr_tab[|q-p|] + f[p]); rows without any zero in any column give
// sqrtf(0 + 1e15f) as OpenCV's "infinite" column value does.
//
// Column pass (segmented for parallelism): per 64-row column segment, first/last zero; per column,
// carry of the nearest zero above/below each segment; per segment, the final vertical distance g
// (uint16, 0xFFFF = no zero in the column) from a 64-bit mask of the segment's zero rows.
// Row pass: one workgroup per row, g staged in LDS; each pixel searches outward while dx^2 is
// below its best squared distance (work ~ distance value).  A pixel whose search would exceed
// SEARCH_CAP columns flags its row; flagged rows are recomputed with the exact integer
// lower-envelope algorithm (one thread per row, O(w)).
#include "common.hpp"

namespace nsc {
namespace {

constexpr int SEG = 64;
constexpr uint16_t NONE = 0xFFFF;
constexpr int SEARCH_CAP = 1024;

__global__ void k_zero_i32(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// first / last zero row inside each segment of each column (-1 = none)
__global__ void k_seg_zeros(const uint8_t* __restrict__ img, int w, int h, int nseg, int16_t* __restrict__ fz,
                            int16_t* __restrict__ lz) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int sg = blockIdx.y;
    if (x >= w) return;
    const int y0 = sg * SEG, n = min(h, y0 + SEG) - y0;
    // 64 loads in flight (rows past the image read the last row and are masked off), a bit per zero row
    uint32_t v[SEG];
#pragma unroll
    for (int j = 0; j < SEG; ++j) v[j] = img[(size_t)min(y0 + j, h - 1) * w + x];
    unsigned long long zm = 0;
#pragma unroll
    for (int j = 0; j < SEG; ++j) zm |= (unsigned long long)(v[j] == 0 && j < n) << j;
    fz[(size_t)sg * w + x] = (int16_t)(zm ? __ffsll((long long)zm) - 1 : -1);
    lz[(size_t)sg * w + x] = (int16_t)(zm ? 63 - __clzll((long long)zm) : -1);
}

// per column: row of the nearest zero strictly before / after each segment (-1 / INT_MAX-ish)
__global__ void k_seg_carry(const int16_t* __restrict__ fz, const int16_t* __restrict__ lz, int w, int nseg,
                            int32_t* __restrict__ above, int32_t* __restrict__ below) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w) return;
    // sixteen segments' entries loaded at a time (clamped indices), then scanned
    constexpr int CH = 16;
    int last = -1;
    for (int s0 = 0; s0 < nseg; s0 += CH) {
        int l[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) l[j] = lz[(size_t)min(s0 + j, nseg - 1) * w + x];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int s = s0 + j;
            if (s >= nseg) break;
            above[(size_t)s * w + x] = last;
            if (l[j] >= 0) last = s * SEG + l[j];
        }
    }
    int first = -1;
    for (int s1 = nseg - 1; s1 >= 0; s1 -= CH) {
        int f[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) f[j] = fz[(size_t)max(s1 - j, 0) * w + x];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const int s = s1 - j;
            if (s < 0) break;
            below[(size_t)s * w + x] = first;
            if (f[j] >= 0) first = s * SEG + f[j];
        }
    }
}

__global__ void k_col_dist(const uint8_t* __restrict__ img, int w, int h, const int32_t* __restrict__ above,
                           const int32_t* __restrict__ below, uint16_t* __restrict__ g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int sg = blockIdx.y;
    if (x >= w) return;
    const int y0 = sg * SEG, n = min(h, y0 + SEG) - y0;
    // the segment's column at once: 64 loads in flight (rows past the image read the last row and
    // are masked off), then a bit per zero row; each row's nearest zero above / below inside the
    // segment is the highest set bit at or below it / the lowest at or above it
    uint32_t v[SEG];
#pragma unroll
    for (int j = 0; j < SEG; ++j) v[j] = img[(size_t)min(y0 + j, h - 1) * w + x];
    unsigned long long zm = 0;
#pragma unroll
    for (int j = 0; j < SEG; ++j) zm |= (unsigned long long)(v[j] == 0 && j < n) << j;
    const int up = above[(size_t)sg * w + x];  // row of the nearest zero above the segment (-1: none)
    const int dn = below[(size_t)sg * w + x];  // row of the nearest zero below it (-1: none)
#pragma unroll
    for (int j = 0; j < SEG; ++j) {
        if (j >= n) break;
        const unsigned long long le = zm & (j == SEG - 1 ? ~0ull : ((2ull << j) - 1)), ge = zm >> j;
        const int zu = le ? y0 + 63 - __clzll((long long)le) : up;
        const int zd = ge ? y0 + j + __ffsll((long long)ge) - 1 : dn;
        const int y = y0 + j;
        const uint32_t du = zu >= 0 ? (uint32_t)(y - zu) : NONE, dd = zd >= 0 ? (uint32_t)(zd - y) : NONE;
        g[(size_t)y * w + x] = (uint16_t)min(du, dd);
    }
}

__device__ __forceinline__ float edt_value(long long dx2, long long g) {
    const float fg = (g == NONE) ? 1e15f : (float)(g * g);
    return sqrtf((float)dx2 + fg);
}

__global__ __launch_bounds__(256) void k_row_search(const uint16_t* __restrict__ g, int w, int h, float* __restrict__ out,
                                                    int* __restrict__ fallback_rows, int* __restrict__ nfallback) {
    extern __shared__ uint16_t grow[];
    const int y = blockIdx.x;
    const uint16_t* gr = g + (size_t)y * w;
    for (int x = threadIdx.x; x < w; x += blockDim.x) grow[x] = gr[x];
    __syncthreads();
    __shared__ int need;
    if (threadIdx.x == 0) need = 0;
    __syncthreads();
    for (int x = threadIdx.x; x < w; x += blockDim.x) {
        // squared distances in 32 bits, exactly: g <= 65534 so g^2 < 2^32 and UINT_MAX marks "no
        // zero"; a candidate dx^2 + v^2 is compared as v^2 < best - dx^2 (dx^2 < best), which can
        // neither overflow nor underflow, and a winner is below the old best
        const uint32_t gx = grow[x];
        uint32_t best = (gx == NONE) ? UINT_MAX : gx * gx;
        int bdx = 0;
        uint32_t bg = gx;
        if (best != 0) {
            int dx = 1;
            for (; dx <= SEARCH_CAP; ++dx) {
                const uint32_t dx2 = (uint32_t)(dx * dx);
                if (dx2 >= best) break;
                if (x - dx < 0 && x + dx >= w) { dx = SEARCH_CAP + 2; break; }
                if (x - dx >= 0) {
                    const uint32_t v = grow[x - dx];
                    if (v != NONE && v * v < best - dx2) { best = dx2 + v * v; bdx = dx; bg = v; }
                }
                if (x + dx < w) {
                    const uint32_t v = grow[x + dx];
                    if (v != NONE && v * v < best - dx2) { best = dx2 + v * v; bdx = dx; bg = v; }
                }
            }
            if (dx == SEARCH_CAP + 1) need = 1;  // search cut short: exact envelope later
        }
        float o;
        if (best == UINT_MAX) o = sqrtf(1e15f);
        else o = edt_value((long long)bdx * bdx, bg);
        out[(size_t)y * w + x] = o;
    }
    __syncthreads();
    if (threadIdx.x == 0 && need) fallback_rows[atomicAdd(nfallback, 1)] = y;
}

// exact lower envelope of the parabolas (x - p)^2 + g(p)^2 over the columns with a zero
// (Felzenszwalb-Huttenlocher with exact integer intersection tests); one workgroup per flagged
// row, the apex stack in LDS, built by one lane and evaluated by all
__global__ __launch_bounds__(256) void k_row_envelope(const uint16_t* __restrict__ g, int w, int h, float* __restrict__ out,
                                                      const int* __restrict__ rows, const int* __restrict__ nrows) {
    extern __shared__ int v[];
    __shared__ int kk;
    if ((int)blockIdx.x >= min(*nrows, h)) return;
    const int y = rows[blockIdx.x];
    if (y < 0 || y >= h) return;
    const uint16_t* gr = g + (size_t)y * w;
    auto f = [&](int p) { return (long long)gr[p] * gr[p]; };
    if (threadIdx.x == 0) {
        int k = -1;
        for (int q = 0; q < w; ++q) {
            if (gr[q] == NONE) continue;
            while (k >= 1) {
                const int p = v[k], r = v[k - 1];
                // drop p when the (r, q) intersection is not right of the (r, p) intersection
                const long long nq = (f(q) + (long long)q * q) - (f(r) + (long long)r * r);
                const long long np = (f(p) + (long long)p * p) - (f(r) + (long long)r * r);
                if (nq * (long long)(p - r) <= np * (long long)(q - r)) --k;
                else break;
            }
            v[++k] = q;
        }
        kk = k;
    }
    __syncthreads();
    const int k = kk;
    for (int x = threadIdx.x; x < w; x += blockDim.x) {
        if (k < 0) { out[(size_t)y * w + x] = sqrtf(1e15f); continue; }
        // binary search for the apex whose parabola is lowest at x (the envelope order is monotone)
        int lo = 0, hi = k;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const int p = v[mid], q = v[mid + 1];
            const long long dp = (long long)(x - p) * (x - p) + f(p);
            const long long dq = (long long)(x - q) * (x - q) + f(q);
            if (dq <= dp) lo = mid + 1;
            else hi = mid;
        }
        const int p = v[lo];
        out[(size_t)y * w + x] = edt_value((long long)(x - p) * (x - p), gr[p]);
    }
}

}  // namespace

int edt_impl(const uint8_t* img, float* out, int w, int h, hipStream_t s) {
    // size limits are checked before anything is queued on the stream
    if (h >= 0xFFFF) {
        set_error("distance transform supports heights below 65535");
        return NSC_ERR_LIMIT;
    }
    if (w > 32768) {
        set_error("distance transform supports widths up to 32768");
        return NSC_ERR_LIMIT;
    }
    const int nseg = (int)cdiv(h, SEG);
    const size_t n = (size_t)w * h;
    auto carve = [&](Carver& c, int32_t*& above, int32_t*& below, int*& rows, int*& nrows, uint16_t*& g, int16_t*& fz,
                     int16_t*& lz) {
        above = c.take<int32_t>((size_t)nseg * w);
        below = c.take<int32_t>((size_t)nseg * w);
        rows = c.take<int>(h);
        nrows = c.take<int>(1);
        g = c.take<uint16_t>(n);
        fz = c.take<int16_t>((size_t)nseg * w);
        lz = c.take<int16_t>((size_t)nseg * w);
    };
    int32_t *above, *below;
    int *rows, *nrows;
    uint16_t* g;
    int16_t *fz, *lz;
    Carver measure(nullptr);
    carve(measure, above, below, rows, nrows, g, fz, lz);
    Scratch sc;
    NSC_CHECK(sc.alloc(measure.size(), s));
    Carver cut(sc.p);
    carve(cut, above, below, rows, nrows, g, fz, lz);
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, nrows, 1);
    dim3 gb(cdiv(w, 64), nseg);
    hipLaunchKernelGGL(k_seg_zeros, gb, dim3(64), 0, s, img, w, h, nseg, fz, lz);
    hipLaunchKernelGGL(k_seg_carry, dim3(cdiv(w, 64)), dim3(64), 0, s, fz, lz, w, nseg, above, below);
    hipLaunchKernelGGL(k_col_dist, gb, dim3(64), 0, s, img, w, h, above, below, g);
    const size_t lds = sizeof(uint16_t) * w;
    hipLaunchKernelGGL(k_row_search, dim3(h), dim3(256), lds, s, g, w, h, out, rows, nrows);
    // rows whose search was cut short (distances beyond SEARCH_CAP): exact envelope; the launch
    // is unconditional and its blocks beyond the flagged count exit at once (no host sync)
    const size_t lds2 = sizeof(int) * w;
    if (lds2 > 64 * 1024) NSC_HIP(hipFuncSetAttribute((const void*)k_row_envelope,
                                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2));
    hipLaunchKernelGGL(k_row_envelope, dim3(h), dim3(256), lds2, s, g, w, h, out, rows, nrows);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

}  // namespace nsc

using namespace nsc;

extern "C" int nsc_dist_transform_u8(const uint8_t* img, float* out, int w, int h, void* stream) {
    NSC_ARG(img && out && w > 0 && h > 0, "dist_transform_u8");
    NSC_CHECK(ensure_init());
    return edt_impl(img, out, w, h, as_stream(stream));
}
