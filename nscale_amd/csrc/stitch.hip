// Whole-slide stitching of overlapping tiles.  nu-segment cuts a slide into tiles segmented
// independently (nu-segment.cpp:274-433 hands out one file per worker); tiles cut with an overlap
// see every nucleus near a seam twice.  Here each tile owns a core: its pixels minus half of the
// overlap on every side that has a neighbour, so that the cores of a regular grid (stride
// tile - overlap) partition the slide.  A nucleus of a tile is kept iff the centre of its box
// (in doubled integer coordinates, no rounding) lies in the tile's core, and kept nuclei get
// consecutive slide-wide ids in label order.  A nucleus smaller than half the overlap lies whole
// in the tile that owns its centre, so each is kept once and never cut; a kept nucleus whose box
// touches a seam-side tile edge is counted as possibly cut (the overlap was too small for it).
//
// Kernels: k_stitch_mark (one thread per box), k_stitch_scan (one workgroup: keep flags to ids),
// k_stitch_owned (one thread per pixel: the tile's labels to slide ids, 0 where dropped).
#include "common.hpp"

namespace nsc {
namespace {

constexpr int kScanThreads = 1024;

struct Core {
    long long cx0, cx1, cy0, cy1;  // doubled slide coordinates, half-open
    int x0, y0, w, h;
    bool left, right, top, bottom;  // sides with a neighbouring tile
};

__global__ void k_fill_i32_stitch(int* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

__global__ void k_stitch_mark(const int32_t* __restrict__ box, int n, Core C, int32_t* __restrict__ flag,
                              int* __restrict__ truncated) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int minx = box[n + i], maxx = box[2 * n + i], miny = box[3 * n + i], maxy = box[4 * n + i];
    const long long cx = 2LL * C.x0 + minx + maxx, cy = 2LL * C.y0 + miny + maxy;
    const bool keep = cx >= C.cx0 && cx < C.cx1 && cy >= C.cy0 && cy < C.cy1;
    flag[i] = keep ? 1 : 0;
    if (keep && ((C.left && minx == 0) || (C.right && maxx == C.w - 1) || (C.top && miny == 0) ||
                 (C.bottom && maxy == C.h - 1)))
        atomicAdd(truncated, 1);
}

// exclusive scan of the flags in chunks of kScanThreads (one workgroup, sequential over chunks):
// ids[i] = first_id + rank for kept boxes, 0 otherwise; *kept = number kept
__global__ __launch_bounds__(kScanThreads) void k_stitch_scan(int32_t* __restrict__ ids, int n, int32_t first_id,
                                                              int* __restrict__ kept) {
    __shared__ int wsum[kScanThreads / 64];
    __shared__ int base;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c = 0; c < n; c += kScanThreads) {
        const int i = c + (int)threadIdx.x;
        const int f = i < n ? ids[i] : 0;
        int v = f;  // inclusive scan within the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o, 64);
            if (lane >= o) v += u;
        }
        if (lane == 63) wsum[wv] = v;
        __syncthreads();
        int before = base;
        for (int k = 0; k < wv; ++k) before += wsum[k];
        if (i < n) ids[i] = f ? first_id + before + v - 1 : 0;
        __syncthreads();
        if (threadIdx.x == kScanThreads - 1) base = before + v;
        __syncthreads();
    }
    if (threadIdx.x == 0) *kept = base;
}

__global__ void k_stitch_owned(const int32_t* __restrict__ lab, size_t lab_ld, int w, int h, int n,
                               const int32_t* __restrict__ ids, int32_t* __restrict__ owned, size_t owned_ld) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w) return;
    for (int y = blockIdx.y; y < h; y += gridDim.y) {
        const int32_t l = *(const int32_t*)((const uint8_t*)lab + lab_ld * y + 4 * (size_t)x);
        *(int32_t*)((uint8_t*)owned + owned_ld * y + 4 * (size_t)x) = (l > 0 && l <= n) ? ids[l - 1] : 0;
    }
}

}  // namespace
}  // namespace nsc

using namespace nsc;

extern "C" int nsc_stitch_tile(const int32_t* labels, size_t labels_ld, const int32_t* bbox5, int n, int x0, int y0,
                               int w, int h, int slide_w, int slide_h, int overlap, int32_t first_id, int32_t* ids,
                               int32_t* owned, size_t owned_ld, int* kept, int* truncated, void* stream) {
    NSC_CHECK(ensure_init());
    NSC_ARG(n >= 0 && w > 0 && h > 0 && x0 >= 0 && y0 >= 0 && x0 + w <= slide_w && y0 + h <= slide_h && kept,
            "stitch_tile: geometry");
    NSC_ARG(overlap >= 0 && overlap < w && overlap < h, "stitch_tile: overlap must be smaller than the tile");
    NSC_ARG(first_id >= 1 && (long long)first_id + n < 0x7FFFFFFFLL, "stitch_tile: ids must stay in 1..2^31-1");
    NSC_ARG(n == 0 || (bbox5 && ids), "stitch_tile: bbox5 / ids");
    NSC_ARG(!owned || labels, "stitch_tile: owned needs labels");
    if (!labels_ld) labels_ld = 4 * (size_t)w;
    if (!owned_ld) owned_ld = 4 * (size_t)w;
    const hipStream_t s = as_stream(stream);
    Core C;
    C.x0 = x0;
    C.y0 = y0;
    C.w = w;
    C.h = h;
    C.left = x0 > 0;
    C.top = y0 > 0;
    C.right = x0 + w < slide_w;
    C.bottom = y0 + h < slide_h;
    const int lo = overlap / 2, hi = overlap - overlap / 2;
    C.cx0 = 2LL * (x0 + (C.left ? lo : 0));
    C.cx1 = 2LL * (x0 + w - (C.right ? hi : 0));
    C.cy0 = 2LL * (y0 + (C.top ? lo : 0));
    C.cy1 = 2LL * (y0 + h - (C.bottom ? hi : 0));
    Scratch sc;
    NSC_CHECK(sc.alloc(2 * sizeof(int), s));
    int* cnt = sc.as<int>();  // [kept, truncated]
    hipLaunchKernelGGL(k_fill_i32_stitch, dim3(1), dim3(64), 0, s, cnt, 2);
    if (n > 0) {
        hipLaunchKernelGGL(k_stitch_mark, dim3(cdiv(n, 256)), dim3(256), 0, s, bbox5, n, C, ids, cnt + 1);
        hipLaunchKernelGGL(k_stitch_scan, dim3(1), dim3(kScanThreads), 0, s, ids, n, first_id, cnt);
    }
    if (owned)
        hipLaunchKernelGGL(k_stitch_owned, dim3(cdiv(w, 256), std::min(h, 4096)), dim3(256), 0, s, labels, labels_ld,
                           w, h, n, ids, owned, owned_ld);
    NSC_HIP(hipGetLastError());
    int host[2];
    NSC_HIP(hipMemcpyAsync(host, cnt, sizeof(host), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    *kept = host[0];
    if (truncated) *truncated = host[1];
    return NSC_OK;
}
