#include <atomic>
// cv::watershed (OpenCV 2.4.9 imgproc/segmentation.cpp) on a BGR image with int32 markers, and
// nscale::watershed2 (MorphologicOperations.cpp:3717-3760) with NeighborOperations::border
// (NeighborOperations.cpp:17-44).
//
// cv::watershed is a sequential priority flood: 256 FIFO levels, pop from the lowest non-empty
// level, a popped pixel takes the common positive label of its 4-neighbours (-1 when they
// disagree) and queues its unknown neighbours at their max-channel colour difference.  The state
// of an unknown pixel only changes through pops of unknown pixels 4-adjacent to it, and labelled
// pixels never change, so the flood restricted to one 4-connected component of unknown pixels is
// independent of every other component: the FIFO order inside a component is the global order
// restricted to it.  The GPU therefore labels the unknown components (CCL) and runs the exact
// flood of each component on its own lane, all components in parallel.
#include <cstdlib>

#include "common.hpp"

namespace nsc {

int ccl_neq(const uint8_t* img, uint8_t key, int32_t* labels, int w, int h, int conn, hipStream_t s);
int relabel_impl(int32_t* labels, int w, int h, int32_t bgval, int* dcount, int32_t* rank, hipStream_t s);
int bbox_impl(const int32_t* canon, int w, int h, int n, int32_t* out5, int32_t* area, hipStream_t s);

namespace {

constexpr int NQ = 256;
constexpr int IN_QUEUE = -2, WSHED = -1;
constexpr uint16_t NOPUSH = 0xFFFF;

__device__ __forceinline__ int c_diff(const uint8_t* a, const uint8_t* b) {
    const int db = abs((int)a[0] - (int)b[0]);
    const int dg = abs((int)a[1] - (int)b[1]);
    const int dr = abs((int)a[2] - (int)b[2]);
    return max(max(db, dg), dr);
}

// initial state: frame -> -1; unknown = interior with marker <= 0 (reset to 0); level of the
// initial push = min colour difference to a positive 4-neighbour (frame counts as -1)
__global__ void k_ws_init(const uint8_t* __restrict__ img, const int32_t* __restrict__ m0, int32_t* __restrict__ m,
                          uint8_t* __restrict__ unknown, uint16_t* __restrict__ lvl, int W, int H) {
    const size_t n = (size_t)W * H;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % W), y = (int)(i / W);
        const bool frame = (x == 0 || y == 0 || x == W - 1 || y == H - 1);
        const int v = m0[i];
        uint16_t L = NOPUSH;
        uint8_t u = 0;
        if (!frame && v <= 0) {
            u = 1;
            const uint8_t* p = img + 3 * i;
            int best = 256;
            auto pos = [&](size_t j, int xx, int yy) {
                return !(xx == 0 || yy == 0 || xx == W - 1 || yy == H - 1) && m0[j] > 0;
            };
            if (pos(i - 1, x - 1, y)) best = min(best, c_diff(p, p - 3));
            if (pos(i + 1, x + 1, y)) best = min(best, c_diff(p, p + 3));
            if (pos(i - W, x, y - 1)) best = min(best, c_diff(p, p - 3 * (size_t)W));
            if (pos(i + W, x, y + 1)) best = min(best, c_diff(p, p + 3 * (size_t)W));
            if (best < 256) L = (uint16_t)best;
        }
        m[i] = frame ? WSHED : (v < 0 ? 0 : v);
        unknown[i] = u;
        lvl[i] = L;
    }
}

// exact flood of one unknown component per thread
__global__ void k_ws_flood(const uint8_t* __restrict__ img, int32_t* m, const int32_t* __restrict__ comp,
                           const uint16_t* __restrict__ lvl, const int32_t* __restrict__ box5, int ncomp, int W,
                           int* __restrict__ next, int* __restrict__ qhead, int* __restrict__ qtail) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncomp) return;
    int* head = qhead + (size_t)c * NQ;
    int* tail = qtail + (size_t)c * NQ;
    for (int l = 0; l < NQ; ++l) { head[l] = -1; tail[l] = -1; }
    const int minx = box5[ncomp + c], maxx = box5[2 * ncomp + c];
    const int miny = box5[3 * ncomp + c], maxy = box5[4 * ncomp + c];
    const int id = c + 1;
    int active = NQ;
    auto push = [&](int l, int p) {
        next[p] = -1;
        if (tail[l] >= 0) next[tail[l]] = p;
        else head[l] = p;
        tail[l] = p;
        if (l < active) active = l;
    };
    // initial pushes in raster order
    for (int y = miny; y <= maxy; ++y)
        for (int x = minx; x <= maxx; ++x) {
            const int i = y * W + x;
            if (comp[i] == id && lvl[i] != NOPUSH) {
                push(lvl[i], i);
                m[i] = IN_QUEUE;
            }
        }
    if (active == NQ) return;
    for (;;) {
        if (head[active] < 0) {
            int l = active + 1;
            while (l < NQ && head[l] < 0) ++l;
            if (l == NQ) break;
            active = l;
        }
        const int p = head[active];
        const int nx = next[p];
        head[active] = nx;
        if (nx < 0) tail[active] = -1;
        int lab = 0, t;
        t = m[p - 1];
        if (t > 0) lab = t;
        t = m[p + 1];
        if (t > 0) { if (lab == 0) lab = t; else if (t != lab) lab = WSHED; }
        t = m[p - W];
        if (t > 0) { if (lab == 0) lab = t; else if (t != lab) lab = WSHED; }
        t = m[p + W];
        if (t > 0) { if (lab == 0) lab = t; else if (t != lab) lab = WSHED; }
        m[p] = lab;
        if (lab == WSHED) continue;
        const uint8_t* ptr = img + 3 * (size_t)p;
        if (m[p - 1] == 0) { push(c_diff(ptr, ptr - 3), p - 1); m[p - 1] = IN_QUEUE; }
        if (m[p + 1] == 0) { push(c_diff(ptr, ptr + 3), p + 1); m[p + 1] = IN_QUEUE; }
        if (m[p - W] == 0) { push(c_diff(ptr, ptr - 3 * (size_t)W), p - W); m[p - W] = IN_QUEUE; }
        if (m[p + W] == 0) { push(c_diff(ptr, ptr + 3 * (size_t)W), p + W); m[p + W] = IN_QUEUE; }
    }
}

// border<int>(padded, -1) at padded (x+1, y+1), written unpadded (the crop of watershed2)
__global__ void k_ws_border_crop(const int32_t* __restrict__ P, int32_t* __restrict__ out, int w, int h) {
    const int W = w + 2;
    const size_t n = (size_t)w * h;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % w) + 1, y = (int)(i / w) + 1;
        const int32_t* r = P + (size_t)y * W;
        const int32_t* u = r - W;
        // all six loads first: a short-circuit condition would put them behind branches
        const int32_t p = r[x], rl = r[x - 1], rr = r[x + 1], uc = u[x], ul = u[x - 1], ur = u[x + 1];
        const bool dam = (uc == -1) & (p != -1) & (((ul != -1) & (rl == -1)) | ((ur != -1) & (rr == -1)));
        out[i] = dam ? -1 : p;
    }
}

// labels (w x h) -> padded marker image with a 0 frame, written twice (P, the flood's in/out image,
// and P0, its copy of the initial markers); BGR image -> padded with (0,0,0)
__global__ void k_ws_pad(const int32_t* __restrict__ lab, const uint8_t* __restrict__ bgr, int32_t* __restrict__ P,
                         int32_t* __restrict__ P0, uint8_t* __restrict__ Pimg, int w, int h) {
    const int W = w + 2, H = h + 2;
    const size_t n = (size_t)W * H;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % W) - 1, y = (int)(i / W) - 1;
        const bool in = x >= 0 && y >= 0 && x < w && y < h;
        // loads from a clamped (always valid) pixel, then selects: no branch around the loads
        const size_t j = (size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1);
        const int32_t lv = lab[j];
        const uint8_t c0 = bgr[3 * j], c1 = bgr[3 * j + 1], c2 = bgr[3 * j + 2];
        P[i] = in ? lv : 0;
        P0[i] = in ? lv : 0;
        Pimg[3 * i] = in ? c0 : 0;
        Pimg[3 * i + 1] = in ? c1 : 0;
        Pimg[3 * i + 2] = in ? c2 : 0;
    }
}

inline dim3 grid1(size_t n) { return dim3(std::min<unsigned>(cdiv(n, 256), 8192u)); }

}  // namespace

int cvws_fast(const uint8_t* img, const int32_t* m0, int32_t* m, int W, int H, int* ok, hipStream_t s);

// how often the decomposed flood applied / fell back (process-wide counters, for reporting)
static std::atomic<int> g_ws_fast{0}, g_ws_fallback{0}, g_ws_last_flags{0};
void ws_stats_record(int ok) { if (ok) ++g_ws_fast; else ++g_ws_fallback; }
void ws_last_flags_record(int flags) { g_ws_last_flags = flags; }
// phase-1 runs per LDS tier, summed over the decomposed floods of the process (reporting only)
static std::atomic<long long> g_ws_tier_runs[8];
void ws_tier_record(const int* runs, int ntiers) {
    for (int t = 0; t < ntiers && t < 8; ++t) g_ws_tier_runs[t] += runs[t];
}

size_t cvws_scratch(int W, int H) {
    const size_t n = (size_t)W * H;
    return n * (sizeof(int32_t) * 3 + 1 + sizeof(uint16_t)) + 64;
}

// markers m (W x H) in/out; the copy of the initial markers, m0, is the start of the scratch (made
// here unless the caller has written it already).  The decomposed flood (cvws_fast) is tried first;
// when its assumptions do not hold it reports so and the plain per-component sequential flood
// below produces the result.  NSC_WATERSHED=seq forces the sequential flood; NSC_WATERSHED=fast
// bounds the latency instead: such an input returns NSC_ERR_LIMIT (the sequential flood of a
// 4096^2 tile's background takes ~20 s).  NSC_WATERSHED_REJECT=1 makes the host treat every
// decomposed result as not applicable (failure-handling tests of both paths).
int cvws_impl(const uint8_t* img, int32_t* m, int W, int H, void* scratch, hipStream_t s, bool m0_written = false) {
    const size_t n = (size_t)W * H;
    int32_t* m0 = static_cast<int32_t*>(scratch);
    int32_t* comp = m0 + n;
    int32_t* rank = comp + n;  // relabel scratch, then the flood's next[] links
    uint16_t* lvl = reinterpret_cast<uint16_t*>(rank + n);
    uint8_t* unknown = reinterpret_cast<uint8_t*>(lvl + n);
    if (!m0_written) NSC_HIP(hipMemcpyAsync(m0, m, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    const char* mode = getenv("NSC_WATERSHED");
    const bool force_seq = mode && mode[0] == 's';
    if (!force_seq) {
        int ok = 0;
        NSC_CHECK(cvws_fast(img, m0, m, W, H, &ok, s));
        const char* reject = getenv("NSC_WATERSHED_REJECT");
        if (reject && reject[0] == '1') ok = 0;
        ws_stats_record(ok);
        if (ok) return NSC_OK;
        if (mode && mode[0] == 'f') {
            set_error("watershed: the decomposed flood does not apply to this input (NSC_WATERSHED=fast)");
            return NSC_ERR_LIMIT;
        }
    }
    hipLaunchKernelGGL(k_ws_init, grid1(n), dim3(256), 0, s, img, m0, m, unknown, lvl, W, H);
    NSC_CHECK(ccl_neq(unknown, 0, comp, W, H, 4, s));
    Scratch cnt;
    NSC_CHECK(cnt.alloc(sizeof(int), s));
    NSC_CHECK(relabel_impl(comp, W, H, -1, cnt.as<int>(), rank, s));
    int ncomp = 0;
    NSC_HIP(hipMemcpyAsync(&ncomp, cnt.p, sizeof(int), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    if (ncomp == 0) return NSC_OK;
    Scratch q;
    NSC_CHECK(q.alloc(sizeof(int32_t) * ((size_t)ncomp * (5 + 2 * NQ)), s));
    int32_t* box5 = q.as<int32_t>();
    int* qhead = box5 + 5 * (size_t)ncomp;
    int* qtail = qhead + (size_t)ncomp * NQ;
    NSC_CHECK(bbox_impl(comp, W, H, ncomp, box5, nullptr, s));
    hipLaunchKernelGGL(k_ws_flood, dim3(cdiv(ncomp, 64)), dim3(64), 0, s, img, m, comp, lvl, box5, ncomp, W, rank,
                       qhead, qtail);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

size_t ws2_scratch(int w, int h) {
    const size_t N = (size_t)(w + 2) * (h + 2);
    return N * (sizeof(int32_t) + 3) + cvws_scratch(w + 2, h + 2);
}

// watershed2 tail: labels of the minima (w x h, bg -1) + BGR -> out (w x h)
int ws2_from_labels(const uint8_t* bgr, const int32_t* labels, int32_t* out, int w, int h, void* scratch, hipStream_t s) {
    const size_t N = (size_t)(w + 2) * (h + 2);
    int32_t* P = static_cast<int32_t*>(scratch);
    uint8_t* Pimg = reinterpret_cast<uint8_t*>(P + N);
    void* rest = Pimg + 3 * N + (16 - (3 * N) % 16) % 16;
    // the flood's copy of the initial markers (the start of its scratch) is written by the padding
    hipLaunchKernelGGL(k_ws_pad, grid1(N), dim3(256), 0, s, labels, bgr, P, static_cast<int32_t*>(rest), Pimg, w, h);
    NSC_CHECK(cvws_impl(Pimg, P, w + 2, h + 2, rest, s, true));
    hipLaunchKernelGGL(k_ws_border_crop, grid1((size_t)w * h), dim3(256), 0, s, P, out, w, h);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

}  // namespace nsc

using namespace nsc;

namespace nsc {
int local_max_impl(const float* img, uint8_t* out, int w, int h, int conn, bool minima, void* scratch, hipStream_t s);
size_t local_max_scratch(int w, int h);
int ccl_u8(const uint8_t* img, int32_t* labels, int w, int h, int conn, bool equal_values, int32_t bgval, hipStream_t s);
}  // namespace nsc

extern "C" {

/* counts of watershed calls served by the decomposed flood / by the sequential fallback */
void nsc_watershed_stats(int* fast, int* fallback, int* last_reason) {
    if (fast) *fast = g_ws_fast;
    if (fallback) *fallback = g_ws_fallback;
    if (last_reason) *last_reason = g_ws_last_flags;
}

int nsc_watershed_tier_runs(long long* runs, int n) {
    if (!runs || n <= 0) return 0;
    const int k = n < 6 ? n : 6;
    for (int t = 0; t < k; ++t) runs[t] = g_ws_tier_runs[t];
    return 6;
}

int nsc_cv_watershed(const uint8_t* bgr, int32_t* markers, int w, int h, void* stream) {
    NSC_ARG(bgr && markers && w > 0 && h > 0, "cv_watershed");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    Scratch sc;
    NSC_CHECK(sc.alloc(cvws_scratch(w, h), s));
    return cvws_impl(bgr, markers, w, h, sc.p, s);
}

int nsc_watershed2(const uint8_t* bgr, const float* img, int32_t* out, int w, int h, int conn, void* stream) {
    NSC_ARG(bgr && img && out && w > 0 && h > 0 && valid_conn(conn), "watershed2");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    const size_t n = (size_t)w * h;
    Scratch a, b;
    NSC_CHECK(a.alloc(std::max(local_max_scratch(w, h), ws2_scratch(w, h)), s));
    NSC_CHECK(b.alloc(n * (1 + sizeof(int32_t)) + 16, s));
    uint8_t* minima = b.as<uint8_t>();
    int32_t* labels = reinterpret_cast<int32_t*>(minima + n + (16 - n % 16) % 16);
    NSC_CHECK(local_max_impl(img, minima, w, h, conn, true, a.p, s));  // localMinima (MO.cpp:3730)
    NSC_CHECK(ccl_u8(minima, labels, w, h, conn, true, -1, s));       // bwlabel2 (MO.cpp:3737)
    return ws2_from_labels(bgr, labels, out, w, h, a.p, s);
}

}  // extern "C"
