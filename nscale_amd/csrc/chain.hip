// HistologicalEntities::segmentNuclei (HistologicalEntities.cpp:473-598, with
// plFindNucleusCandidates :115-304 and plSeparateNuclei :308-449) on one tile, every
// intermediate resident in a reusable device workspace.  Pixel predicates and masks are fused
// into small elementwise kernels; the morphology, reconstruction, labelling, distance and
// watershed steps call the operator implementations directly.
#include <cfloat>
#include <climits>
#include <cstring>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "common.hpp"

namespace nsc {

int morph_impl(const uint8_t* in, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside, bool erode,
               hipStream_t s);
int morph_open_impl(const uint8_t* in, uint8_t* out, uint8_t* tmp, int w, int h, const uint8_t* kernel, int k,
                    hipStream_t s);
int bwareaopen_impl(const uint8_t* img, uint8_t* out, int w, int h, int lo, int hi, int conn, int* dcount,
                    int32_t* labels, int32_t* area, hipStream_t s);
int fill_holes_impl(const uint8_t* img, uint8_t* out, int w, int h, int conn, int32_t* labels, uint8_t* flag,
                    hipStream_t s);
int bwselect_impl(const uint8_t* bw, const uint8_t* seeds, uint8_t* out, int w, int h, int conn, int32_t* labels,
                  uint8_t* flag, hipStream_t s);
int edt_impl(const uint8_t* img, float* out, int w, int h, hipStream_t s);
int imhmin_impl(const float* img, float* out, int w, int h, float hval, int conn, float* s1, float* s2, hipStream_t s);
int local_max_impl(const float* img, uint8_t* out, int w, int h, int conn, bool minima, void* scratch, hipStream_t s);
size_t local_max_scratch(int w, int h);
size_t ws2_scratch(int w, int h);
int ws2_from_labels(const uint8_t* bgr, const int32_t* labels, int32_t* out, int w, int h, void* scratch, hipStream_t s);
int ccl_u8(const uint8_t* img, int32_t* labels, int w, int h, int conn, bool equal_values, int32_t bgval, hipStream_t s);
int relabel_impl(int32_t* labels, int w, int h, int32_t bgval, int* dcount, int32_t* rank, hipStream_t s);
int root_count_impl(const int32_t* labels, int w, int h, int* dcount, hipStream_t s);
int bbox_impl(const int32_t* canon, int w, int h, int n, int32_t* out5, int32_t* area, hipStream_t s);
int bbox_literal_impl(const int32_t* L, int w, int h, int32_t bgval, int32_t* out5, int cap, int* compcount,
                      int32_t* firstpos, int32_t* R, int32_t* rank, int* cnt, hipStream_t s);

namespace {

const uint8_t kDisk19[361] = {  // HistologicalEntities.cpp:169-188 (MATLAB strel('disk',10))
    0,0,0,0,1,1,1,1,1,1,1,1,1,1,1,0,0,0,0, 0,0,0,1,1,1,1,1,1,1,1,1,1,1,1,1,0,0,0,
    0,0,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,0,0, 0,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,0,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1, 1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1, 1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1, 1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1, 1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1, 1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1, 0,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,0,
    0,0,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,0,0, 0,0,0,1,1,1,1,1,1,1,1,1,1,1,1,1,0,0,0,
    0,0,0,0,1,1,1,1,1,1,1,1,1,1,1,0,0,0,0};
const uint8_t kCross3[9] = {0, 1, 0, 1, 1, 1, 0, 1, 0};  // getStructuringElement(MORPH_ELLIPSE, 3x3)

inline dim3 grid1(size_t n) { return dim3(std::min<unsigned>(cdiv(n, 256), 8192u)); }

// getBackground (HE.cpp:93-109) count, getRBC planes (HE.cpp:33-82), rc = 255 - R (HE.cpp:158)
__device__ __forceinline__ void s1_pixel(uint8_t B, uint8_t G, uint8_t R, uint8_t blue, uint8_t green, uint8_t red, float t1,
                                         float t2, uint32_t& bg, uint32_t& b1, uint32_t& b2, uint32_t& rb, uint32_t& rcv,
                                         int k, int& local) {
    const bool isbg = (B > blue) && (G > green) && (R > red);
    local += isbg;
    const float bd = (float)B + FLT_EPSILON;
    const float gd = (float)G + FLT_EPSILON;
    const float rd = (float)R;
    const float q = rd / gd;  // IEEE division, correctly rounded (no fast-math)
    bg |= (isbg ? 255u : 0u) << (8 * k);
    b1 |= ((q > t1) ? 255u : 0u) << (8 * k);
    b2 |= ((q > t2) ? 255u : 0u) << (8 * k);
    rb |= ((rd / bd > 1.0f) ? 255u : 0u) << (8 * k);
    rcv |= (uint32_t)(uint8_t)(255 - R) << (8 * k);
}

// four pixels per thread: 12 BGR bytes in as three words, one word out per plane
__global__ void k_s1_pixels(const uint8_t* __restrict__ bgr, size_t n, uint8_t blue, uint8_t green, uint8_t red, float t1,
                            float t2, uint8_t* __restrict__ bgmask, uint8_t* __restrict__ bw1, uint8_t* __restrict__ bw2,
                            uint8_t* __restrict__ r2b, uint8_t* __restrict__ rc, int* __restrict__ bg_count, int vec) {
    int local = 0;
    const size_t n4 = vec ? n / 4 : 0;  // word access needs a 4-byte aligned input
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += (size_t)gridDim.x * blockDim.x) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(bgr + 12 * q);
        const uint32_t w0 = src[0], w1 = src[1], w2 = src[2];
        uint8_t px[12];
        for (int k = 0; k < 4; ++k) {
            px[k] = (uint8_t)(w0 >> (8 * k));
            px[4 + k] = (uint8_t)(w1 >> (8 * k));
            px[8 + k] = (uint8_t)(w2 >> (8 * k));
        }
        uint32_t bg = 0, b1 = 0, b2 = 0, rb = 0, rcv = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) s1_pixel(px[3 * k], px[3 * k + 1], px[3 * k + 2], blue, green, red, t1, t2, bg, b1, b2, rb, rcv, k, local);
        if (bgmask) reinterpret_cast<uint32_t*>(bgmask)[q] = bg;
        reinterpret_cast<uint32_t*>(bw1)[q] = b1;
        reinterpret_cast<uint32_t*>(bw2)[q] = b2;
        reinterpret_cast<uint32_t*>(r2b)[q] = rb;
        reinterpret_cast<uint32_t*>(rc)[q] = rcv;
    }
    // tail pixels (n not a multiple of 4)
    for (size_t i = 4 * n4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t bg = 0, b1 = 0, b2 = 0, rb = 0, rcv = 0;
        s1_pixel(bgr[3 * i], bgr[3 * i + 1], bgr[3 * i + 2], blue, green, red, t1, t2, bg, b1, b2, rb, rcv, 0, local);
        if (bgmask) bgmask[i] = (uint8_t)bg;
        bw1[i] = (uint8_t)b1;
        bw2[i] = (uint8_t)b2;
        r2b[i] = (uint8_t)rb;
        rc[i] = (uint8_t)rcv;
    }
    block_count_add(local, bg_count);
}

// diffIm = sat(rc - rc_recon) (HE.cpp:211); thresholds at G1 / G2 (:227, :279)
__global__ void k_diff_thr(const uint8_t* __restrict__ rc, const uint8_t* __restrict__ rec, size_t n, uint8_t g1,
                           uint8_t g2, uint8_t* __restrict__ diff, uint8_t* __restrict__ t1, uint8_t* __restrict__ t2) {
    // four pixels per thread (the planes are workspace buffers, 256-byte aligned)
    const size_t n4 = n / 4;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += (size_t)gridDim.x * blockDim.x) {
        const uint32_t a = reinterpret_cast<const uint32_t*>(rc)[q], b = reinterpret_cast<const uint32_t*>(rec)[q];
        uint32_t dv = 0, o1 = 0, o2 = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int d = (int)((a >> (8 * k)) & 255) - (int)((b >> (8 * k)) & 255);
            const uint32_t v = d > 0 ? (uint32_t)d : 0u;
            dv |= v << (8 * k);
            o1 |= (v > g1 ? 255u : 0u) << (8 * k);
            o2 |= (v > g2 ? 255u : 0u) << (8 * k);
        }
        if (diff) reinterpret_cast<uint32_t*>(diff)[q] = dv;
        reinterpret_cast<uint32_t*>(t1)[q] = o1;
        reinterpret_cast<uint32_t*>(t2)[q] = o2;
    }
    for (size_t i = 4 * n4 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int d = (int)rc[i] - (int)rec[i];
        const uint8_t v = d > 0 ? (uint8_t)d : 0;
        if (diff) diff[i] = v;
        t1[i] = v > g1 ? 255 : 0;
        t2[i] = v > g2 ? 255 : 0;
    }
}

// rbc = rbcsel & r2b; seg_norbc &= (rbc == 0)   (HE.cpp:76, :295)
__global__ void k_norbc(uint8_t* __restrict__ seg, const uint8_t* __restrict__ rbcsel, const uint8_t* __restrict__ r2b,
                        uint8_t* __restrict__ rbc_out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint8_t rbc = rbcsel[i] & r2b[i];
        if (rbc_out) rbc_out[i] = rbc;
        seg[i] = seg[i] & (rbc == 0 ? 255 : 0);
    }
}

// dist -> distance (HE.cpp:366-374) -> imhmin operands mask = -distance, marker = mask - 1
// (MO.cpp:3355-3356); nuclei = img masked by seg_big (HE.cpp:407)
__global__ void k_dist_prep(const float* __restrict__ dist, const uint8_t* __restrict__ seg_big,
                            const uint8_t* __restrict__ bgr, size_t n, float* __restrict__ distance,
                            float* __restrict__ hmask, float* __restrict__ hmarker, uint8_t* __restrict__ nuclei) {
    // every load unconditional, then selects (a load under the mask's condition compiles to a
    // branch around it, and the loads of a pixel then wait for one another)
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint8_t sb = seg_big[i];
        const float dv = dist[i];
        const uint8_t c0 = bgr[3 * i], c1 = bgr[3 * i + 1], c2 = bgr[3 * i + 2];
        const bool f = sb != 0;
        const float d = f ? -dv : -FLT_MAX;
        if (distance) distance[i] = d;
        const float m = -d;
        hmask[i] = m;
        hmarker[i] = m - 1.0f;
        nuclei[3 * i] = f ? c0 : 0;
        nuclei[3 * i + 1] = f ? c1 : 0;
        nuclei[3 * i + 2] = f ? c2 : 0;
    }
}

__global__ void k_negate_f(float* a, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = -a[i];
}

// seg_nonoverlap = seg_big where watermask >= 0  (HE.cpp:423-424)
__global__ void k_nonoverlap(const uint8_t* __restrict__ seg_big, const int32_t* __restrict__ wm, uint8_t* __restrict__ out,
                             size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int32_t m = wm[i];
        const uint8_t sb = seg_big[i];  // unconditional: no branch around the load
        out[i] = m >= 0 ? sb : 0;
    }
}

}  // namespace

// ------------------------------------------------------------------------------------------
struct Workspace {
    int max_w = 0, max_h = 0;
    size_t n_cap = 0, big_bytes = 0;  // pixel capacity of the planes, bytes of the shared scratch
    void* base = nullptr;
    size_t bytes = 0;
    // u8 planes
    uint8_t *bgm, *bw1, *bw2, *r2b, *rc, *rc_open, *rc_recon, *diff, *t1, *t2, *fill1, *bw1_t, *rbcsel, *rbc, *seg_norbc,
        *seg_nohole, *seg_open, *seg_big_t, *seg_big, *nonov, *seg, *fin, *tmp8, *nuclei;
    // 32-bit planes
    int32_t *L, *area, *wm;
    float *dist, *distance, *hmask, *hmarker, *d2;
    void* big;  // shared by local minima and the watershed (max of their scratch)
    int* counts;  // [0] background, [1] compcount S1, [2] compcount final-area, [3] ncomp
    Arena arena;  // operator scratch during a chain call (separate allocation)
};

// operator scratch the chain needs at most, per pixel of the largest tile (measured peak of the
// watershed's decomposition plus margin); beyond it operators fall back to the stream pool
constexpr size_t kArenaPerPx = 144;

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
static size_t big_scratch(int w, int h) { return al256(std::max(local_max_scratch(w, h), ws2_scratch(w, h))); }

size_t workspace_bytes(int mw, int mh) {
    const size_t n = (size_t)mw * mh;
    return 23 * al256(n) + al256(3 * n) + 8 * al256(4 * n) + big_scratch(mw, mh) + 256;
}

size_t workspace_arena_bytes(int mw, int mh) { return kArenaPerPx * (size_t)mw * mh + ((size_t)64 << 20); }

// a tile fits a workspace when its planes and its shared scratch fit (orientation-free)
bool workspace_fits(const nsc_workspace* ws, int w, int h);

static int ws_alloc(Workspace& W, int mw, int mh) {
    const size_t n = (size_t)mw * mh;
    const size_t u8n = al256(n), b3 = al256(3 * n), i32 = al256(4 * n);
    const size_t big = big_scratch(mw, mh);
    W.bytes = workspace_bytes(mw, mh);
    W.n_cap = n;
    W.big_bytes = big;
    NSC_HIP(hipMalloc(&W.base, W.bytes));
    W.arena.cap = workspace_arena_bytes(mw, mh);
    if (hipMalloc(&W.arena.base, W.arena.cap) != hipSuccess) {  // optional: the pool serves instead
        (void)hipGetLastError();
        W.arena.base = nullptr;
        W.arena.cap = 0;
    }
    uint8_t* p = static_cast<uint8_t*>(W.base);
    uint8_t** u8s[] = {&W.bgm, &W.bw1, &W.bw2, &W.r2b, &W.rc, &W.rc_open, &W.rc_recon, &W.diff, &W.t1, &W.t2,
                       &W.fill1, &W.bw1_t, &W.rbcsel, &W.rbc, &W.seg_norbc, &W.seg_nohole, &W.seg_open, &W.seg_big_t,
                       &W.seg_big, &W.nonov, &W.seg, &W.fin, &W.tmp8};
    for (auto q : u8s) { *q = p; p += u8n; }
    W.nuclei = p; p += b3;
    W.L = reinterpret_cast<int32_t*>(p); p += i32;
    W.area = reinterpret_cast<int32_t*>(p); p += i32;
    W.wm = reinterpret_cast<int32_t*>(p); p += i32;
    W.dist = reinterpret_cast<float*>(p); p += i32;
    W.distance = reinterpret_cast<float*>(p); p += i32;
    W.hmask = reinterpret_cast<float*>(p); p += i32;
    W.hmarker = reinterpret_cast<float*>(p); p += i32;
    W.d2 = reinterpret_cast<float*>(p); p += i32;
    W.big = p; p += big;
    W.counts = reinterpret_cast<int*>(p);
    W.max_w = mw;
    W.max_h = mh;
    return NSC_OK;
}

}  // namespace nsc

struct nsc_workspace : nsc::Workspace {};

bool nsc::workspace_fits(const nsc_workspace* ws, int w, int h) {
    return ws && (size_t)w * h <= ws->n_cap && big_scratch(w, h) <= ws->big_bytes;
}

using namespace nsc;

extern "C" {

void nsc_seg_params_default(nsc_seg_params* p) {
    p->blue = 220; p->green = 220; p->red = 220;
    p->T1 = 5.0; p->T2 = 4.0;
    p->G1 = 80;
    p->minSize = 11; p->maxSize = 1000;
    p->G2 = 45;
    p->minSizePl = 30; p->minSizeSeg = 21; p->maxSizeSeg = 1000;
    p->fillHolesConnectivity = 4; p->reconConnectivity = 8; p->watershedConnectivity = 8;
}

int nsc_workspace_create(int max_w, int max_h, nsc_workspace** ws) {
    NSC_ARG(ws && max_w > 0 && max_h > 0, "workspace_create");
    NSC_LABEL_SIZE(max_w, max_h);
    NSC_CHECK(ensure_init());
    auto* W = new nsc_workspace();
    int rc = ws_alloc(*W, max_w, max_h);
    if (rc != NSC_OK) { delete W; return rc; }
    *ws = W;
    return NSC_OK;
}

int nsc_workspace_destroy(nsc_workspace* ws) {
    if (!ws) return NSC_OK;
    if (ws->base) (void)hipFree(ws->base);
    if (ws->arena.base) (void)hipFree(ws->arena.base);
    delete ws;
    return NSC_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// The chain, stage by stage.  A call's context carries the workspace, the stream and the optional
// per-stage dumps and timers (the reference's IntermediateResultHandler ids and logger stages).
namespace nsc {
namespace {

// label column of a box block built on canonical labels: the box of canonical label k = canon[r]
// of root r (L[r] == r) gets label r
__global__ void k_root_label_column(const int32_t* __restrict__ L, const int32_t* __restrict__ canon, int n,
                                    int32_t* __restrict__ out5) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (L[i] == i) {
            const int k = canon[i];
            if (k >= 1) out5[k - 1] = i;
        }
    }
}

// The reference's logger stage name (SimpleCSVLogger::logTimeSinceLastLog, HistologicalEntities.cpp
// :121-588) of the stage that follows the mark with stage id `id`; null after the last one.
static const char* stage_after(int id) {
    switch (id) {
    case 0: return "background";      // toRGB + background ratio (:121-140)
    case 1: return "RBC";             // :144
    case 2: return "open19";          // invert + open19 (:159-194)
    case 3: return "reconToNuclei";   // :213
    case 26: return "threshold1";     // :228
    case 4: return "fillHoles1";      // :234
    case 7: return "areaThreshold1";  // :268
    case 8: return "blobsGt45";       // :299
    case 10: return "fillHoles2";     // :495
    case 12: return "openBlobs";      // :520
    case 13: return "30To1000";       // :323
    case 14: return "dilate";         // :334
    case 15: return "distTransform";  // :376
    case 16: return "imhmin";         // :385
    case 18: return "watershed";      // watershed + water to mask (:419-425)
    case 20: return "20To1000";       // :556
    case 23: return "fillHolesLast";  // :580
    case 24: return "bwlabel2";       // :588
    default: return nullptr;
    }
}

struct Chain {
    Workspace& W;
    hipStream_t s;
    void** stage;
    double* stage_ms;
    std::vector<std::pair<int, hipEvent_t>> marks;
    bool range_open = false;  // a roctx range named for the current stage is pushed
    Chain(Workspace& w, hipStream_t st, void** sg, double* ms) : W(w), s(st), stage(sg), stage_ms(ms) {}
    ~Chain() {
        if (range_open) roctxRangePop();
        for (auto& m : marks) (void)hipEventDestroy(m.second);
    }
    // stage boundary: closes the roctx range of the stage that ends here (host time: the interval in
    // which its kernels were queued) and opens the next stage's, under the reference's logger names;
    // with timers, the GPU time since the previous mark is charged to stage id
    void mark(int id) {
        if (range_open) roctxRangePop();
        const char* next = stage_after(id);
        range_open = next != nullptr;
        if (next) roctxRangePushA(next);
        if (!stage_ms) return;
        hipEvent_t e;
        if (hipEventCreate(&e) == hipSuccess) {
            (void)hipEventRecord(e, s);
            marks.push_back({id, e});
        }
    }
    int save(int id, const void* src, size_t bytes) {
        if (stage && stage[id]) NSC_HIP(hipMemcpyAsync(stage[id], src, bytes, hipMemcpyDeviceToDevice, s));
        return NSC_OK;
    }
    bool want(int id) const { return stage && stage[id]; }
    void finish_timing() {
        if (!stage_ms || marks.empty()) return;
        (void)hipEventSynchronize(marks.back().second);
        for (size_t k = 1; k < marks.size(); ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, marks[k - 1].second, marks[k].second);
            stage_ms[marks[k].first] += ms;
        }
        for (auto& m : marks) (void)hipEventDestroy(m.second);
        marks.clear();
    }
};

// plFindNucleusCandidates (HistologicalEntities.cpp:115-304): candidates in W.seg_norbc.
// Returns NSC_SEG_CONTINUE or the early status (background, no candidates).
int stage_s1(Chain& C, const uint8_t* bgr, int w, int h, const nsc_seg_params& p) {
    Workspace& W = C.W;
    hipStream_t s = C.s;
    const size_t n = (size_t)w * h;
    C.mark(0);
    NSC_CHECK(C.save(0, bgr, 3 * n));
    NSC_HIP(hipMemsetAsync(W.counts, 0, 4 * sizeof(int), s));
    hipLaunchKernelGGL(k_s1_pixels, grid_counted(cdiv(n, 4)), dim3(256), 0, s, bgr, n, p.blue, p.green, p.red, (float)p.T1,
                       (float)p.T2, C.want(1) ? W.bgm : nullptr, W.bw1, W.bw2, W.r2b, W.rc, &W.counts[0],
                       (reinterpret_cast<uintptr_t>(bgr) & 3) == 0 ? 1 : 0);
    NSC_HIP(hipGetLastError());
    int bgArea = 0;
    NSC_HIP(hipMemcpyAsync(&bgArea, &W.counts[0], sizeof(int), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    const float ratio = (float)bgArea / (float)(w * h);  // HE.cpp:125
    C.mark(1);
    if (ratio >= 0.99) return NSC_SEG_BACKGROUND;
    if (ratio >= 0.9) return NSC_SEG_BACKGROUND_LIKELY;
    NSC_CHECK(C.save(1, W.bgm, n));

    // getRBC: bwselect(bw2, bw1, 8) & imR2B -- an empty bw1 selects nothing, which is the
    // reference's zero branch (HE.cpp:74-80)
    NSC_CHECK(bwselect_impl(W.bw2, W.bw1, W.rbcsel, w, h, 8, W.L, W.tmp8, s));
    C.mark(2);
    // rc_open = morphOpen(rc, disk19); rc_recon = imreconstruct(rc_open, rc, reconConnectivity)
    NSC_CHECK(morph_open_impl(W.rc, W.rc_open, W.tmp8, w, h, kDisk19, 19, s));
    NSC_CHECK(C.save(3, W.rc_open, n));
    C.mark(3);
    NSC_CHECK(recon_impl<uint8_t>(W.rc_open, W.rc, W.rc_recon, w, h, p.reconConnectivity, s));
    C.mark(26);  // timing id 26: the imreconstruct<uchar>(rc_open, rc) call alone
    NSC_CHECK(C.save(4, W.rc_recon, n));
    hipLaunchKernelGGL(k_diff_thr, grid1(n), dim3(256), 0, s, W.rc, W.rc_recon, n, p.G1, p.G2, W.diff, W.t1, W.t2);
    NSC_CHECK(C.save(5, W.diff, n));
    NSC_CHECK(C.save(6, W.t1, n));
    C.mark(4);
    NSC_CHECK(fill_holes_impl(W.t1, W.fill1, w, h, p.fillHolesConnectivity, W.L, W.tmp8, s));
    NSC_CHECK(C.save(7, W.fill1, n));
    C.mark(7);
    NSC_CHECK(bwareaopen_impl(W.fill1, W.bw1_t, w, h, p.minSize, p.maxSize, 8, &W.counts[1], W.L, W.area, s));
    NSC_CHECK(C.save(8, W.bw1_t, n));
    NSC_CHECK(C.save(9, W.t2, n));
    C.mark(8);
    NSC_CHECK(bwselect_impl(W.t2, W.bw1_t, W.seg_norbc, w, h, 8, W.L, W.tmp8, s));
    NSC_CHECK(C.save(10, W.seg_norbc, n));
    hipLaunchKernelGGL(k_norbc, grid1(n), dim3(256), 0, s, W.seg_norbc, W.rbcsel, W.r2b, C.want(2) ? W.rbc : nullptr, n);
    NSC_CHECK(C.save(2, W.rbc, n));
    NSC_CHECK(C.save(11, W.seg_norbc, n));
    C.mark(10);
    int kept = 0;  // bwareaopen2(fill1, 11..1000) count (HE.cpp:272-274)
    NSC_HIP(hipMemcpyAsync(&kept, &W.counts[1], sizeof(int), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    return kept == 0 ? NSC_SEG_NO_CANDIDATES_LEFT : NSC_SEG_CONTINUE;
}

// plSeparateNuclei (HistologicalEntities.cpp:308-449) on seg_open: seg_nonoverlap in W.nonov.
int stage_a4(Chain& C, const uint8_t* bgr, const uint8_t* seg_open, int w, int h, int minSizePl, int wsConn) {
    Workspace& W = C.W;
    hipStream_t s = C.s;
    const size_t n = (size_t)w * h;
    NSC_CHECK(bwareaopen_impl(seg_open, W.seg_big_t, w, h, minSizePl, INT_MAX, 8, nullptr, W.L, W.area, s));
    NSC_CHECK(C.save(14, W.seg_big_t, n));
    C.mark(14);
    NSC_CHECK(morph_impl(W.seg_big_t, W.seg_big, w, h, kCross3, 3, 0, false, s));  // HE.cpp:330
    NSC_CHECK(C.save(15, W.seg_big, n));
    C.mark(15);
    NSC_CHECK(edt_impl(W.seg_big, W.dist, w, h, s));  // HE.cpp:359
    NSC_CHECK(C.save(16, W.dist, 4 * n));
    hipLaunchKernelGGL(k_dist_prep, grid1(n), dim3(256), 0, s, W.dist, W.seg_big, bgr, n, C.want(17) ? W.distance : nullptr,
                       W.hmask, W.hmarker, W.nuclei);
    NSC_CHECK(C.save(17, W.distance, 4 * n));
    C.mark(16);
    NSC_CHECK(recon_impl<float>(W.hmarker, W.hmask, W.d2, w, h, 8, s));  // imhmin(distance, 1, 8) HE.cpp:384
    hipLaunchKernelGGL(k_negate_f, grid1(n), dim3(256), 0, s, W.d2, n);
    NSC_CHECK(C.save(18, W.d2, 4 * n));
    NSC_CHECK(C.save(19, W.nuclei, 3 * n));
    C.mark(18);
    // watershed2(nuclei, distance2, watershedConnectivity)  HE.cpp:414
    NSC_CHECK(local_max_impl(W.d2, W.tmp8, w, h, wsConn, true, W.big, s));
    NSC_CHECK(ccl_u8(W.tmp8, W.L, w, h, wsConn, true, -1, s));
    NSC_CHECK(ws2_from_labels(W.nuclei, W.L, W.wm, w, h, W.big, s));
    NSC_CHECK(C.save(20, W.wm, 4 * n));
    hipLaunchKernelGGL(k_nonoverlap, grid1(n), dim3(256), 0, s, W.seg_big, W.wm, W.nonov, n);
    NSC_CHECK(C.save(21, W.nonov, n));
    C.mark(20);
    NSC_HIP(hipGetLastError());
    return NSC_SEG_CONTINUE;
}

// boxes of the final labels (contiguous lab) as io->bbox_mode asks; nuclei = number of roots.
// Called after the chain's last synchronisation, with W.L holding the canonical relabel of lab
// unless lab itself was relabelled (io->canonical).
int chain_boxes(Chain& C, int32_t* lab, int w, int h, nsc_seg_io* io, int nuclei) {
    Workspace& W = C.W;
    hipStream_t s = C.s;
    const int n = w * h;
    io->compcount = 0;
    if (io->bbox_mode == NSC_BBOX_REFERENCE_CPU) {
        // ConnComponents::boundingBox(labels, bgval 0) on the reference CPU labels (HE.cpp:590-591)
        int cnt = 0;
        NSC_CHECK(bbox_literal_impl(lab, w, h, 0, io->bbox5, io->max_boxes, &cnt, W.L, W.area, W.wm, &W.counts[4], s));
        io->compcount = cnt;
        if (cnt > io->max_boxes || (cnt > 0 && !io->bbox5)) {
            set_error("segment_nuclei: %d boxes exceed the capacity of %d", cnt, io->max_boxes);
            return NSC_ERR_LIMIT;
        }
        NSC_HIP(stream_sync(s));
        return NSC_OK;
    }
    if (io->bbox_mode == NSC_BBOX_ALL) {
        io->compcount = nuclei;
        if (nuclei == 0) return NSC_OK;
        if (nuclei > io->max_boxes || !io->bbox5) {
            set_error("segment_nuclei: %d boxes exceed the capacity of %d", nuclei, io->max_boxes);
            return NSC_ERR_LIMIT;
        }
        if (io->canonical) {
            NSC_CHECK(bbox_impl(lab, w, h, nuclei, io->bbox5, io->area, s));
        } else {
            // W.L holds the canonical copy the nuclei count was taken from
            NSC_CHECK(bbox_impl(W.L, w, h, nuclei, io->bbox5, io->area, s));
            hipLaunchKernelGGL(k_root_label_column, grid1(n), dim3(256), 0, s, lab, W.L, n, io->bbox5);
        }
        NSC_HIP(hipGetLastError());
        NSC_HIP(stream_sync(s));
    }
    return NSC_OK;
}

int segment_body(const uint8_t* bgr_in, int w, int h, const nsc_seg_params& p, int32_t* labels_out, nsc_seg_io* io,
                 Workspace& W, void** stage, double* stage_ms, hipStream_t s) {
    const size_t n = (size_t)w * h;
    W.arena.top = 0;
    W.arena.depth = 0;
    ArenaScope arena_scope(&W.arena);
    Chain C(W, s, stage, stage_ms);
    const size_t bgr_step = (io && io->bgr_step) ? io->bgr_step : (size_t)3 * w;
    const size_t lab_step = (io && io->labels_step) ? io->labels_step : (size_t)4 * w;
    NSC_ARG(bgr_step >= (size_t)3 * w && lab_step >= (size_t)4 * w && lab_step % 4 == 0, "segment_nuclei: row steps");
    // pitched images: continuous copies in the workspace arena (released in reverse order at return)
    Scratch bgr_c, lab_c;
    const uint8_t* bgr = bgr_in;
    if (bgr_step != (size_t)3 * w) {
        NSC_CHECK(bgr_c.alloc(3 * n, s));
        NSC_HIP(hipMemcpy2DAsync(bgr_c.p, (size_t)3 * w, bgr_in, bgr_step, (size_t)3 * w, h, hipMemcpyDeviceToDevice, s));
        bgr = bgr_c.as<uint8_t>();
    }
    int32_t* labels = labels_out;
    if (lab_step != (size_t)4 * w) {
        NSC_CHECK(lab_c.alloc(4 * n, s));
        labels = lab_c.as<int32_t>();
    }

    int st = stage_s1(C, bgr, w, h, p);
    if (st != NSC_SEG_CONTINUE) {
        C.finish_timing();
        return st;
    }
    // ---- segmentNuclei body (HE.cpp:494-499) ----
    NSC_CHECK(fill_holes_impl(W.seg_norbc, W.seg_nohole, w, h, 4, W.L, W.tmp8, s));  // HE.cpp:494
    NSC_CHECK(C.save(12, W.seg_nohole, n));
    C.mark(12);
    NSC_CHECK(morph_open_impl(W.seg_nohole, W.seg_open, W.tmp8, w, h, kCross3, 3, s));  // HE.cpp:498-499
    NSC_CHECK(C.save(13, W.seg_open, n));
    C.mark(13);
    st = stage_a4(C, bgr, W.seg_open, w, h, p.minSizePl, p.watershedConnectivity);
    if (st != NSC_SEG_CONTINUE) return st;

    NSC_CHECK(bwareaopen_impl(W.nonov, W.seg, w, h, p.minSizeSeg, p.maxSizeSeg, 4, &W.counts[2], W.L, W.area, s));  // :548
    NSC_CHECK(C.save(23, W.seg, n));
    C.mark(23);
    NSC_CHECK(fill_holes_impl(W.seg, W.fin, w, h, p.fillHolesConnectivity, W.L, W.tmp8, s));  // :573
    NSC_CHECK(C.save(24, W.fin, n));
    C.mark(24);
    NSC_CHECK(ccl_u8(W.fin, labels, w, h, 8, true, -1, s));  // bwlabel2(final, 8) :585
    NSC_CHECK(C.save(25, labels, 4 * n));
    // number of nuclei = number of roots (rank scratch in W.area)
    const bool canonical = io && io->canonical;
    const bool cpu_boxes = io && io->bbox_mode == NSC_BBOX_REFERENCE_CPU;
    if (canonical && !cpu_boxes) {
        NSC_CHECK(relabel_impl(labels, w, h, -1, &W.counts[3], W.area, s));
    } else if (io && io->bbox_mode == NSC_BBOX_ALL) {
        // the boxes of the reference's labels come from a canonical copy
        NSC_HIP(hipMemcpyAsync(W.L, labels, 4 * n, hipMemcpyDeviceToDevice, s));
        NSC_CHECK(relabel_impl(W.L, w, h, -1, &W.counts[3], W.area, s));
    } else {
        NSC_CHECK(root_count_impl(labels, w, h, &W.counts[3], s));  // the nucleus count only
    }
    C.mark(25);
    NSC_HIP(hipGetLastError());
    int c[4] = {0, 0, 0, 0};
    NSC_HIP(hipMemcpyAsync(c, W.counts, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    C.finish_timing();
    if (c[2] == 0) return NSC_SEG_NO_CANDIDATES_LEFT;  // HE.cpp:554-556
    if (io) {
        io->nuclei = c[3];
        NSC_CHECK(chain_boxes(C, labels, w, h, io, c[3]));
        if (canonical && cpu_boxes) NSC_CHECK(relabel_impl(labels, w, h, -1, &W.counts[3], W.area, s));
    }
    if (labels != labels_out)
        NSC_HIP(hipMemcpy2DAsync(labels_out, lab_step, labels, (size_t)4 * w, (size_t)4 * w, h, hipMemcpyDeviceToDevice, s));
    NSC_HIP(stream_sync(s));
    return NSC_SEG_SUCCESS;
}

// A failed call may leave work queued on the stream against the workspace: wait for it before
// returning, so a retry (possibly on another stream) cannot reuse buffers still in use.
int settle(int rc, hipStream_t s) {
    if (rc <= NSC_ERR_ARG) {
        (void)hipStreamSynchronize(s);
        // work the chain forked onto this thread's side streams (and never joined on error)
        for (hipStream_t side : side_stream().side)
            if (side) (void)hipStreamSynchronize(side);
    }
    return rc;
}

int check_chain_args(const uint8_t* bgr, int w, int h, const nsc_seg_params* prm, const nsc_workspace* ws) {
    NSC_ARG(bgr && prm && ws && w > 0 && h > 0, "segment_nuclei");
    NSC_ARG(workspace_fits(ws, w, h), "tile larger than the workspace");
    NSC_LABEL_SIZE(w, h);
    NSC_ARG(valid_conn(prm->fillHolesConnectivity) && valid_conn(prm->reconConnectivity) &&
                valid_conn(prm->watershedConnectivity), "connectivity");
    return ensure_init();
}

}  // namespace
}  // namespace nsc

extern "C" {

int nsc_segment_nuclei_io(const uint8_t* bgr, int w, int h, const nsc_seg_params* prm, int32_t* labels, nsc_seg_io* io,
                          nsc_workspace* ws, void** stage, double* stage_ms, void* stream) {
    if (!bgr || w <= 0 || h <= 0) return NSC_SEG_INVALID_IMAGE;  // HE.cpp:477 (no image data)
    NSC_ARG(labels, "segment_nuclei: labels");
    NSC_CHECK(check_chain_args(bgr, w, h, prm, ws));
    NSC_ARG(!io || (io->bbox_mode >= NSC_BBOX_NONE && io->bbox_mode <= NSC_BBOX_ALL), "segment_nuclei: bbox_mode");
    if (io) {
        io->compcount = 0;
        io->nuclei = 0;
    }
    hipStream_t s = as_stream(stream);
    return settle(segment_body(bgr, w, h, *prm, labels, io, *ws, stage, stage_ms, s), s);
}

int nsc_segment_nuclei(const uint8_t* bgr, int w, int h, const nsc_seg_params* prm, int32_t* labels, int* compcount,
                       nsc_workspace* ws, void** stage, double* stage_ms, void* stream) {
    nsc_seg_io io;
    memset(&io, 0, sizeof(io));
    const int st = nsc_segment_nuclei_io(bgr, w, h, prm, labels, &io, ws, stage, stage_ms, stream);
    if (compcount && st == NSC_SEG_SUCCESS) *compcount = io.nuclei;
    return st;
}

int nsc_pl_find_nucleus_candidates(const uint8_t* bgr, size_t bgr_step, int w, int h, const nsc_seg_params* prm,
                                   uint8_t* seg_norbc, size_t seg_step, nsc_workspace* ws, void* stream) {
    if (!bgr || w <= 0 || h <= 0) return NSC_SEG_INVALID_IMAGE;
    NSC_ARG(seg_norbc, "plFindNucleusCandidates: seg_norbc");
    NSC_CHECK(check_chain_args(bgr, w, h, prm, ws));
    if (!bgr_step) bgr_step = (size_t)3 * w;
    if (!seg_step) seg_step = (size_t)w;
    NSC_ARG(bgr_step >= (size_t)3 * w && seg_step >= (size_t)w, "plFindNucleusCandidates: row steps");
    hipStream_t s = as_stream(stream);
    Workspace& W = *ws;
    auto body = [&]() -> int {
        W.arena.top = 0;
        W.arena.depth = 0;
        ArenaScope arena_scope(&W.arena);
        Chain C(W, s, nullptr, nullptr);
        Scratch bgr_c;
        const uint8_t* in = bgr;
        if (bgr_step != (size_t)3 * w) {
            NSC_CHECK(bgr_c.alloc((size_t)3 * w * h, s));
            NSC_HIP(hipMemcpy2DAsync(bgr_c.p, (size_t)3 * w, bgr, bgr_step, (size_t)3 * w, h, hipMemcpyDeviceToDevice, s));
            in = bgr_c.as<uint8_t>();
        }
        const int st = stage_s1(C, in, w, h, *prm);
        if (st != NSC_SEG_CONTINUE) return st;
        NSC_HIP(hipMemcpy2DAsync(seg_norbc, seg_step, W.seg_norbc, (size_t)w, (size_t)w, h, hipMemcpyDeviceToDevice, s));
        NSC_HIP(stream_sync(s));
        return st;
    };
    return settle(body(), s);
}

int nsc_pl_separate_nuclei(const uint8_t* bgr, size_t bgr_step, const uint8_t* seg_open, size_t open_step, int w, int h,
                           int minSizePl, int watershedConnectivity, uint8_t* seg_nonoverlap, size_t out_step,
                           nsc_workspace* ws, void* stream) {
    if (!bgr || w <= 0 || h <= 0) return NSC_SEG_INVALID_IMAGE;
    NSC_ARG(seg_open && seg_nonoverlap && ws && valid_conn(watershedConnectivity), "plSeparateNuclei");
    NSC_ARG(workspace_fits(ws, w, h), "tile larger than the workspace");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    if (!bgr_step) bgr_step = (size_t)3 * w;
    if (!open_step) open_step = (size_t)w;
    if (!out_step) out_step = (size_t)w;
    NSC_ARG(bgr_step >= (size_t)3 * w && open_step >= (size_t)w && out_step >= (size_t)w, "plSeparateNuclei: row steps");
    hipStream_t s = as_stream(stream);
    Workspace& W = *ws;
    auto body = [&]() -> int {
        W.arena.top = 0;
        W.arena.depth = 0;
        ArenaScope arena_scope(&W.arena);
        Chain C(W, s, nullptr, nullptr);
        Scratch bgr_c;
        const uint8_t* in = bgr;
        if (bgr_step != (size_t)3 * w) {
            NSC_CHECK(bgr_c.alloc((size_t)3 * w * h, s));
            NSC_HIP(hipMemcpy2DAsync(bgr_c.p, (size_t)3 * w, bgr, bgr_step, (size_t)3 * w, h, hipMemcpyDeviceToDevice, s));
            in = bgr_c.as<uint8_t>();
        }
        NSC_HIP(hipMemcpy2DAsync(W.seg_open, (size_t)w, seg_open, open_step, (size_t)w, h, hipMemcpyDeviceToDevice, s));
        const int st = stage_a4(C, in, W.seg_open, w, h, minSizePl, watershedConnectivity);
        if (st != NSC_SEG_CONTINUE) return st;
        NSC_HIP(hipMemcpy2DAsync(seg_nonoverlap, out_step, W.nonov, (size_t)w, (size_t)w, h, hipMemcpyDeviceToDevice, s));
        NSC_HIP(stream_sync(s));
        return st;
    };
    return settle(body(), s);
}

}  // extern "C"
