// Exact parallel decomposition of cv::watershed's sequential flood (OpenCV 2.4.9
// imgproc/segmentation.cpp) for marker images whose unknown region contains a large black
// region B -- the shape watershed2 gets inside segmentNuclei: the background plateau holds label
// 0 (it contains pixel (0,0), MorphologicOperations.cpp:3737 labels a component by its smallest
// index) and is therefore "unknown", and the masked colour image is (0,0,0) there
// (HistologicalEntities.cpp:400-407).  Every other unknown pixel is in F (non-black).
//
// Facts the decomposition rests on (each B-B colour difference is 0, every B-F difference is >0):
//  1. Runs = 4-connected components of F.  Until the first pop of a B pixel that receives a
//     positive label (call it b1, at time T1), runs do not interact: a B pixel popped as -1 changes
//     nothing any other pixel reads.  Each run is flooded on its own lane ("phase 1"); B
//     neighbours are recorded as leaf items (when, at what level and by whom they are queued).
//  2. The global pop order of the merged runs is a total order computable from per-run data: for
//     items x, y of different runs, x pops first iff the largest (level, push time) key on x's
//     push chain is smaller than y's, comparing push times recursively through the pushers
//     (a pop can only be delayed by a larger key on its own chain).  `before()` implements it.
//  3. b1 = the earliest B item whose neighbours popped before it agree on one label Y1.  At T1 the
//     level-0 FIFO holds nothing else, so the flood from b1 runs to completion through every B
//     pixel not yet queued that it reaches (label Y1: such a pixel has no labelled F neighbour,
//     which would have queued it) before anything else pops.
//  4. After that, the unresolved pixels fall into 4-connected components that are again
//     independent ("phase 2"), each flooded on its own lane from the queue state at T1.  The only
//     order not known locally is the FIFO order among F pixels queued by the flood from b1 at the
//     same level; for those the lane checks that every permutation gives the same labels.
// Anything outside these assumptions (a capacity, a comparison that does not terminate, a tie
// whose order matters) raises a flag and the caller re-runs the plain sequential flood.
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "common.hpp"

namespace nsc {

int ccl_neq(const uint8_t* img, uint8_t key, int32_t* labels, int w, int h, int conn, hipStream_t s);
int relabel_impl(int32_t* labels, int w, int h, int32_t bgval, int* dcount, int32_t* rank, hipStream_t s);
int bbox_impl(const int32_t* canon, int w, int h, int n, int32_t* out5, int32_t* area, hipStream_t s);

namespace {

constexpr int NQ = 256;
constexpr int IN_QUEUE = -2, WSHED = -1;
constexpr uint16_t NOPUSH = 0xFFFF;
enum { C_FRAME = 0, C_SEED = 1, C_F = 2, C_B = 3 };
enum { FB_LEAF_CAP = 1, FB_CMP = 2, FB_TIES = 4, FB_HEAP = 8, FB_TIEPERM = 16 };
constexpr int CMP_MAX_STEPS = 1 << 22;

struct Dev {
    // image
    const uint8_t* img;
    int W, H, N;
    const int32_t* m0;
    int32_t* m;
    const uint8_t* cls;
    const uint16_t* lvl0;
    const int32_t* run;  // canonical run id (1..R) for F pixels
    // F items (indexed by pixel)
    uint8_t* st;     // 0 none, 1 queued, 2 popped
    uint8_t* ilvl;
    int32_t* pseq;
    int32_t* parent;  // pusher pixel, -1 = initial push
    int32_t* bnode;   // item id
    int32_t* popidx;
    int32_t* lab;     // label decided in the run's flood
    int32_t* qnext;   // FIFO links, indexed by item id
    // leaves
    int32_t* head;    // per B pixel: first leaf slot or -1
    int32_t* lf_b;
    int32_t* lf_run;
    uint8_t* lf_lvl;
    int32_t* lf_pseq;
    int32_t* lf_parent;
    int32_t* lf_bnode;
    int32_t* lf_popidx;
    int32_t* lf_next;
    int* nleaf;
    int leafcap;
    int* flags;
    // per item, packed after phase 1 for the pop-order walks: {run, bnode, popidx, parent} and level
    int4* wr;
    uint8_t* wl;
};

// A value every lane of the wave holds (or the one active lane's), as a wave-uniform value: the
// compiler then keeps it in a scalar register and branches on it with scalar branches.
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The label a popped pixel takes from its four neighbours (cv::watershed): the common positive
// label, WSHED where two positive labels differ, 0 where none is positive.  Branch-free: the
// largest label, and the smallest positive one found as the unsigned minimum of t - 1 (0, WSHED and
// IN_QUEUE wrap to the top).
__device__ __forceinline__ int pop_label(int a, int b, int c, int d) {
    const int mx = max(max(a, b), max(c, d));
    const unsigned mn1 = min(min((unsigned)(a - 1), (unsigned)(b - 1)), min((unsigned)(c - 1), (unsigned)(d - 1)));
    return mx <= 0 ? 0 : ((int)mn1 + 1 == mx ? mx : WSHED);
}

__device__ __forceinline__ int c_diff(const uint8_t* a, const uint8_t* b) {
    const int db = abs((int)a[0] - (int)b[0]);
    const int dg = abs((int)a[1] - (int)b[1]);
    const int dr = abs((int)a[2] - (int)b[2]);
    return max(max(db, dg), dr);
}

// ---- item accessors: id < N -> F pixel or initially queued B pixel; id >= N -> leaf slot ----
__device__ __forceinline__ int it_lvl(const Dev& D, int x) {
    if (x >= D.N) return D.lf_lvl[x - D.N];
    return D.cls[x] == C_F ? D.ilvl[x] : D.lvl0[x];
}
__device__ __forceinline__ int it_parent(const Dev& D, int x) {
    if (x >= D.N) return D.lf_parent[x - D.N];
    return D.cls[x] == C_F ? D.parent[x] : -1;
}
__device__ __forceinline__ int it_pseq(const Dev& D, int x) {
    if (x >= D.N) return D.lf_pseq[x - D.N];
    return D.cls[x] == C_F ? D.pseq[x] : 0;
}

// x pops before y in the merged phase-1 flood (x != y); sets FB_CMP if the walk is too long.
// Reads the packed item records (k_pack_items), so a step is two dependent loads: the records of
// x and y, then those of their chains' largest-key items.
__device__ bool before_from(const Dev& D, int x, int y, int step0) {
    for (int step = step0; step < CMP_MAX_STEPS; ++step) {
        if (x == y) return false;
        const int4 rx = D.wr[x], ry = D.wr[y];  // {run, bnode, popidx, parent}
        if (rx.x == ry.x) return rx.z < ry.z;
        const int a = rx.y, b = ry.y;
        const int4 ra = D.wr[a], rb = D.wr[b];
        const int la = D.wl[a], lb = D.wl[b];
        if (la != lb) return la < lb;
        const int pa = ra.w, pb = rb.w;
        if (pa < 0 && pb < 0) return a < b;  // both queued in the initial raster scan
        if (pa < 0) return true;
        if (pb < 0) return false;
        x = pa;
        y = pb;
    }
    atomicOr(D.flags, FB_CMP);
    return false;
}
__device__ __forceinline__ bool before(const Dev& D, int x, int y) { return before_from(D, x, y, 0); }

// What the first step of before() reads about an item: its run and pop index, its chain's
// largest-key item, that item's level and pusher.  A tournament that compares the same winners
// level after level loads this once per candidate and carries it along.
struct ItemKey {
    int item, run, popidx, top, lvl, par;
};
__device__ __forceinline__ ItemKey item_key(const Dev& D, int x) {
    const int4 rx = D.wr[x];
    const int4 ra = D.wr[rx.y];
    return ItemKey{x, rx.x, rx.z, rx.y, (int)D.wl[rx.y], ra.w};
}
// before(D, kx.item, ky.item), its first step taken from the keys
__device__ __forceinline__ bool before_keyed(const Dev& D, const ItemKey& kx, const ItemKey& ky) {
    if (kx.item == ky.item) return false;
    if (kx.run == ky.run) return kx.popidx < ky.popidx;
    if (kx.lvl != ky.lvl) return kx.lvl < ky.lvl;
    if (kx.par < 0 && ky.par < 0) return kx.top < ky.top;
    if (kx.par < 0) return true;
    if (ky.par < 0) return false;
    return before_from(D, kx.par, ky.par, 1);
}

// ---- classification ----
// nF: this workgroup's count goes to partial[blockIdx.x] (summed by k_sum_partials): one atomic
// per workgroup on a single counter would serialise over a grid this large.
__global__ void k_classify(Dev D, uint8_t* cls, uint16_t* lvl0, uint8_t* fmask, int* partial) {
    const int W = D.W, H = D.H;
    int local = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < D.N; i += gridDim.x * blockDim.x) {
        const int x = i % W, y = i / W;
        const bool frame = (x == 0 || y == 0 || x == W - 1 || y == H - 1);
        // the pixel's and its four neighbours' markers and colours, all loads issued together
        // (a frame pixel, whose neighbours may leave the image, reads itself five times and
        // ignores the copies); the decisions below are selects on loaded values
        const int nb[4] = {frame ? i : i - 1, frame ? i : i + 1, frame ? i : i - W, frame ? i : i + W};
        const int v = D.m0[i];
        int vn[4];
        uint8_t pc[3], nc[4][3];
#pragma unroll
        for (int k = 0; k < 4; ++k) vn[k] = D.m0[nb[k]];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            pc[ch] = D.img[3 * (size_t)i + ch];
#pragma unroll
            for (int k = 0; k < 4; ++k) nc[k][ch] = D.img[3 * (size_t)nb[k] + ch];
        }
        int best = 256;  // smallest colour step to a neighbouring seed off the frame
        const int nx[4] = {x - 1, x + 1, x, x}, ny[4] = {y, y, y - 1, y + 1};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool pos = !(nx[k] == 0 || ny[k] == 0 || nx[k] == W - 1 || ny[k] == H - 1) && vn[k] > 0;
            const int d = c_diff(pc, nc[k]);
            best = pos ? min(best, d) : best;
        }
        const bool black = pc[0] == 0 && pc[1] == 0 && pc[2] == 0;
        const uint8_t c = frame ? C_FRAME : (v > 0 ? C_SEED : (black ? C_B : C_F));
        const uint16_t L = (frame || v > 0 || best >= 256) ? NOPUSH : (uint16_t)best;
        cls[i] = c;
        lvl0[i] = L;
        fmask[i] = (c == C_F) ? 1 : 0;
        D.st[i] = 0;
        D.head[i] = -1;
        D.m[i] = frame ? WSHED : (v < 0 ? 0 : v);
        local += (c == C_F);
    }
    const int bs = block_sum(local);
    if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// *out = sum of partial[0..n) (one workgroup of 1024 threads)
__global__ __launch_bounds__(1024) void k_sum_partials(const int* __restrict__ partial, int n, int* out) {
    int v = 0;
    for (int i0 = threadIdx.x; i0 < n; i0 += 4 * blockDim.x) {  // four loads in flight per thread
        int p[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) p[j] = i0 + j * (int)blockDim.x < n ? partial[i0 + j * blockDim.x] : 0;
        v += p[0] + p[1] + p[2] + p[3];
    }
    const int t = block_sum(v);
    if (threadIdx.x == 0) *out = t;
}

// ---- phase 1: one lane floods one run ----
__global__ void k_phase1(Dev D, const int32_t* __restrict__ box5, int nrun, const int* __restrict__ list,
                         const int* __restrict__ count, int* __restrict__ qh, int* __restrict__ qt) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= *count) return;
    const int r = list[idx];
    const int W = D.W;
    const int id = r + 1;
    int* head = qh + (size_t)r * NQ;  // per run (run ids are unique across the tiers)
    int* tail = qt + (size_t)r * NQ;
    for (int l = 0; l < NQ; ++l) { head[l] = -1; tail[l] = -1; }
    int active = NQ, seq = 0, popc = 0;
    auto enq = [&](int l, int x) {
        D.qnext[x] = -1;
        if (tail[l] >= 0) D.qnext[tail[l]] = x;
        else head[l] = x;
        tail[l] = x;
        if (l < active) active = l;
    };
    // key of an F item or leaf of this run: (level, pseq)
    auto key_less = [&](int a, int b) {
        const int la = it_lvl(D, a), lb = it_lvl(D, b);
        if (la != lb) return la < lb;
        return it_pseq(D, a) < it_pseq(D, b);
    };
    const int minx = box5[nrun + r], maxx = box5[2 * nrun + r];
    const int miny = box5[3 * nrun + r], maxy = box5[4 * nrun + r];
    for (int y = miny; y <= maxy; ++y)
        for (int x = minx; x <= maxx; ++x) {
            const int i = y * W + x;
            if (D.cls[i] == C_F && D.run[i] == id && D.lvl0[i] != NOPUSH) {
                D.ilvl[i] = (uint8_t)D.lvl0[i];
                D.pseq[i] = seq++;
                D.parent[i] = -1;
                D.bnode[i] = i;
                D.st[i] = 1;
                enq(D.lvl0[i], i);
            }
        }
    const int dirs[4] = {-1, 1, -W, W};
    for (;;) {
        if (active >= NQ) break;
        if (head[active] < 0) {
            int l = active + 1;
            while (l < NQ && head[l] < 0) ++l;
            if (l == NQ) break;
            active = l;
        }
        const int x = head[active];
        const int nx = D.qnext[x];
        head[active] = nx;
        if (nx < 0) tail[active] = -1;
        if (x >= D.N) {  // leaf: record when it pops, nothing else
            D.lf_popidx[x - D.N] = popc++;
            continue;
        }
        const int p = x;
        D.popidx[p] = popc++;
        int lab = 0;
        for (int d = 0; d < 4; ++d) {
            const int s = p + dirs[d];
            int t = 0;
            const uint8_t cs = D.cls[s];
            if (cs == C_SEED) t = D.m0[s];
            else if (cs == C_F && D.st[s] == 2) t = D.lab[s];
            if (t > 0) { if (lab == 0) lab = t; else if (t != lab) lab = WSHED; }
        }
        D.lab[p] = lab;
        D.st[p] = 2;
        if (lab == WSHED) continue;
        const uint8_t* ptr = D.img + 3 * (size_t)p;
        const int bp = D.bnode[p];
        for (int d = 0; d < 4; ++d) {
            const int s = p + dirs[d];
            const uint8_t cs = D.cls[s];
            if (cs == C_F) {
                if (D.st[s] != 0) continue;
                const int l = c_diff(ptr, D.img + 3 * (size_t)s);
                D.ilvl[s] = (uint8_t)l;
                D.pseq[s] = seq++;
                D.parent[s] = p;
                D.bnode[s] = key_less(bp, s) ? s : bp;
                D.st[s] = 1;
                enq(l, s);
            } else if (cs == C_B && D.lvl0[s] == NOPUSH) {
                bool done = false;  // queued already by an earlier labelled neighbour of this run
                for (int e = 0; e < 4; ++e) {
                    const int f2 = s + dirs[e];
                    if (f2 != p && D.cls[f2] == C_F && D.run[f2] == id && D.st[f2] == 2 && D.lab[f2] > 0) done = true;
                }
                if (done) continue;
                const int slot = atomicAdd(D.nleaf, 1);
                if (slot >= D.leafcap) { atomicOr(D.flags, FB_LEAF_CAP); continue; }
                const int l = c_diff(ptr, D.img + 3 * (size_t)s);
                D.lf_b[slot] = s;
                D.lf_run[slot] = id;
                D.lf_lvl[slot] = (uint8_t)l;
                D.lf_pseq[slot] = seq++;
                D.lf_parent[slot] = p;
                D.lf_popidx[slot] = -1;
                const int lid = D.N + slot;
                D.lf_bnode[slot] = key_less(bp, lid) ? lid : bp;  // linked into head[s] by k_link_leaves
                enq(l, lid);
            }
        }
    }
}

// Store to element i of a per-pixel or per-leaf array through a 32-bit byte offset (the
// decomposition runs on images below 2^28 pixels, so every offset fits): with a uniform base the
// compiler uses the scalar-base form of the vector store instead of forming a 64-bit address in
// vector registers for each store.
template <typename T, typename V>
__device__ __forceinline__ void gst(T* base, uint32_t i, V v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (size_t)(i * (uint32_t)sizeof(T))) = (T)v;
}

// ---- phase 1 with the run's neighbourhood staged in LDS ----
// Same flood as k_phase1, for runs whose bounding box plus a 2-pixel margin has at most LCAP
// pixels.  The wave stages the frame into LDS (class, in-run flag, seed labels, and the push level
// of every pixel towards its right and lower neighbour, from the colours), numbers the initial
// items in raster order with ballots and links them into their level FIFOs chunk by chunk; lane 0
// then runs the flood with every read in LDS, writing the per-item records to global memory as it
// goes.  Each queued F pixel carries the key (level, push sequence) and frame index of its chain's
// largest-key item, so a push compares keys without following pointers.  The 2-pixel margin keeps
// every neighbour of a leaf candidate inside the frame.  Leaf slots are reserved per run up front
// (one per B pixel of the frame that could become a leaf), so no atomic sits on the flood's
// critical path; unused slots keep lf_b = -1.
// LDS per frame pixel: 15 bytes (class 1, label 4, chain key 4, push levels 2, chain item or leaf
// index 2, FIFO link 2), so a CU holds as many concurrent floods as its LDS allows for the tier.
constexpr uint16_t NIL16 = 0xFFFF;
// phase-1 FIFO links: an item index below 0x7FFF (frames hold at most 6400 pixels), P1_LEAF when
// the item is a leaf, P1_NIL for none
constexpr uint32_t P1_NIL = 0x7FFF, P1_LEAF = 0x8000;
enum { K_INRUN = 4, K_ST1 = 8, K_ST2 = 16, K_LV = 32 };
constexpr int P1_MARGIN = 2;

__device__ __forceinline__ int cdiff_packed(uint32_t a, uint32_t b) {
    const int d0 = abs((int)(a & 255) - (int)(b & 255));
    const int d1 = abs((int)((a >> 8) & 255) - (int)((b >> 8) & 255));
    const int d2 = abs((int)((a >> 16) & 255) - (int)((b >> 16) & 255));
    return max(max(d0, d1), d2);
}

template <int LCAP>
__global__ __launch_bounds__(64) void k_phase1_lds(Dev D, const int32_t* __restrict__ box5, int nrun,
                                                   const int* __restrict__ list, const int* __restrict__ count,
                                                   int* __restrict__ next_run) {
    __shared__ uint8_t code[LCAP];
    __shared__ int32_t lk[LCAP];     // known label: seed label, label of a popped F pixel, else 0
    __shared__ uint32_t bkey[LCAP];  // queued F pixel: (level << 16 | push seq) of its chain's largest-key item
                                     // (the pixel's colour while the frame is staged)
    __shared__ uint16_t dif[LCAP];   // push level towards the right (low byte) and lower (high byte) neighbour
    __shared__ uint16_t bloc[LCAP];  // queued F pixel: frame index of that item; B pixel: its leaf index
    __shared__ uint16_t qn[LCAP];    // FIFO link of the item at this pixel (an F item, or a B pixel's leaf)
    __shared__ uint16_t qhead[NQ], qtail[NQ];
    __shared__ int sh_base, sh_it;

    // a fixed grid takes the tier's runs from a work counter, one at a time, so a workgroup that
    // drew short runs takes more of them (a fixed stride over the list left the tier's time set by
    // the workgroup that drew several long floods); launching one workgroup per run of the whole
    // image (most of them idle for this tier) would hold the large LDS allocation of every idle
    // workgroup across the device
    const int nr = *count;
    for (;;) {
    if (threadIdx.x == 0) sh_it = atomicAdd(next_run, 1);
    __syncthreads();
    const int it = sh_it;
    if (it >= nr) break;
    [&]() {  // one run; `return` ends the run
    const int r = list[it];
    const int lane = threadIdx.x;
    const int W = D.W, H = D.H, id = r + 1;
    const int ox = box5[nrun + r] - P1_MARGIN, oy = box5[3 * nrun + r] - P1_MARGIN;
    const int lw = box5[2 * nrun + r] - ox + 1 + P1_MARGIN, lh = box5[4 * nrun + r] - oy + 1 + P1_MARGIN;
    const int L = lw * lh;  // <= LCAP (bucketed by the host)
    // k / lw by multiply-high (exact for k < 2^16)
    const uint32_t magic = (uint32_t)((0x100000000ull + (unsigned)lw - 1) / (unsigned)lw);
    const int gbase = oy * W + ox;
    auto gidx = [&](int k) {
        const int ky = (int)__umulhi((uint32_t)k, magic);
        return gbase + ky * W + (k - ky * lw);
    };

    for (int k = lane; k < L; k += 64) {
        const int ky = (int)__umulhi((uint32_t)k, magic), kx = k - ky * lw;
        const int gx = ox + kx, gy = oy + ky;
        uint8_t cd = C_FRAME;
        int32_t l0 = 0;
        uint32_t cc = 0;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {  // the margin may step one pixel off the image
            const int g = gy * W + gx;
            const uint8_t c = D.cls[g];
            cd = c;
            if (c == C_F && D.run[g] == id) cd |= K_INRUN;
            if (c == C_SEED) l0 = D.m0[g];
            if (D.lvl0[g] != NOPUSH) cd |= K_LV;
            const uint8_t* q = D.img + 3 * (size_t)g;
            cc = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16);
        }
        code[k] = cd;
        lk[k] = l0;
        bkey[k] = cc;
    }
    for (int l = lane; l < NQ; l += 64) { qhead[l] = P1_NIL; qtail[l] = P1_NIL; }
    __syncthreads();
    // push levels between 4-neighbours (a pair that wraps around a frame row is never used: pushes
    // start at run pixels, which lie the margin away from the frame's edge)
    for (int k = lane; k < L; k += 64) {
        const uint32_t c = bkey[k];
        const int rt = k + 1 < L ? cdiff_packed(c, bkey[k + 1]) : 0;
        const int dn = k + lw < L ? cdiff_packed(c, bkey[k + lw]) : 0;
        dif[k] = (uint16_t)(rt | (dn << 8));
    }
    __syncthreads();
    // leaf candidates (B, never initially queued, 4-adjacent to the run) and initial items in
    // raster order: an initial item's records are written here, and the chunk's items of one level
    // are linked behind that level's FIFO tail (the first of them reads the tail, the last moves it)
    int cnt = 0, ninit = 0;
    for (int k0 = 0; k0 < L; k0 += 64) {
        const int k = k0 + lane;
        bool cand = false, init = false;
        if (k < L) {
            const uint8_t cd = code[k];
            const int ky = (int)__umulhi((uint32_t)k, magic), kx = k - ky * lw;
            if ((cd & 3) == C_B && !(cd & K_LV) && kx > 0 && kx < lw - 1 && ky > 0 && ky < lh - 1)
                cand = ((code[k - 1] | code[k + 1] | code[k - lw] | code[k + lw]) & K_INRUN) != 0;
            init = (cd & K_INRUN) && (cd & K_LV);
        }
        const unsigned long long bc = __ballot(cand), bi = __ballot(init);
        cnt += __popcll(bc);
        if (!bi) continue;
        int lvl = -1, g = 0;
        if (init) {
            g = gidx(k);
            lvl = D.lvl0[g];
        }
        int prevk = -1, nextk = -1;  // this chunk's previous and next initial item of the same level
        for (unsigned long long m = bi; m; m &= m - 1) {
            const int j = __ffsll((long long)m) - 1;
            const int lj = __builtin_amdgcn_readlane(lvl, j);
            if (lj == lvl) {
                if (j < lane) prevk = k0 + j;
                else if (j > lane && nextk < 0) nextk = k0 + j;
            }
        }
        if (init) {
            const int q = ninit + __popcll(bi & ((1ull << lane) - 1ull));
            bkey[k] = ((uint32_t)lvl << 16) | (uint32_t)q;
            bloc[k] = (uint16_t)k;
            code[k] |= K_ST1;
            qn[k] = nextk < 0 ? P1_NIL : (uint16_t)nextk;
            gst(D.ilvl, g, (uint8_t)lvl);
            gst(D.pseq, g, q);
            gst(D.parent, g, -1);
            gst(D.bnode, g, g);
            gst(D.st, g, 1);
            if (prevk < 0) {
                const uint16_t t = qtail[lvl];
                if (t == P1_NIL) qhead[lvl] = (uint16_t)k;
                else qn[t] = (uint16_t)k;
            }
            if (nextk < 0) qtail[lvl] = (uint16_t)k;
        }
        ninit += __popcll(bi);
    }
    if (lane == 0) sh_base = atomicAdd(D.nleaf, cnt);
    __syncthreads();
    const int base = sh_base;
    // every reserved slot below the capacity reads as unused until a leaf fills it
    for (int j = base + lane; j < min(base + cnt, D.leafcap); j += 64) gst(D.lf_b, j, -1);
    if (base + cnt > D.leafcap) {
        if (lane == 0) atomicOr(D.flags, FB_LEAF_CAP);
        return;
    }
    __syncthreads();
    if (lane != 0) return;

    int active = 0;  // no queued item below this level
    int hx = qhead[0];  // head of the active level's FIFO (a copy of qhead[active])
    // FIFO of level l; the flood resumes at the lowest queued level.  Links and heads carry the
    // queued item's kind (P1_LEAF), so a leaf's pop knows what it pops before reading anything.
    auto enq = [&](int l, int x, uint32_t kind) {
        qn[x] = P1_NIL;
        const uint32_t t = uni(qtail[l]);
        const uint32_t xv = (uint32_t)x | kind;
        if (t != P1_NIL) qn[t] = (uint16_t)xv;
        else qhead[l] = (uint16_t)xv;
        qtail[l] = (uint16_t)x;
        if (t == P1_NIL && l <= active) {  // nothing is queued below `active`: l was empty
            active = l;
            hx = (int)xv;
        }
    };
    int seq = ninit, popc = 0, nlf = 0;
    const int ld[4] = {-1, 1, -lw, lw};
    const int gd[4] = {-1, 1, -W, W};
    for (;;) {
        int xv = hx;
        if (xv == P1_NIL) {
            int l = active + 1;
            while (l < NQ && uni(qhead[l]) == P1_NIL) ++l;
            if (l == NQ) break;
            active = l;
            xv = uni(qhead[l]);
        }
        const int x = xv & (int)P1_NIL;
        if (xv & P1_LEAF) {  // a leaf: record when it pops
            uint32_t nx = qn[x];
            int bl = bloc[x];
            asm volatile("" : "+v"(nx), "+v"(bl));
            nx = uni(nx);
            bl = uni(bl);
            qhead[active] = (uint16_t)nx;
            hx = (int)nx;
            if (nx == P1_NIL) qtail[active] = P1_NIL;
            gst(D.lf_popidx, base + bl, popc++);
            continue;
        }
        // an F item: its link, class and chain entry and its neighbours' values, in one round
        // trip (the compiler would otherwise sink the push-side reads past the dam test and fetch
        // them in a second round trip)
        uint32_t nx = qn[x], cpc = code[x];
        int nl[4];
        uint32_t nc[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            nl[d] = lk[x + ld[d]];
            nc[d] = code[x + ld[d]];
        }
        uint32_t dleft = dif[x - 1], dhere = dif[x], dup = dif[x - lw];
        uint32_t bk = bkey[x];
        int bl = bloc[x];
        asm volatile("" : "+v"(nx), "+v"(cpc), "+v"(nl[0]), "+v"(nl[1]), "+v"(nl[2]), "+v"(nl[3]), "+v"(nc[0]),
                     "+v"(nc[1]), "+v"(nc[2]), "+v"(nc[3]), "+v"(dleft), "+v"(dhere), "+v"(dup), "+v"(bk), "+v"(bl));
        // one lane floods: every value it read is the same for the wave, so the bookkeeping below
        // runs on the scalar unit with scalar branches instead of masked vector code
        nx = uni(nx);
        cpc = uni(cpc);
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            nl[d] = uni(nl[d]);
            nc[d] = uni(nc[d]);
        }
        dleft = uni(dleft);
        dhere = uni(dhere);
        dup = uni(dup);
        bk = uni(bk);
        bl = uni(bl);
        qhead[active] = (uint16_t)nx;
        hx = (int)nx;
        if (nx == P1_NIL) qtail[active] = P1_NIL;
        const int p = x;
        const int lvd[4] = {(int)(dleft & 255), (int)(dhere & 255), (int)(dup >> 8), (int)(dhere >> 8)};
        const int gp = gidx(p);
        gst(D.popidx, gp, popc++);
        const int lab = pop_label(nl[0], nl[1], nl[2], nl[3]);
        lk[p] = lab;
        code[p] = (uint8_t)((cpc & ~K_ST1) | K_ST2);
        gst(D.lab, gp, lab);
        gst(D.st, gp, 2);
        if (lab == WSHED) continue;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int sk = p + ld[d];
            // a neighbour's state may have changed since the batch read only through this pop's own
            // earlier pushes, which go to other pixels
            const uint32_t cs = nc[d];
            if ((cs & 3) == C_F) {
                if (cs & (K_ST1 | K_ST2)) continue;
                const int l = lvd[d];
                const int gs = gp + gd[d];
                const uint32_t key = ((uint32_t)l << 16) | (uint32_t)seq;
                const bool newer = bk < key;  // the new item has the larger key on its chain
                bkey[sk] = newer ? key : bk;
                bloc[sk] = (uint16_t)(newer ? sk : bl);
                code[sk] = cs | K_ST1;
                gst(D.ilvl, gs, (uint8_t)l);
                gst(D.pseq, gs, seq);
                gst(D.parent, gs, gp);
                gst(D.bnode, gs, newer ? gs : gidx(bl));
                gst(D.st, gs, 1);
                ++seq;
                enq(l, sk, 0);
            } else if ((cs & 3) == C_B && !(cs & (K_LV | K_ST1))) {
                // K_ST1 on a B pixel: queued already, by an earlier labelled neighbour of this run
                // (every popped F pixel of the run is labelled or a dam, and a labelled one queues
                // its B neighbours)
                if (nlf >= cnt) { atomicOr(D.flags, FB_LEAF_CAP); continue; }
                const int j = nlf++;
                const int slot = base + j;
                bloc[sk] = (uint16_t)j;
                code[sk] = cs | K_ST1;
                const int l = lvd[d];
                const uint32_t key = ((uint32_t)l << 16) | (uint32_t)seq;
                gst(D.lf_b, slot, gp + gd[d]);
                gst(D.lf_run, slot, id);
                gst(D.lf_lvl, slot, (uint8_t)l);
                gst(D.lf_pseq, slot, seq);
                gst(D.lf_parent, slot, gp);
                gst(D.lf_popidx, slot, -1);
                gst(D.lf_bnode, slot, bk < key ? D.N + slot : gidx(bl));
                ++seq;
                enq(l, sk, P1_LEAF);
            }
        }
        if (seq >= 65535) { atomicOr(D.flags, FB_LEAF_CAP); return; }
    }
    }();
    __syncthreads();  // the LDS arrays are reused by the next run
    }
}

// LDS bytes of one k_phase1_lds<LCAP> workgroup (the host sizes each tier's grid from it)
constexpr int p1_lds_bytes(int lcap) { return 15 * lcap + 4 * NQ + 8; }

// grid of a k_phase1_lds<LCAP> launch: the workgroups the device holds at once (LDS-bound), at most
// one per run of the image
template <int LCAP>
int p1_grid(int nrun) {
    static const int g = [] {
        int dev = 0, per_cu = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_phase1_lds<LCAP>, 64, 0) != hipSuccess) {
            (void)hipGetLastError();
            return 256 * std::max(1, 160 * 1024 / p1_lds_bytes(LCAP));
        }
        return std::max(1, per_cu) * std::max(1, cus);
    }();
    return std::max(1, std::min(nrun, g));
}

// wave-aggregated append of the lanes whose `take` is set
__device__ __forceinline__ void wave_append(bool take, int v, int* __restrict__ list, int* __restrict__ count) {
    const unsigned long long m = __ballot(take);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader);
    if (take) list[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
}

// every used leaf slot joins the list of its B pixel (list order is irrelevant to its readers);
// the first leaf of a pixel also puts the pixel on the list of B pixels with queue entries
// (one append per workgroup and iteration: an atomic per wave on the single list counter
// serialises over tens of thousands of waves)
__global__ void k_link_leaves(Dev D, int* __restrict__ blist, int* __restrict__ bcount) {
    const int n = min(*D.nleaf, D.leafcap);
    const int stride = gridDim.x * blockDim.x;
    for (int k0 = blockIdx.x * blockDim.x; k0 < n; k0 += stride) {  // block-uniform trip count
        const int sl = k0 + threadIdx.x;
        bool first = false;
        int b = -1;
        if (sl < n) {
            b = D.lf_b[sl];
            if (b >= 0 && b < D.N) {
                const int old = atomicExch(&D.head[b], sl);
                D.lf_next[sl] = old;
                first = old < 0;
            }
        }
        block_append(first, b, blist, bcount);
    }
}

// The item records before() walks, from the phase-1 results: F pixels (run, chain's largest-key
// item, pop index, pusher, level), B pixels queued in the initial scan (run -(x + 2), themselves,
// pop index 0, no pusher, initial level) and used leaf slots (ids N + slot).  Other ids are never
// visited by a walk and keep whatever the memory held.
__global__ void k_pack_items(Dev D) {
    const int N = D.N;
    const int nl = min(*D.nleaf, D.leafcap);
    const size_t total = (size_t)N + nl;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        if (t < (size_t)N) {
            const int x = (int)t;
            const uint8_t c = D.cls[x];
            if (c == C_F) {
                D.wr[x] = make_int4(D.run[x], D.bnode[x], D.popidx[x], D.parent[x]);
                D.wl[x] = D.ilvl[x];
            } else if (c == C_B && D.lvl0[x] != NOPUSH) {
                D.wr[x] = make_int4(-(x + 2), x, 0, -1);
                D.wl[x] = (uint8_t)D.lvl0[x];
            }
        } else {
            const int sl = (int)(t - N);
            if (D.lf_b[sl] < 0) continue;
            D.wr[t] = make_int4(D.lf_run[sl], D.lf_bnode[sl], D.lf_popidx[sl], D.lf_parent[sl]);
            D.wl[t] = D.lf_lvl[sl];
        }
    }
}

// B pixels queued in the initial scan (next to a seed)
__global__ void k_collect_b0(Dev D, int* __restrict__ blist, int* __restrict__ bcount) {
    const int stride = gridDim.x * blockDim.x;
    const int n_up = ((D.N + 63) / 64) * 64;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n_up; i += stride) {
        const int j = min(i, D.N - 1);  // both loads unconditional (clamped), then the test
        const uint8_t c = D.cls[j];
        const uint16_t l0 = D.lvl0[j];
        const bool take = i < D.N && c == C_B && l0 != NOPUSH;
        wave_append(take, i, blist, bcount);
    }
}

// runs by the size of their LDS frame (box + 1-pixel margin): bucket b holds frames of at most
// kRunTier[b] pixels; the last bucket floods from global memory
constexpr int kRunTiers = 6;

__constant__ int kRunTier[kRunTiers - 1] = {512, 1024, 2048, 4096, 6400};
// (one append per wave and bucket: an atomic per run on the few bucket counters serialises)
__global__ void k_bucket_runs(const int32_t* __restrict__ box5, int nrun, int* __restrict__ lists, int* __restrict__ counts) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    int b = -1;
    if (r < nrun) {
        const long long L = (long long)(box5[2 * nrun + r] - box5[nrun + r] + 1 + 2 * P1_MARGIN) *
                            (box5[4 * nrun + r] - box5[3 * nrun + r] + 1 + 2 * P1_MARGIN);
        b = 0;
        while (b < kRunTiers - 1 && L > kRunTier[b]) ++b;
    }
    for (int t = 0; t < kRunTiers; ++t) wave_append(b == t, r, lists + (size_t)t * nrun, &counts[t]);
}

// Label a B pixel b (never on the frame) takes when its queue entry r pops: the agreeing positive
// label of its seed neighbours and of its F neighbours popped (labelled) before r, WSHED where
// they disagree, 0 for none.  The four neighbours' class, marker, state and label are loaded
// together and used through selects (loads behind the class test were a dependent round trip
// per neighbour); only the pop-order walk stays conditional.
__device__ int b_label(const Dev& D, int b, int r, int W) {
    const int dirs[4] = {-1, 1, -W, W};
    uint8_t cs[4], ss[4];
    int ms[4], ls[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const int s = b + dirs[d];
        cs[d] = D.cls[s];
        ss[d] = D.st[s];
        ms[d] = D.m0[s];
        ls[d] = D.lab[s];
    }
    int lab = 0;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        const bool seed = cs[d] == C_SEED;
        const bool fcand = cs[d] == C_F && ss[d] == 2 && ls[d] > 0;
        int t = seed ? ms[d] : 0;
        if (!seed && fcand && before(D, b + dirs[d], r)) t = ls[d];
        if (t > 0) { if (lab == 0) lab = t; else if (t != lab) lab = WSHED; }
    }
    return lab;
}

// ---- B items: the real queue entry of each B pixel and its label decision at pop time ----
// over the B pixels that have a queue entry (listed by k_collect_b0 / k_link_leaves); real[] of
// every other pixel stays -1
__global__ void k_bresolve(Dev D, int32_t* __restrict__ real, int* __restrict__ cand, int* __restrict__ ncand,
                           const int* __restrict__ blist, const int* __restrict__ bcount) {
    const int W = D.W;
    const int nb = *bcount;
    // block-uniform trip count (block_append synchronises the block)
    for (int k0 = blockIdx.x * blockDim.x; k0 < nb; k0 += gridDim.x * blockDim.x) {
        const int k = k0 + threadIdx.x;
        int b = -1, lab = 0;
        if (k < nb) {
            b = blist[k];
            int r = -1;
            if (D.lvl0[b] != NOPUSH) r = b;
            else {
                for (int sl = D.head[b]; sl >= 0; sl = D.lf_next[sl]) {
                    const int lid = D.N + sl;
                    if (r < 0 || before(D, D.lf_parent[sl], it_parent(D, r))) r = lid;
                }
            }
            real[b] = r;
            if (r >= 0) lab = b_label(D, b, r, W);
        }
        block_append(lab > 0, b, cand, ncand);
    }
}

// earliest candidate per block (candidates: B pixels whose earlier-popped neighbours agree on a
// positive label), -1 for a block without candidates
constexpr int B1_BLOCKS = 1024;
__global__ __launch_bounds__(256) void k_b1_partial(Dev D, const int32_t* __restrict__ real, const int* __restrict__ cand,
                                                    const int* __restrict__ ncand, int* __restrict__ partial) {
    __shared__ int best[256];
    __shared__ ItemKey bkey[256];
    const int n = *ncand;
    int mine = -1;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += B1_BLOCKS * 256) {
        const int b = cand[k];
        if (mine < 0 || before(D, real[b], real[mine])) mine = b;
    }
    best[threadIdx.x] = mine;
    if (mine >= 0) bkey[threadIdx.x] = item_key(D, real[mine]);
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if ((int)threadIdx.x < stride) {
            const int a = best[threadIdx.x], c = best[threadIdx.x + stride];
            if (a < 0 || (c >= 0 && before_keyed(D, bkey[threadIdx.x + stride], bkey[threadIdx.x]))) {
                best[threadIdx.x] = c;
                bkey[threadIdx.x] = bkey[threadIdx.x + stride];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = best[0];
}

// info[0] = b1 pixel (-1 none), info[1] = b1 item, info[2] = Y1; `cand` holds the B1_BLOCKS
// per-block winners (-1 = none)
__global__ __launch_bounds__(1024) void k_b1_reduce(Dev D, const int32_t* __restrict__ real, const int* __restrict__ cand,
                                                    int* __restrict__ info) {
    __shared__ int best[1024];
    __shared__ ItemKey bkey[1024];
    const int n = B1_BLOCKS;
    int mine = -1;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int b = cand[k];
        if (b < 0) continue;
        if (mine < 0 || before(D, real[b], real[mine])) mine = b;
    }
    best[threadIdx.x] = mine;
    if (mine >= 0) bkey[threadIdx.x] = item_key(D, real[mine]);
    __syncthreads();
    for (int stride = 512; stride > 0; stride >>= 1) {
        if ((int)threadIdx.x < stride) {
            const int a = best[threadIdx.x], c = best[threadIdx.x + stride];
            if (a < 0 || (c >= 0 && before_keyed(D, bkey[threadIdx.x + stride], bkey[threadIdx.x]))) {
                best[threadIdx.x] = c;
                bkey[threadIdx.x] = bkey[threadIdx.x + stride];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int b1 = best[0];
        info[0] = b1;
        info[1] = b1 >= 0 ? real[b1] : -1;
        int lab = 0;
        if (b1 >= 0) lab = b_label(D, b1, real[b1], D.W);
        info[2] = lab;
        // a level-0 b1 could share the level-0 FIFO with other items: outside the decomposition
        if (b1 >= 0 && it_lvl(D, real[b1]) == 0) atomicOr(D.flags, FB_TIES);
    }
}

// queue state at T1 (after b1's pop); P pixels (B, never queued) -> pmask
__global__ void k_t1_state(Dev D, const int32_t* __restrict__ real, const int* __restrict__ info, uint8_t* __restrict__ pmask,
                           uint8_t* __restrict__ kind) {
    const int b1 = info[0], b1i = info[1], Y1 = info[2];
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < D.N; p += gridDim.x * blockDim.x) {
        uint8_t pm = 0, kd = 0;
        // the pixel's class, state, label, pusher and queue entry are loaded together (each one
        // behind the class test would be another dependent round trip)
        const uint8_t c = D.cls[p], stp = D.st[p];
        const int labp = D.lab[p], parp = D.parent[p], r = real[p];
        if (c == C_F) {
            int v = 0;
            if (stp == 2 && (b1 < 0 || before(D, p, b1i))) v = labp;
            else if (stp >= 1 && (parp < 0 || b1 < 0 || before(D, parp, b1i))) { v = IN_QUEUE; kd = 1; }
            D.m[p] = v;
        } else if (c == C_B) {
            int v = 0;
            if (p == b1) v = Y1;
            else if (r >= 0) {
                const int par = D.wr[r].w;  // the queue entry's pusher, from its packed record
                const bool queued = par < 0 || b1 < 0 || before(D, par, b1i);
                if (queued) {
                    if (b1 < 0 || before(D, r, b1i)) v = WSHED;
                    else { v = IN_QUEUE; kd = 1; }
                }
            }
            if (v == 0) pm = 1;
            D.m[p] = v;
        }
        pmask[p] = pm;
        kind[p] = kd;
    }
}

// flood from b1: the never-queued B pixels it reaches are exactly those at finite geodesic
// distance from b1 through never-queued B pixels; they take Y1.  An F pixel still unqueued next to
// one of them (or to b1) is queued by that flood (kind 2) from its neighbour nearest to b1: the
// FIFO position of that push is ordered first by this distance (the flood's layer).
__global__ void k_geo_init(Dev D, const int* __restrict__ info, uint8_t* __restrict__ pmask, unsigned* __restrict__ dist) {
    const int b1 = info[0];
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < D.N; p += gridDim.x * blockDim.x) {
        const bool src = (p == b1);
        if (src) pmask[p] = 1;
        dist[p] = src ? 0u : 0xFFFFFFFFu;
    }
}
__global__ void k_bfs_mark(Dev D, const int* __restrict__ info, const unsigned* __restrict__ dist, uint8_t* __restrict__ kind,
                           int32_t* __restrict__ pdist) {
    const int b1 = info[0], Y1 = info[2];
    const int W = D.W;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < D.N; p += gridDim.x * blockDim.x) {
        if (b1 < 0 || Y1 <= 0) continue;
        if (D.cls[p] != C_F || D.m[p] != 0) continue;
        const int dirs[4] = {-1, 1, -W, W};
        unsigned best = 0xFFFFFFFFu;
        for (int d = 0; d < 4; ++d) {
            const int s = p + dirs[d];
            if (D.cls[s] == C_B) best = min(best, dist[s]);
        }
        if (best != 0xFFFFFFFFu) { kind[p] = 2; pdist[p] = (int)best; }
    }
}
__global__ void k_bfs_commit(Dev D, const int* __restrict__ info, const unsigned* __restrict__ dist, const uint8_t* __restrict__ kind,
                             uint8_t* __restrict__ umask) {
    const int b1 = info[0], Y1 = info[2];
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < D.N; p += gridDim.x * blockDim.x) {
        // the four loads first (under the conditions they would wait for one another)
        const int m0 = D.m[p];
        const uint8_t c = D.cls[p], kd = kind[p];
        const unsigned dp = dist[p];
        int v = m0;
        if (b1 >= 0 && Y1 > 0 && c == C_B && v == 0 && dp != 0xFFFFFFFFu) v = Y1;
        if (kd == 2) v = IN_QUEUE;
        if (v != m0) D.m[p] = v;
        umask[p] = (c != C_FRAME && c != C_SEED && (v == 0 || v == IN_QUEUE)) ? 1 : 0;
    }
}

// ---- phase 2: one lane floods one component of unresolved pixels ----
struct P2 {
    const int32_t* comp;
    const int32_t* box5;
    const int32_t* area;
    const int32_t* off;
    const uint8_t* kind;
    int32_t* scratch;  // 3 * area per component: item list, heap, saved labels
    int ncomp;
};

// Order of two queued items of the same level by push time: phase-1 items by pusher order, items
// queued by the flood from b1 after them (their mutual order is not known: they are ordered by the
// b1 flood's layer, then in raster order, and the caller checks that the order does not matter).
// The data this reads of one item is packed once, so that a sort compares cached values:
// bit 63 = queued by the b1 flood (kind 2); then, for kind 2, its label (the layer) in the low 32
// bits; else (pusher + 1) in bits 32..62 and the push sequence in the low 32 bits.
__device__ long long push_key(const Dev& D, const int32_t* real, const uint8_t* kind, int a) {
    if (kind[a] == 2) return (long long)(1ull << 63) | (long long)(uint32_t)D.lab[a];
    const int ia = D.cls[a] == C_B ? real[a] : a;
    const int pa = it_parent(D, ia);
    return ((long long)(uint32_t)(pa + 1) << 32) | (long long)(uint32_t)it_pseq(D, ia);
}
// a pushed before b, from their packed data
__device__ bool push_before_key(const Dev& D, const int32_t* real, int a, long long ka, int b, long long kb) {
    const bool ba = ka < 0, bb = kb < 0;
    if (ba != bb) return bb;
    if (ba) {
        const int la = (int)(uint32_t)ka, lb = (int)(uint32_t)kb;
        if (la != lb) return la < lb;
        return a < b;
    }
    const int pa = (int)((ka >> 32) & 0x7FFFFFFF) - 1, pb = (int)((kb >> 32) & 0x7FFFFFFF) - 1;
    if (pa < 0 && pb < 0) {
        const int ia = D.cls[a] == C_B ? real[a] : a;
        const int ib = D.cls[b] == C_B ? real[b] : b;
        return ia < ib;
    }
    if (pa < 0) return true;
    if (pb < 0) return false;
    if (pa == pb) return (int)(uint32_t)ka < (int)(uint32_t)kb;
    return before(D, pa, pb);
}

__device__ int item_level(const Dev& D, const int32_t* real, const uint8_t* kind, int p) {
    if (kind[p] == 2) {
        const uint8_t* c = D.img + 3 * (size_t)p;
        return max(max((int)c[0], (int)c[1]), (int)c[2]);
    }
    return D.cls[p] == C_B ? it_lvl(D, real[p]) : D.ilvl[p];
}

// binary heap on (level << 24 | seq) packed keys, values = pixel ids
__device__ __forceinline__ void heap_push(long long* hk, int* hv, int& n, long long k, int v) {
    int i = n++;
    while (i > 0) {
        const int pi = (i - 1) >> 1;
        if (hk[pi] <= k) break;
        hk[i] = hk[pi]; hv[i] = hv[pi]; i = pi;
    }
    hk[i] = k; hv[i] = v;
}
__device__ __forceinline__ int heap_pop(long long* hk, int* hv, int& n) {
    const int top = hv[0];
    const long long k = hk[--n];
    const int v = hv[n];
    int i = 0;
    while (true) {
        int c = 2 * i + 1;
        if (c >= n) break;
        if (c + 1 < n && hk[c + 1] < hk[c]) ++c;
        if (hk[c] >= k) break;
        hk[i] = hk[c]; hv[i] = hv[c]; i = c;
    }
    if (n > 0) { hk[i] = k; hv[i] = v; }
    return top;
}

constexpr int MAXG = 8, MAXGL = 5, MAXCOMBO = 4096;
struct TieRec {
    int c, ni, ntie, ncombo;
    long long base;  // offset (ints) of the record's region in the tie pool
    int gfirst[MAXG], glen[MAXG];
};

__device__ __forceinline__ int fact(int n) { int f = 1; for (int k = 2; k <= n; ++k) f *= k; return f; }

// flood one component from its sorted queue `order` (ni items); the state of the component's
// pixels lives either in D.m (st == nullptr) or in a private bbox-indexed array st.  Returns
// false on heap overflow.
__device__ bool flood_component(const Dev& D, const P2& Q, int c, const int* order, int ni, int* st, int* hv,
                                long long* hk, int A, const int32_t* real) {
    const int W = D.W, id = c + 1;
    const int minx = Q.box5[Q.ncomp + c], maxx = Q.box5[2 * Q.ncomp + c];
    const int miny = Q.box5[3 * Q.ncomp + c];
    const int bw = maxx - minx + 1;
    auto loc = [&](int i) -> int* {  // private cell of a component pixel, nullptr otherwise
        if (!st) return nullptr;
        const int x = i % W, y = i / W;
        if (x < minx || x > maxx || y < miny || Q.comp[i] != id) return nullptr;
        return st + (y - miny) * bw + (x - minx);
    };
    auto get = [&](int i) { int* q = loc(i); return q ? *q : D.m[i]; };
    auto put = [&](int i, int v) { int* q = loc(i); if (q) *q = v; else D.m[i] = v; };
    const int dirs[4] = {-1, 1, -W, W};
    int hn = 0;
    long long seqc = 0;
    for (int k = 0; k < ni; ++k)
        heap_push(hk, hv, hn, ((long long)item_level(D, real, Q.kind, order[k]) << 40) | (seqc++), order[k]);
    while (hn > 0) {
        const int p = heap_pop(hk, hv, hn);
        int lab = 0;
        for (int d = 0; d < 4; ++d) {
            const int t = get(p + dirs[d]);
            if (t > 0) { if (lab == 0) lab = t; else if (t != lab) lab = WSHED; }
        }
        put(p, lab);
        if (lab == WSHED) continue;
        const uint8_t* ptr = D.img + 3 * (size_t)p;
        for (int d = 0; d < 4; ++d) {
            const int s2 = p + dirs[d];
            if (get(s2) == 0) {
                if (hn >= A) return false;
                const int l = c_diff(ptr, D.img + 3 * (size_t)s2);
                heap_push(hk, hv, hn, ((long long)l << 40) | (seqc++), s2);
                put(s2, IN_QUEUE);
            }
        }
    }
    return true;
}

// components deferred to the LDS flood: a record per component, a work item per (record, arrangement)
constexpr int FL_SMALL = 1024, FL_BIG = 4096;  // frame pixels (box + 1-pixel margin) of the two LDS tiers
constexpr int FL_MIN_AREA = 24;                // smaller tie-free components flood in their lane
struct FloodQ {
    TieRec* recs;
    int* nrec;
    int reccap;
    int* work[2];
    int* nwork;  // [2]
    int workcap;
    int* res;    // per-arrangement results of tie components (pool)
    int* recbad;  // per record: some arrangement differs from arrangement 0
};

// boxes larger than this are scanned by the whole wave (64 pixels a step) instead of one lane, and
// item lists longer than P2_COOP_SORT (up to 64 * P2_SORT_PER) are sorted by the whole wave
constexpr int P2_COOP_BOX = 512, P2_COOP_SORT = 16, P2_SORT_PER = 8;

// One lane per component; launched with 64-thread blocks (one wave), which the cooperative scan of
// large boxes relies on.
__global__ void k_phase2(Dev D, P2 Q, const int32_t* __restrict__ real, long long* __restrict__ heapk, TieRec* ties,
                         int* nties, int tiecap, long long* pool_used, long long poolcap, FloodQ F) {
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = c < Q.ncomp;
    const int W = D.W;
    const int id = c + 1;
    const int cc = valid ? c : 0;  // a valid index for the loads of lanes without a component
    const int A = Q.area[cc];
    int32_t* items = Q.scratch + 3 * (size_t)Q.off[cc];
    int* hv = items + A;
    long long* hk = heapk + (size_t)Q.off[cc];
    const int minx = Q.box5[Q.ncomp + cc], maxx = Q.box5[2 * Q.ncomp + cc];
    const int miny = Q.box5[3 * Q.ncomp + cc], maxy = Q.box5[4 * Q.ncomp + cc];
    const int bw = maxx - minx + 1, B = bw * (maxy - miny + 1);
    // collect queued items (raster order).  Boxes above P2_COOP_BOX pixels: the wave scans each
    // such box in turn, 64 box pixels a step, and compacts the hits with a ballot (a lane alone
    // would wait on one memory round trip per eight pixels of a box of up to thousands).
    int ni = 0;
    const bool coop = valid && B > P2_COOP_BOX;
    unsigned long long pend = __ballot(coop);
    while (pend) {
        const int src = __ffsll((long long)pend) - 1;
        pend &= pend - 1;
        const int sc = __shfl(c, src), sminx = __shfl(minx, src), sminy = __shfl(miny, src);
        const int sbw = __shfl(bw, src), sB = __shfl(B, src);
        int32_t* sitems = Q.scratch + 3 * (size_t)Q.off[sc];
        int cnt = 0;
        for (int k0 = 0; k0 < sB; k0 += 64) {
            const int k = k0 + lane;
            bool take = false;
            int gi = 0;
            if (k < sB) {
                gi = (sminy + k / sbw) * W + sminx + k % sbw;
                take = Q.comp[gi] == sc + 1 && D.m[gi] == IN_QUEUE;
            }
            const unsigned long long tb = __ballot(take);
            if (take) sitems[cnt + __popcll(tb & ((1ull << lane) - 1ull))] = gi;
            cnt += __popcll(tb);
        }
        if (lane == src) ni = cnt;
    }
    // smaller boxes: the lane scans its own, eight pixels' loads in flight at a time (a load per
    // pixel with its test in the same iteration waits on every pixel in turn)
    if (valid && !coop)
        for (int y = miny; y <= maxy; ++y)
            for (int x0 = minx; x0 <= maxx; x0 += 8) {
                int cv[8], mv[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int i = y * W + min(x0 + j, maxx);
                    cv[j] = Q.comp[i];
                    mv[j] = D.m[i];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (x0 + j <= maxx && cv[j] == id && mv[j] == IN_QUEUE) items[ni++] = y * W + x0 + j;
            }
    // Sort by (level, push time): the FIFO pop order, a strict total order.  Each item's level and
    // push data are computed once (into hv[] and hk[], free until the flood below, and ni <= A)
    // and move with the item.  Few items: the lane's own insertion sort.  More (the components of
    // the LDS floods, up to ~100 items): the wave ranks each item by counting the items that
    // precede it (one item per lane, all compared at once) and scatters the items to their ranks
    // -- the same order, without one lane's ~n^2/2 dependent comparisons.
    const bool coopsort = valid && ni > P2_COOP_SORT && ni <= 64 * P2_SORT_PER;
    if (valid && !coopsort) {
        for (int a = 0; a < ni; ++a) {
            const int v = items[a];
            hv[a] = item_level(D, real, Q.kind, v);
            hk[a] = push_key(D, real, Q.kind, v);
        }
        for (int a = 1; a < ni; ++a) {
            const int v = items[a], lv = hv[a];
            const long long kv = hk[a];
            int b = a - 1;
            while (b >= 0) {
                const int u = items[b], lu = hv[b];
                const long long ku = hk[b];
                const bool gt = (lu > lv) || (lu == lv && push_before_key(D, real, v, kv, u, ku));
                if (!gt) break;
                items[b + 1] = u;
                hv[b + 1] = lu;
                hk[b + 1] = ku;
                --b;
            }
            items[b + 1] = v;
            hv[b + 1] = lv;
            hk[b + 1] = kv;
        }
    }
    __syncthreads();  // every item list (stored by every lane in the scans) is complete
    unsigned long long pend_sort = __ballot(coopsort);
    while (pend_sort) {
        const int src = __ffsll((long long)pend_sort) - 1;
        pend_sort &= pend_sort - 1;
        const int sc = __shfl(c, src), sni = __shfl(ni, src), sA = __shfl(A, src);
        int32_t* sitems = Q.scratch + 3 * (size_t)Q.off[sc];
        int* shv = sitems + sA;
        long long* shk = heapk + (size_t)Q.off[sc];
        for (int k = lane; k < sni; k += 64) {
            const int v = sitems[k];
            shv[k] = item_level(D, real, Q.kind, v);
            shk[k] = push_key(D, real, Q.kind, v);
        }
        __syncthreads();
        int rk[P2_SORT_PER], vv[P2_SORT_PER], lvv[P2_SORT_PER];
        long long kvv[P2_SORT_PER];
#pragma unroll
        for (int q = 0; q < P2_SORT_PER; ++q) {
            const int k = lane + 64 * q;
            rk[q] = -1;
            if (k >= sni) continue;
            const int v = sitems[k], lv = shv[k];
            const long long kv = shk[k];
            int rank = 0;
            for (int j = 0; j < sni; ++j) {
                if (j == k) continue;
                const int u = sitems[j], lu = shv[j];
                rank += (lu < lv) || (lu == lv && push_before_key(D, real, u, shk[j], v, kv)) ? 1 : 0;
            }
            rk[q] = rank;
            vv[q] = v;
            lvv[q] = lv;
            kvv[q] = kv;
        }
        __syncthreads();  // every comparison has read the lists before they are overwritten
#pragma unroll
        for (int q = 0; q < P2_SORT_PER; ++q) {
            if (rk[q] < 0) continue;
            sitems[rk[q]] = vv[q];
            shv[rk[q]] = lvv[q];
            shk[rk[q]] = kvv[q];
        }
        __syncthreads();  // the sorted list is complete before its owner reads it
    }
    if (!valid) return;
    // groups of items queued by the b1 flood at the same level and layer: order unknown
    TieRec R;
    R.c = c; R.ni = ni; R.ntie = 0;
    long long ncombo = 1;
    for (int a = 0; a < ni;) {
        const int v = items[a];
        if (Q.kind[v] != 2) { ++a; continue; }
        const int lv = hv[a];  // cached levels, aligned with the sorted items
        int e = a + 1;
        while (e < ni && Q.kind[items[e]] == 2 && hv[e] == lv && D.lab[items[e]] == D.lab[v]) ++e;
        if (e - a >= 2) {
            if (R.ntie == MAXG || e - a > MAXGL) { atomicOr(D.flags, FB_TIES); return; }
            R.gfirst[R.ntie] = a;
            R.glen[R.ntie] = e - a;
            ++R.ntie;
            ncombo *= fact(e - a);
            if (ncombo > MAXCOMBO) { atomicOr(D.flags, FB_TIES); return; }
        }
        a = e;
    }
    R.ncombo = (int)ncombo;
    const int L = (maxx - minx + 3) * (maxy - miny + 3);
    if (L <= FL_BIG && (R.ntie > 0 || A > FL_MIN_AREA)) {  // LDS flood, one wave per arrangement
        long long base = 0;
        if (ncombo > 1) {
            const long long need = ncombo * (long long)L;
            base = atomicAdd((unsigned long long*)pool_used, (unsigned long long)need);
            if (base + need > poolcap) { atomicOr(D.flags, FB_TIES); return; }
        }
        R.base = base;
        const int rec = atomicAdd(F.nrec, 1);
        if (rec >= F.reccap) { atomicOr(D.flags, FB_TIES); return; }
        F.recs[rec] = R;
        const int tier = L <= FL_SMALL ? 0 : 1;
        const int w0 = atomicAdd(&F.nwork[tier], (int)ncombo);
        if (w0 + ncombo > F.workcap) { atomicOr(D.flags, FB_TIES); return; }
        for (int k = 0; k < (int)ncombo; ++k) F.work[tier][w0 + k] = rec * MAXCOMBO + k;
        return;
    }
    if (R.ntie == 0) {
        if (!flood_component(D, Q, c, items, ni, nullptr, hv, hk, A, real)) atomicOr(D.flags, FB_HEAP);
        return;
    }
    // defer to the parallel permutation check (one wave per component)
    const long long per = (long long)B + ni + A + 2LL * A + 4;  // B: box pixels
    const long long need = 64 * per + B + 64;
    const long long base = atomicAdd((unsigned long long*)pool_used, (unsigned long long)need);
    const int k = atomicAdd(nties, 1);
    if (k >= tiecap) { atomicOr(D.flags, FB_TIES); return; }
    if (base + need > poolcap) {  // slot taken but no room: leave an empty record
        atomicOr(D.flags, FB_TIES);
        R.c = -1;
    }
    R.base = base;
    ties[k] = R;
}

// one wave per deferred component: lane t floods the arrangements t, t+64, ... of the tie groups
// on a private copy; every result must equal arrangement 0, which is then committed
__global__ __launch_bounds__(64) void k_phase2_ties(Dev D, P2 Q, const int32_t* __restrict__ real, const TieRec* __restrict__ ties,
                                                   const int* __restrict__ nties, int* pool) {
    const int k = blockIdx.x;
    if (k >= min(*nties, (int)gridDim.x)) return;
    const TieRec R = ties[k];
    if (R.c < 0) return;
    const int c = R.c, id = c + 1, W = D.W, lane = threadIdx.x;
    const int A = Q.area[c];
    const int minx = Q.box5[Q.ncomp + c], maxx = Q.box5[2 * Q.ncomp + c];
    const int miny = Q.box5[3 * Q.ncomp + c], maxy = Q.box5[4 * Q.ncomp + c];
    const int bw = maxx - minx + 1, bh = maxy - miny + 1, B = bw * bh;
    const int ni = R.ni;
    const int32_t* items = Q.scratch + 3 * (size_t)Q.off[c];
    const long long per = (long long)B + ni + A + 2LL * A + 4;
    int* mine = pool + R.base + lane * per;
    int* st = mine;
    int* order = st + B;
    int* hv = order + ni;
    long long* hk = reinterpret_cast<long long*>(hv + A + ((reinterpret_cast<uintptr_t>(hv + A) & 7) ? 1 : 0));
    int* canon = pool + R.base + 64 * per;
    __shared__ int bad;
    if (lane == 0) bad = 0;
    __syncthreads();
    const int rounds = (R.ncombo + 63) / 64;
    for (int r = 0; r < rounds; ++r) {
        const int combo = r * 64 + lane;
        const bool act = combo < R.ncombo;
        if (act) {
            for (int q = 0; q < ni; ++q) order[q] = items[q];
            int code = combo;
            for (int g = 0; g < R.ntie; ++g) {  // arrangement = factorial-base digits per group
                const int L = R.glen[g];
                const int f = fact(L);
                int digit = code % f;
                code /= f;
                int pool_el[MAXGL];
                for (int q = 0; q < L; ++q) pool_el[q] = items[R.gfirst[g] + q];
                for (int q = 0; q < L; ++q) {
                    const int fq = fact(L - 1 - q);
                    const int idx = digit / fq;
                    digit %= fq;
                    order[R.gfirst[g] + q] = pool_el[idx];
                    for (int z = idx; z < L - 1 - q; ++z) pool_el[z] = pool_el[z + 1];
                }
            }
            for (int y = 0; y < bh; ++y)
                for (int x = 0; x < bw; ++x) st[y * bw + x] = 0;
            for (int q = 0; q < ni; ++q) {
                const int i = items[q];
                st[(i / W - miny) * bw + (i % W - minx)] = IN_QUEUE;
            }
            if (!flood_component(D, Q, c, order, ni, st, hv, hk, A, real)) atomicOr(&bad, 1);
        }
        if (r == 0) {
            if (lane == 0)
                for (int q = 0; q < B; ++q) canon[q] = st[q];
            __syncthreads();
        }
        if (act) {
            for (int y = 0; y < bh; ++y)
                for (int x = 0; x < bw; ++x) {
                    const int i = (miny + y) * W + (minx + x);
                    if (Q.comp[i] == id && st[y * bw + x] != canon[y * bw + x]) atomicOr(&bad, 1);
                }
        }
    }
    __syncthreads();
    if (lane == 0) {
        if (bad) { atomicOr(D.flags, FB_TIEPERM); return; }
        for (int y = 0; y < bh; ++y)
            for (int x = 0; x < bw; ++x) {
                const int i = (miny + y) * W + (minx + x);
                if (Q.comp[i] == id) D.m[i] = canon[y * bw + x];
            }
    }
}

// one wave floods one arrangement of one deferred component with its frame (box + 1-pixel
// margin) in LDS: the same flood as flood_component (lowest level first, FIFO within a level,
// neighbours in the order left, right, up, down), lane 0 driving it
template <int LCAP>
__global__ __launch_bounds__(64) void k_flood_lds(Dev D, P2 Q, const int32_t* __restrict__ real, FloodQ F, int tier) {
    __shared__ int32_t st[LCAP];
    __shared__ uint32_t col[LCAP];
    __shared__ uint16_t qn[LCAP];
    __shared__ uint8_t inc[LCAP];
    __shared__ uint16_t ord[LCAP];
    __shared__ uint8_t olv[LCAP];
    __shared__ uint16_t qhead[NQ], qtail[NQ];
    __shared__ int bad;
    const int nw = min(F.nwork[tier], F.workcap);
    for (int wi = blockIdx.x; wi < nw; wi += gridDim.x) {
    const int w = F.work[tier][wi];
    const TieRec R = F.recs[w / MAXCOMBO];
    const int combo = w % MAXCOMBO;
    const int c = R.c, id = c + 1, W = D.W, lane = threadIdx.x;
    const int ox = Q.box5[Q.ncomp + c] - 1, oy = Q.box5[3 * Q.ncomp + c] - 1;
    const int lw = Q.box5[2 * Q.ncomp + c] - ox + 2, lh = Q.box5[4 * Q.ncomp + c] - oy + 2;
    const int L = lw * lh;  // <= LCAP; the frame lies inside the image (components avoid the frame)
    const int ni = R.ni;
    const int32_t* items = Q.scratch + 3 * (size_t)Q.off[c];
    auto gidx = [&](int k) { return (oy + k / lw) * W + ox + k % lw; };
    if (lane == 0) bad = 0;
    for (int k = lane; k < L; k += 64) {
        const int g = gidx(k);
        inc[k] = Q.comp[g] == id ? 1 : 0;
        st[k] = D.m[g];
        const uint8_t* q = D.img + 3 * (size_t)g;
        col[k] = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16);
    }
    for (int l = lane; l < NQ; l += 64) { qhead[l] = NIL16; qtail[l] = NIL16; }
    for (int q = lane; q < ni; q += 64) {
        const int v = items[q];
        ord[q] = (uint16_t)((v / W - oy) * lw + (v % W - ox));
        olv[q] = (uint8_t)item_level(D, real, Q.kind, v);
    }
    __syncthreads();
    if (lane == 0) {
        // arrangement `combo`: factorial-base digits, one per tie group (members share their level)
        int code = combo;
        for (int g = 0; g < R.ntie; ++g) {
            const int Lg = R.glen[g], f0 = R.gfirst[g];
            const int f = fact(Lg);
            int digit = code % f;
            code /= f;
            uint16_t pool_el[MAXGL];
            for (int q = 0; q < Lg; ++q) pool_el[q] = ord[f0 + q];
            for (int q = 0; q < Lg; ++q) {
                const int fq = fact(Lg - 1 - q);
                const int idx = digit / fq;
                digit %= fq;
                ord[f0 + q] = pool_el[idx];
                for (int z = idx; z < Lg - 1 - q; ++z) pool_el[z] = pool_el[z + 1];
            }
        }
        int active = NQ;
        int hx = NIL16;  // head of the active level's FIFO (a copy of qhead[active])
        auto enq = [&](int l, int x) {
            qn[x] = NIL16;
            const uint32_t t = uni(qtail[l]);
            if (t != NIL16) qn[t] = (uint16_t)x;
            else qhead[l] = (uint16_t)x;
            qtail[l] = (uint16_t)x;
            if (t == NIL16 && l <= active) {  // nothing is queued below `active`: l was empty
                active = l;
                hx = x;
            }
        };
        for (int q = 0; q < ni; ++q) enq(olv[q], ord[q]);
        const int ld[4] = {-1, 1, -lw, lw};
        for (;;) {
            int p = hx;
            if (p == NIL16) {
                int l = active + 1;
                while (l < NQ && uni(qhead[l]) == NIL16) ++l;
                if (l >= NQ) break;
                active = l;
                p = uni(qhead[l]);
            }
            // everything the pop reads, in one round trip (this pop's own writes go to p and to
            // distinct neighbours, so the values stay current for its pushes)
            uint32_t nx = qn[p], cp = col[p];
            int ns[4];
            uint32_t ninc[4], ncol[4];
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                ns[d] = st[p + ld[d]];
                ninc[d] = inc[p + ld[d]];
                ncol[d] = col[p + ld[d]];
            }
            asm volatile("" : "+v"(nx), "+v"(cp), "+v"(ns[0]), "+v"(ns[1]), "+v"(ns[2]), "+v"(ns[3]), "+v"(ninc[0]),
                         "+v"(ninc[1]), "+v"(ninc[2]), "+v"(ninc[3]), "+v"(ncol[0]), "+v"(ncol[1]), "+v"(ncol[2]),
                         "+v"(ncol[3]));
            // one lane floods: the values are the wave's, kept in scalar registers
            nx = uni(nx);
            cp = uni(cp);
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                ns[d] = uni(ns[d]);
                ninc[d] = uni(ninc[d]);
                ncol[d] = uni(ncol[d]);
            }
            qhead[active] = (uint16_t)nx;
            hx = (int)nx;
            if (nx == NIL16) qtail[active] = NIL16;
            const int lab = pop_label(ns[0], ns[1], ns[2], ns[3]);
            st[p] = lab;
            if (lab == WSHED) continue;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const int s2 = p + ld[d];
                if (ns[d] != 0) continue;
                if (!ninc[d]) { bad = 1; continue; }  // an unresolved pixel outside the component
                st[s2] = IN_QUEUE;
                enq(cdiff_packed(cp, ncol[d]), s2);
            }
        }
        if (bad) atomicOr(D.flags, FB_HEAP);
    }
    __syncthreads();
    if (R.ncombo == 1) {
        for (int k = lane; k < L; k += 64)
            if (inc[k]) D.m[gidx(k)] = st[k];
    } else {
        int* out = F.res + R.base + (long long)combo * L;
        for (int k = lane; k < L; k += 64) out[k] = st[k];
    }
    __syncthreads();  // the LDS frame is reused by the next work item
    }
}

// Tie components flooded in LDS: every arrangement must give arrangement 0's labels, which are then
// committed.  The comparison takes one workgroup per (record, arrangement) work item, across the
// device (a workgroup per record would read every arrangement of a record through one CU).
__global__ __launch_bounds__(256) void k_tie_check(P2 Q, FloodQ F, int W) {
    const int n0 = min(F.nwork[0], F.workcap), n1 = min(F.nwork[1], F.workcap);
    for (int t = blockIdx.x; t < n0 + n1; t += gridDim.x) {
        const int w = t < n0 ? F.work[0][t] : F.work[1][t - n0];
        const int ri = w / MAXCOMBO, combo = w % MAXCOMBO;
        const TieRec R = F.recs[ri];
        if (R.ncombo == 1 || combo == 0) continue;
        const int c = R.c, id = c + 1;
        const int ox = Q.box5[Q.ncomp + c] - 1, oy = Q.box5[3 * Q.ncomp + c] - 1;
        const int lw = Q.box5[2 * Q.ncomp + c] - ox + 2, lh = Q.box5[4 * Q.ncomp + c] - oy + 2;
        const int L = lw * lh;
        const int* res = F.res + R.base;
        bool diff = false;
        for (int k = threadIdx.x; k < L; k += blockDim.x) {
            const int g = (oy + k / lw) * W + ox + k % lw;
            if (Q.comp[g] == id && res[(long long)combo * L + k] != res[k]) diff = true;
        }
        if (diff) atomicOr(&F.recbad[ri], 1);
    }
}

__global__ __launch_bounds__(256) void k_tie_commit(Dev D, P2 Q, FloodQ F) {
    const int nr = min(*F.nrec, F.reccap);
    for (int ri = blockIdx.x; ri < nr; ri += gridDim.x) {
        const TieRec R = F.recs[ri];
        if (R.ncombo == 1) continue;
        if (F.recbad[ri]) {
            if (threadIdx.x == 0) atomicOr(D.flags, FB_TIEPERM);
            continue;
        }
        const int c = R.c, id = c + 1, W = D.W;
        const int ox = Q.box5[Q.ncomp + c] - 1, oy = Q.box5[3 * Q.ncomp + c] - 1;
        const int lw = Q.box5[2 * Q.ncomp + c] - ox + 2, lh = Q.box5[4 * Q.ncomp + c] - oy + 2;
        const int L = lw * lh;
        const int* res = F.res + R.base;
        for (int k = threadIdx.x; k < L; k += blockDim.x) {
            const int g = (oy + k / lw) * W + ox + k % lw;
            if (Q.comp[g] == id) D.m[g] = res[k];
        }
    }
}

inline dim3 grid1(size_t n) { return dim3(std::min<unsigned>(cdiv(n, 256), 8192u)); }

}  // namespace

// Returns NSC_OK with *ok = 1 when the decomposition produced the exact result in m, *ok = 0
// when the caller must run the sequential flood instead (m is then undefined).
void ws_last_flags_record(int flags);
void ws_tier_record(const int* runs, int ntiers);

int geodist_impl(const uint8_t* pass, unsigned* dist, int w, int h, hipStream_t s, const int* seed);

int cvws_fast(const uint8_t* img, const int32_t* m0, int32_t* m, int W, int H, int* ok, hipStream_t s) {
    const size_t N = (size_t)W * H;
    const bool dbg = getenv("NSC_WS_DEBUG") != nullptr;
    hipEvent_t ev0 = nullptr;
    auto tick = [&](const char* what, const int* dcount) {
        if (!dbg) return;
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s);
        (void)hipEventSynchronize(e);
        float ms = 0.f;
        if (ev0) (void)hipEventElapsedTime(&ms, ev0, e);
        int c = -1;
        if (dcount) (void)hipMemcpy(&c, dcount, sizeof(int), hipMemcpyDeviceToHost);
        fprintf(stderr, "[ws] %-12s %8.3f ms  count %d\n", what, ms, c);
        if (ev0) (void)hipEventDestroy(ev0);
        ev0 = e;
    };
    tick("start", nullptr);
    *ok = 0;
    if (N >= (size_t)INT_MAX / 8) return NSC_OK;
    Carver measure(nullptr);
    auto carve = [&](Carver& c, uint8_t*& cls, uint16_t*& lvl0, uint8_t*& fmask, int32_t*& run, uint8_t*& st,
                     uint8_t*& ilvl, int32_t*& pseq, int32_t*& parent, int32_t*& bnode, int32_t*& popidx, int32_t*& lab,
                     int32_t*& head, int32_t*& real, int32_t*& lab2, uint8_t*& kind, int*& small) {
        cls = c.take<uint8_t>(N);
        lvl0 = c.take<uint16_t>(N);
        fmask = c.take<uint8_t>(N);
        run = c.take<int32_t>(N);
        st = c.take<uint8_t>(N);
        ilvl = c.take<uint8_t>(N);
        pseq = c.take<int32_t>(N);
        parent = c.take<int32_t>(N);
        bnode = c.take<int32_t>(N);
        popidx = c.take<int32_t>(N);
        lab = c.take<int32_t>(N);
        head = c.take<int32_t>(N);
        real = c.take<int32_t>(N);
        lab2 = c.take<int32_t>(N);
        kind = c.take<uint8_t>(N);
        small = c.take<int>(64 + B1_BLOCKS);  // counters, then the per-block b1 winners
    };
    uint8_t *cls, *fmask, *st, *ilvl, *kind;
    uint16_t* lvl0;
    int32_t *run, *pseq, *parent, *bnode, *popidx, *lab, *head, *real, *lab2;
    int* small;
    carve(measure, cls, lvl0, fmask, run, st, ilvl, pseq, parent, bnode, popidx, lab, head, real, lab2, kind, small);
    Scratch sc;
    NSC_CHECK(sc.alloc(measure.size(), s));
    Carver cut(sc.p);
    carve(cut, cls, lvl0, fmask, run, st, ilvl, pseq, parent, bnode, popidx, lab, head, real, lab2, kind, small);
    // small[]: 0 nF, 1 nrun, 2 nleaf, 3 flags, 4 ncand, 8..10 info, 16 ncomp, 20 deferred tie components,
    // 24..25 tie pool bump pointer (long long), 26..31 run bucket sizes, 32..34 LDS flood queue counts,
    // 35 B pixels with queue entries, 40..44 phase-1 tier work counters
    NSC_HIP(hipMemsetAsync(small, 0, 64 * sizeof(int), s));

    Dev D{};
    D.img = img; D.W = W; D.H = H; D.N = (int)N; D.m0 = m0; D.m = m; D.cls = cls; D.lvl0 = lvl0; D.run = run;
    D.st = st; D.ilvl = ilvl; D.pseq = pseq; D.parent = parent; D.bnode = bnode; D.popidx = popidx; D.lab = lab;
    D.head = head; D.nleaf = &small[2]; D.flags = &small[3];
    {  // real[] is unused until B resolution: it holds the per-workgroup counts
        const unsigned nb = (unsigned)std::min<size_t>(cdiv(N, 256), 1u << 14);
        hipLaunchKernelGGL(k_classify, dim3(nb), dim3(256), 0, s, D, cls, lvl0, fmask, real);
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1024), 0, s, real, (int)nb, &small[0]);
    }
    tick("classify", &small[0]);
    NSC_CHECK(ccl_neq(fmask, 0, run, W, H, 4, s));
    NSC_CHECK(relabel_impl(run, W, H, -1, &small[1], lab2, s));
    tick("runs", &small[1]);
    int hostc[2];
    NSC_HIP(hipMemcpyAsync(hostc, small, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    const int nF = hostc[0], nrun = hostc[1];
    // phase-1 scratch: run boxes, per-run FIFO heads/tails, queue links, leaf table
    const int leafcap = std::max(1024, std::min((int)std::min(N, (size_t)INT_MAX / 4), 4 * nF + 1024));
    Carver m2(nullptr);
    auto carve2 = [&](Carver& c, int32_t*& box5, int*& qh, int*& qt, int32_t*& qnext, int32_t*& lf) {
        box5 = c.take<int32_t>(5 * (size_t)std::max(nrun, 1));
        qh = c.take<int>((size_t)std::max(nrun, 1) * (NQ + kRunTiers));  // run buckets, then FIFO heads
        qt = c.take<int>((size_t)std::max(nrun, 1) * NQ);
        qnext = c.take<int32_t>(N + (size_t)leafcap);
        lf = c.take<int32_t>((size_t)leafcap * 8);
    };
    int32_t *box5, *qnext, *lf;
    int *qh, *qt;
    carve2(m2, box5, qh, qt, qnext, lf);
    Scratch sc2;
    NSC_CHECK(sc2.alloc(m2.size(), s));
    Carver cut2(sc2.p);
    carve2(cut2, box5, qh, qt, qnext, lf);
    D.qnext = qnext;
    D.leafcap = leafcap;
    D.lf_b = lf;
    D.lf_run = lf + (size_t)leafcap;
    D.lf_pseq = lf + 2 * (size_t)leafcap;
    D.lf_parent = lf + 3 * (size_t)leafcap;
    D.lf_bnode = lf + 4 * (size_t)leafcap;
    D.lf_popidx = lf + 5 * (size_t)leafcap;
    D.lf_next = lf + 6 * (size_t)leafcap;
    D.lf_lvl = reinterpret_cast<uint8_t*>(lf + 7 * (size_t)leafcap);
    if (nrun > 0) {
        NSC_CHECK(bbox_impl(run, W, H, nrun, box5, nullptr, s));
        int* lists = reinterpret_cast<int*>(qh);  // kRunTiers * nrun ints
        int* bcount = &small[26];                  // kRunTiers counters
        int* qh2 = qh + (size_t)kRunTiers * nrun;  // FIFO heads for the global-flood bucket
        hipLaunchKernelGGL(k_bucket_runs, dim3(cdiv(nrun, 256)), dim3(256), 0, s, box5, nrun, lists, bcount);
        if (dbg) {
            int bc[kRunTiers];
            (void)hipMemcpy(bc, bcount, sizeof(bc), hipMemcpyDeviceToHost);
            fprintf(stderr, "[ws] run tiers %d %d %d %d %d %d\n", bc[0], bc[1], bc[2], bc[3], bc[4], bc[5]);
            std::vector<int32_t> hb((size_t)5 * nrun);
            (void)hipMemcpy(hb.data(), box5, sizeof(int32_t) * hb.size(), hipMemcpyDeviceToHost);
            long long area[kRunTiers] = {}, amax[kRunTiers] = {};
            for (int r = 0; r < nrun; ++r) {
                const long long a = (long long)(hb[2 * (size_t)nrun + r] - hb[nrun + r] + 1 + 2 * P1_MARGIN) *
                                    (hb[4 * (size_t)nrun + r] - hb[3 * (size_t)nrun + r] + 1 + 2 * P1_MARGIN);
                const int tier[kRunTiers - 1] = {512, 1024, 2048, 4096, 6400};
                int b = 0;
                while (b < kRunTiers - 1 && a > tier[b]) ++b;
                area[b] += a;
                amax[b] = std::max(amax[b], a);
            }
            for (int b = 0; b < kRunTiers; ++b)
                fprintf(stderr, "[ws] tier %d: runs %d, box px %lld, max %lld\n", b, bc[b], area[b], amax[b]);
        }
        auto L = [&](int b) { return lists + (size_t)b * nrun; };
        // the few large runs set the latency: the largest tiers start first, each on a side stream,
        // beside the many small runs on the caller's stream (the 1024 tier, then the 512 tier: the
        // tiers share the LDS, so the order decides little; measured against the 512 tier after the
        // 2048 tier, this was 0.07 ms shorter on the same box)
        // grids: the workgroups the device holds at once per tier (idle ones exit at once); on a
        // serving thread the tier counts are read and each grid holds one workgroup per run at most
        // (idle workgroups of the large tiers would wait for a CU with their LDS free, beside the
        // other slots' kernels), and empty tiers launch nothing
        int tier_runs[kRunTiers];
        const bool exact = thread_serving();
        if (exact) {
            NSC_HIP(hipMemcpyAsync(tier_runs, bcount, sizeof(tier_runs), hipMemcpyDeviceToHost, s));
            NSC_HIP(stream_sync(s));
        }
        auto grid = [&](int b, int resident) { return exact ? std::min(tier_runs[b], resident) : resident; };
        SideStream& ss = side_stream();
        NSC_CHECK(ss.fork(s));
        int* next_run = &small[40];  // per-tier work counters (zeroed with small[])
        if (const int g = grid(4, std::min(nrun, 64)))
            hipLaunchKernelGGL(k_phase1_lds<6400>, dim3(g), dim3(64), 0, ss.side[0], D, box5, nrun, L(4), bcount + 4,
                               next_run + 4);
        if (const int g = exact ? (int)cdiv(tier_runs[5], 64) : (int)cdiv(nrun, 64))
            hipLaunchKernelGGL(k_phase1, dim3(g), dim3(64), 0, ss.side[0], D, box5, nrun, L(5), bcount + 5, qh2, qt);
        if (const int g = grid(3, p1_grid<4096>(nrun)))
            hipLaunchKernelGGL(k_phase1_lds<4096>, dim3(g), dim3(64), 0, ss.side[1], D, box5, nrun, L(3), bcount + 3,
                               next_run + 3);
        if (const int g = grid(2, p1_grid<2048>(nrun)))
            hipLaunchKernelGGL(k_phase1_lds<2048>, dim3(g), dim3(64), 0, ss.side[2], D, box5, nrun, L(2), bcount + 2,
                               next_run + 2);
        if (const int g = grid(1, p1_grid<1024>(nrun)))
            hipLaunchKernelGGL(k_phase1_lds<1024>, dim3(g), dim3(64), 0, s, D, box5, nrun, L(1), bcount + 1, next_run + 1);
        if (const int g = grid(0, p1_grid<512>(nrun)))
            hipLaunchKernelGGL(k_phase1_lds<512>, dim3(g), dim3(64), 0, s, D, box5, nrun, L(0), bcount, next_run);
        NSC_CHECK(ss.join(s));
    }
    // B pixels with a queue entry: initially queued ones, then those with leaves (the phase-1
    // queue links are dead from here on and hold the list)
    int* blist = qnext;  // N + leafcap ints
    hipLaunchKernelGGL(k_collect_b0, grid1(N), dim3(256), 0, s, D, blist, &small[35]);
    if (nrun > 0) hipLaunchKernelGGL(k_link_leaves, grid1(N), dim3(256), 0, s, D, blist, &small[35]);
    // item records for the pop-order walks (every walk comes after this point)
    Scratch sc_items;
    NSC_CHECK(sc_items.alloc((sizeof(int4) + 1) * ((size_t)N + leafcap) + 256, s));
    D.wr = sc_items.as<int4>();
    D.wl = reinterpret_cast<uint8_t*>(D.wr + ((size_t)N + leafcap));
    hipLaunchKernelGGL(k_pack_items, grid1(N), dim3(256), 0, s, D);
    tick("phase1", &small[2]);
    // B resolution and b1
    int* cand = lab2;  // reuse (N ints)
    NSC_HIP(hipMemsetAsync(real, 0xFF, sizeof(int32_t) * N, s));  // -1 everywhere
    hipLaunchKernelGGL(k_bresolve, grid_counted(N), dim3(256), 0, s, D, real, cand, &small[4], blist, &small[35]);
    tick("bresolve", &small[4]);
    {
        int* partial = small + 64;
        hipLaunchKernelGGL(k_b1_partial, dim3(B1_BLOCKS), dim3(256), 0, s, D, real, cand, &small[4], partial);
        hipLaunchKernelGGL(k_b1_reduce, dim3(1), dim3(1024), 0, s, D, real, partial, &small[8]);
    }
    tick("b1", &small[8]);
    hipLaunchKernelGGL(k_t1_state, grid1(N), dim3(256), 0, s, D, real, &small[8], fmask, kind);
    // flood from b1 over the never-queued B pixels (fmask marks them): geodesic distance from b1
    {
        unsigned* dist = reinterpret_cast<unsigned*>(head);  // head[] is dead after k_bresolve
        hipLaunchKernelGGL(k_geo_init, grid1(N), dim3(256), 0, s, D, &small[8], fmask, dist);
        const int grc = geodist_impl(fmask, dist, W, H, s, &small[8]);  // the seed: b1
        if (grc == NSC_ERR_LIMIT) {  // outside the tile search's range: the sequential flood decides
            clear_error();  // the call still succeeds: no error text left behind
            ws_last_flags_record(FB_CMP);
            return NSC_OK;
        }
        NSC_CHECK(grc);
        tick("geodist", nullptr);
        hipLaunchKernelGGL(k_bfs_mark, grid1(N), dim3(256), 0, s, D, &small[8], dist, kind, lab /* pdist; lab[] is dead */);
        hipLaunchKernelGGL(k_bfs_commit, grid1(N), dim3(256), 0, s, D, &small[8], dist, kind, fmask);
    }
    // phase 2
    NSC_CHECK(ccl_neq(fmask, 0, lab2, W, H, 4, s));
    NSC_CHECK(relabel_impl(lab2, W, H, -1, &small[16], head /* rank scratch: head[] is dead after k_bresolve */, s));
    int sm16[16];  // small[16..31]: ncomp and, at 26..31, the phase-1 run counts per tier
    NSC_HIP(hipMemcpyAsync(sm16, &small[16], sizeof(sm16), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    const int ncomp = sm16[0];
    ws_tier_record(sm16 + 10, kRunTiers);
    if (ncomp > 0) {
        Scratch sc3;
        Carver m3(nullptr);
        auto carve3 = [&](Carver& c, int32_t*& b5, int32_t*& area, int32_t*& off, int32_t*& scr, long long*& hk) {
            b5 = c.take<int32_t>(5 * (size_t)ncomp);
            area = c.take<int32_t>(ncomp);
            off = c.take<int32_t>(ncomp);
            scr = c.take<int32_t>(3 * N);
            hk = c.take<long long>(N);
        };
        int32_t *b5, *area, *off, *scr;
        long long* hk;
        carve3(m3, b5, area, off, scr, hk);
        NSC_CHECK(sc3.alloc(m3.size(), s));
        Carver cut3(sc3.p);
        carve3(cut3, b5, area, off, scr, hk);
        NSC_CHECK(bbox_impl(lab2, W, H, ncomp, b5, area, s));
        if (dbg) {  // phase-2 components by where they flood (tie-free ones; a tie moves a small one to LDS)
            std::vector<int32_t> hb(5 * (size_t)ncomp), ha(ncomp);
            (void)hipMemcpy(hb.data(), b5, sizeof(int32_t) * hb.size(), hipMemcpyDeviceToHost);
            (void)hipMemcpy(ha.data(), area, sizeof(int32_t) * ha.size(), hipMemcpyDeviceToHost);
            long long nbig = 0, maxa_big = 0, maxl = 0, nsmall = 0, nlds = 0, n1 = 0, n2 = 0;
            for (int c = 0; c < ncomp; ++c) {
                n1 += ha[c] == 1;
                n2 += ha[c] == 2;
                const long long L = (long long)(hb[2 * ncomp + c] - hb[ncomp + c] + 3) * (hb[4 * ncomp + c] - hb[3 * ncomp + c] + 3);
                maxl = std::max(maxl, L);
                if (L > FL_BIG) { ++nbig; maxa_big = std::max<long long>(maxa_big, ha[c]); }
                else if (ha[c] > FL_MIN_AREA) ++nlds;
                else ++nsmall;
            }
            fprintf(stderr, "[ws] p2 comps %d: lane-flooded %lld (area 1: %lld, 2: %lld), LDS %lld, global %lld (max area %lld), max frame %lld\n",
                    ncomp, nsmall, n1, n2, nlds, nbig, maxa_big, maxl);
        }
        {  // offsets of the components' scratch: exclusive scan of their areas
            size_t tb = 0;
            NSC_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, area, off, ncomp, s));
            Scratch st;
            NSC_CHECK(st.alloc(tb, s));
            NSC_HIP(hipcub::DeviceScan::ExclusiveSum(st.p, tb, area, off, ncomp, s));
        }
        P2 Q{lab2, b5, area, off, kind, scr, ncomp};
        tick("p2 setup", &small[16]);
        const int tiecap = 4096;
        const long long poolcap = std::max<long long>(16ll << 20, 4ll * (long long)N);  // ints
        Scratch sc4;
        NSC_CHECK(sc4.alloc(sizeof(TieRec) * tiecap + sizeof(int) * poolcap + 256, s));
        TieRec* ties = reinterpret_cast<TieRec*>(sc4.p);
        int* pool = reinterpret_cast<int*>(reinterpret_cast<char*>(sc4.p) + ((sizeof(TieRec) * tiecap + 255) & ~size_t(255)));
        long long* pool_used = reinterpret_cast<long long*>(&small[24]);  // 8-byte aligned: small is 256-aligned
        // records and work lists of the LDS flood
        const int workcap = ncomp + (1 << 18);
        Scratch sc5;
        NSC_CHECK(sc5.alloc(sizeof(TieRec) * (size_t)ncomp + 2 * sizeof(int) * (size_t)workcap + sizeof(int) * (size_t)ncomp + 768,
                            s));
        FloodQ F{};
        F.recs = reinterpret_cast<TieRec*>(sc5.p);
        F.reccap = ncomp;
        F.work[0] = reinterpret_cast<int*>(reinterpret_cast<char*>(sc5.p) + ((sizeof(TieRec) * (size_t)ncomp + 255) & ~size_t(255)));
        F.work[1] = F.work[0] + workcap;
        F.workcap = workcap;
        F.nrec = &small[32];
        F.nwork = &small[33];  // [33], [34]
        F.res = pool;
        F.recbad = reinterpret_cast<int*>(reinterpret_cast<char*>(F.work[1] + workcap) + 256 -
                                          (reinterpret_cast<uintptr_t>(F.work[1] + workcap) & 255));
        NSC_HIP(hipMemsetAsync(F.recbad, 0, sizeof(int) * (size_t)ncomp, s));
        hipLaunchKernelGGL(k_phase2, dim3(cdiv(ncomp, 64)), dim3(64), 0, s, D, Q, real, hk, ties, &small[20], tiecap,
                           pool_used, poolcap, F);
        tick("phase2", &small[20]);
        hipLaunchKernelGGL(k_flood_lds<FL_BIG>, dim3(256), dim3(64), 0, s, D, Q, real, F, 1);
        hipLaunchKernelGGL(k_flood_lds<FL_SMALL>, dim3(1024), dim3(64), 0, s, D, Q, real, F, 0);
        hipLaunchKernelGGL(k_tie_check, dim3(2048), dim3(256), 0, s, Q, F, W);
        hipLaunchKernelGGL(k_tie_commit, dim3(1024), dim3(256), 0, s, D, Q, F);
        tick("lds flood", &small[32]);
        hipLaunchKernelGGL(k_phase2_ties, dim3(tiecap), dim3(64), 0, s, D, Q, real, ties, &small[20], pool);
        tick("ties", &small[3]);
    }
    if (ev0) (void)hipEventDestroy(ev0);
    int flags = 0;
    NSC_HIP(hipGetLastError());
    NSC_HIP(hipMemcpyAsync(&flags, &small[3], sizeof(int), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    *ok = (flags == 0) ? 1 : 0;
    ws_last_flags_record(flags);
    return NSC_OK;
}

}  // namespace nsc
