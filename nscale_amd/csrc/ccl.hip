// Connected components and everything built on them:
//   ConnComponents::label / bwlabel2 / relabel / areaThresholdLabeled / boundingBox
//   (ConnComponents.cpp:24-284, MorphologicOperations.cpp:3092-3117, 3281-3314), and the binary
//   reconstruction family imreconstructBinary / imfill / imfillHoles / bwselect
//   (MorphologicOperations.cpp:2768-3031), which are component selections: a pixel of the
//   passable set is reached iff its component contains a seed.
//
// Labelling is union-find with the smallest raster index as root (the reference's merge links
// the larger root under the smaller, ConnComponents.cpp:34-43, so its labels are also the
// smallest index of each component):
//   1. ccl_local: 64x64 tiles in LDS, pixels start at the first pixel of their row run, runs
//      of adjacent rows unite with atomicMin on local indices (local raster order = global
//      raster order inside a tile), then each pixel stores the global index of its tile-local
//      root;
//   2. ccl_seams: pixels on tile seams unite with their cross-tile neighbours in global memory
//      (atomicMin union, retry while the target stopped being a root);
//   3. ccl_flatten: every pixel follows its chain to the root (a read-only walk).
// Pixel classes: FG_NEQ  -> foreground = value != key; neighbours join when both foreground
//                FG_EQ   -> foreground = value != 0; neighbours join when values are equal
//                FG_FEQ  -> foreground = cls != 0; neighbours join when float values are equal
#include <algorithm>
#include <climits>
#include <type_traits>

#include "common.hpp"

namespace nsc {
namespace {

constexpr int CT_W = 64, CT_H = 64, CT_N = CT_W * CT_H;
enum { FG_NEQ = 0, FG_EQ = 1, FG_FEQ = 2 };

struct Px {
    const uint8_t* img;
    const float* fv;
    uint8_t key;
};


// find with intermediate pointer jumping (each visited node is re-pointed to its grandparent).
// Parents are always smaller than children, and a non-root never becomes a root again, so
// re-pointing a non-root to any ancestor keeps every set intact; a concurrent union that lowers
// the same pointer continues its own loop with the value it replaced (see unite), so a lost
// update only shortens a path.
// relaxed atomic accesses: other threads update the forest concurrently, so no access may be
// kept in a register by the compiler
__device__ __forceinline__ int ld_wg(const int* a) { return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ __forceinline__ void st_wg(int* a, int v) { __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }

__device__ __forceinline__ int lfind(int* L, int x) {
    int p = ld_wg(&L[x]);
    if (p == x) return x;
    int q = ld_wg(&L[p]);
    while (p != q) {
        st_wg(&L[x], q);
        x = p;
        p = q;
        q = ld_wg(&L[p]);
    }
    return p;
}

__device__ void lunite(int* L, int a, int b) {
    while (true) {
        a = lfind(L, a);
        b = lfind(L, b);
        if (a == b) return;
        if (a > b) { int t = a; a = b; b = t; }
        int old = atomicMin(&L[b], a);
        if (old == b) return;
        b = old;
    }
}

// relaxed device-scope accesses: unions from every workgroup of the launch update the forest
__device__ __forceinline__ int ld_rel(const int* a) { return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void st_rel(int* a, int v) { __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ __forceinline__ int gfind(int32_t* L, int x) {
    int p = ld_rel(&L[x]);
    if (p == x) return x;
    int q = ld_rel(&L[p]);
    while (p != q) {
        st_rel(&L[x], q);
        x = p;
        p = q;
        q = ld_rel(&L[p]);
    }
    return p;
}

// Union in global memory of two ancestors a, b of the pixels to join (the labels a seam pixel has
// already read, usually tile-local roots that are still roots): both root checks are issued before
// either is used, and a find runs only for an ancestor that stopped being a root.
__device__ void gunite(int32_t* L, int a, int b) {
    const int pa = ld_rel(&L[a]), pb = ld_rel(&L[b]);
    if (pa != a) a = gfind(L, a);
    if (pb != b) b = gfind(L, b);
    while (true) {
        if (a == b) return;
        if (a > b) { int t = a; a = b; b = t; }
        const int old = atomicMin(&L[b], a);
        if (old == b) return;
        b = old;
        a = gfind(L, a);
        b = gfind(L, b);
    }
}

// One 64 x 64 tile per workgroup; a row of the tile is one wave (lane = column), and thread t
// handles rows t / 64, t / 64 + 4, ... .  Horizontal joins need no union: each pixel starts out
// pointing at the first pixel of its run of left-joined pixels (from a ballot of the row).  A
// vertical or diagonal join is united only where no join further left already connects the
// same two runs:
//   up (k, k-64):        skipped when k joins k-1, k-1 joins up, and k-64 joins k-65;
//   up-left (k, k-65):   skipped when k joins up and k-64 joins k-65, or k joins k-1 and k-1
//                        joins up;
//   up-right (k, k-63):  skipped when k joins up and k-63 joins k-64, or k+1 joins k and k+1
//                        joins up.
// Every skipped join is implied by a path through joins that are united or skipped only by a rule
// that looks further left, so the components are those of all joins.
template <int MODE, int CONN>
__global__ __launch_bounds__(256) void ccl_local(Px P, int32_t* __restrict__ L, int w, int h, int tiles_x,
                                                 uint8_t* __restrict__ clear8) {
    constexpr int PER = CT_N / 256;  // rows per thread
    using V = typename std::conditional<MODE == FG_FEQ, float, uint32_t>::type;
    __shared__ int ll[CT_N];
    // pixel values the joins compare (equal-value modes)
    __shared__ typename std::conditional<MODE == FG_FEQ, float, uint8_t>::type lv[MODE == FG_NEQ ? 1 : CT_N];
    __shared__ unsigned long long jrow[CT_H];  // bit x of row y: pixel x joins pixel x - 1
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int x0 = tx * CT_W, y0 = ty * CT_H;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // every load first, from clamped (always valid) addresses: a load per iteration with its use in
    // the same iteration compiles to one memory round trip after another
    uint8_t c[PER];
    float f[PER];
    const auto gimg = as_global(P.img);
    const size_t gxc = (size_t)clampi(x0 + lane, 0, w - 1);
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const size_t i = (size_t)clampi(y0 + wv + 4 * j, 0, h - 1) * w + gxc;
        c[j] = gimg[i];
        if (MODE == FG_FEQ) f[j] = as_global(P.fv)[i];
    }
    auto val = [&](int j) -> V {
        if constexpr (MODE == FG_FEQ) return f[j];
        else return (V)c[j];
    };
    auto left_of = [&](V v) -> V { return from_left((V)0, v); };
    unsigned fgbits = 0;  // bit j: row wv + 4 j of this column is foreground
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = threadIdx.x + 256 * j, ly = wv + 4 * j;
        const bool in = x0 + lane < w && y0 + ly < h;
        const bool fg = in && (MODE == FG_NEQ ? c[j] != P.key : c[j] != 0);
        fgbits |= (fg ? 1u : 0u) << j;
        if (MODE == FG_EQ) lv[k] = c[j];
        if (MODE == FG_FEQ) lv[k] = f[j];
        // the left neighbour is lane - 1 of the same wave (lane 0 receives 0: no left join).  Every
        // cross-lane read runs on all lanes (none of them under a condition): a lane that is
        // switched off supplies nothing to its neighbour.
        const bool fg_left = from_left(0, fg ? 1 : 0) != 0;
        const V v_left = left_of(val(j));
        bool jl = fg && fg_left;
        if (MODE != FG_NEQ) jl = jl && val(j) == v_left;
        const unsigned long long m = __ballot(jl);
        // run start: the last position <= lane without a left join (bit 0 is never a join)
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const int start = 63 - __clzll((long long)(~m & upto));
        ll[k] = fg ? ly * CT_W + start : -1;
        if (lane == 0) jrow[ly] = m;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = threadIdx.x + 256 * j, ly = wv + 4 * j;
        if (ly == 0) continue;  // uniform over the wave
        const bool fg = (fgbits >> j) & 1u;
        const bool up = ll[k - CT_W] >= 0;
        V vu = 0;
        if (MODE != FG_NEQ) vu = (V)lv[k - CT_W];
        const unsigned long long jr = jrow[ly], ju_row = jrow[ly - 1];
        const bool jl = (jr >> lane) & 1ull;
        const bool jl_up = (ju_row >> lane) & 1ull;  // k - 64 joins k - 65
        bool ju = fg && up;
        if (MODE != FG_NEQ) ju = ju && val(j) == vu;
        const int ju_i = ju ? 1 : 0, up_i = up ? 1 : 0;
        const bool ju_left = from_left(0, ju_i) != 0;
        const bool do_up = ju && !(jl && ju_left && jl_up);
        bool do_ul = false, do_ur = false;
        if (CONN == 8) {
            const bool up_left = from_left(0, up_i) != 0, up_right = from_right(0, up_i) != 0;
            const bool ju_right = from_right(0, ju_i) != 0;
            const V vul = left_of(vu), vur = from_right((V)0, vu);
            bool jul = fg && up_left, jur = fg && up_right;
            if (MODE != FG_NEQ) {
                jul = jul && val(j) == vul;
                jur = jur && val(j) == vur;
            }
            const bool jl_upright = lane < 63 && ((ju_row >> (lane + 1)) & 1ull);  // k - 63 joins k - 64
            const bool jl_right = lane < 63 && ((jr >> (lane + 1)) & 1ull);        // k + 1 joins k
            do_ul = jul && !((ju && jl_up) || (jl && ju_left));
            do_ur = jur && !((ju && jl_upright) || (jl_right && ju_right));
        }
        if (do_up) lunite(ll, k, k - CT_W);
        if (do_ul) lunite(ll, k, k - CT_W - 1);
        if (do_ur) lunite(ll, k, k - CT_W + 1);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int k = threadIdx.x + 256 * j, ly = wv + 4 * j;
        const int gx = x0 + lane, gy = y0 + ly;
        int r = ll[k];
        int32_t out = -1;
        if (r >= 0) {
            r = lfind(ll, k);
            out = (int32_t)((size_t)(y0 + r / CT_W) * w + (x0 + r % CT_W));
            // clear8: entry cleared at every tile-local root (every final root is one), so a per-root
            // table filled while flattening needs no image-sized fill
            if (clear8 && r == k) clear8[out] = 0;
        }
        if (gx < w && gy < h) L[(size_t)gy * w + gx] = out;
    }
}

// unions across tile seams: every pixel in a tile's first row, first column or last column
// joins its "backward" neighbours (left, up, up-left, up-right) that lie in another tile.  Along a
// seam, consecutive pixels usually link the same two tile-local components: a pixel skips a union
// when the previous pixel of the seam joins, in the same direction, two pixels whose labels are the
// ones it sees (the labels read are ancestors of those pixels, so the previous pixel's union joins
// the same two sets).
template <int MODE, int CONN>
__global__ __launch_bounds__(256) void ccl_seams(Px P, int32_t* L, int w, int h, int tiles_x, int ntiles) {
    constexpr int ND = CONN == 8 ? 4 : 2;
    // with 4-connectivity a last-column pixel has no backward neighbour in another tile (left and
    // up are in its own tile; the tile's top row covers the corner), so only 8 needs that column
    const int per_tile = CONN == 8 ? CT_W + 2 * CT_H : CT_W + CT_H;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= ntiles * per_tile) return;
    const int t = gid / per_tile, e = gid % per_tile;
    const int tx = t % tiles_x, ty = t / tiles_x;
    int lx, ly, sx, sy;  // position on the seam and the step to the previous seam pixel
    if (e < CT_W) { lx = e; ly = 0; sx = -1; sy = 0; }
    else if (e < CT_W + CT_H) { lx = 0; ly = e - CT_W; sx = 0; sy = -1; }
    else { lx = CT_W - 1; ly = e - CT_W - CT_H; sx = 0; sy = -1; }
    const bool has_prev = (sx ? lx : ly) > 0;
    const int x = tx * CT_W + lx, y = ty * CT_H + ly;
    if (x >= w || y >= h) return;
    const int px = x + sx, py = y + sy;
    const int dxs[4] = {-1, 0, -1, 1}, dys[4] = {0, -1, -1, -1};
    // every label and value this pixel needs, loaded together from clamped (always valid)
    // addresses; the checks below discard what lies outside the image or inside the tile
    auto at = [&](int qx, int qy) { return (size_t)clampi(qy, 0, h - 1) * w + clampi(qx, 0, w - 1); };
    const size_t i = (size_t)y * w + x, ip = at(px, py);
    int lj[ND], ljp[ND];
    size_t j[ND], jp[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        j[d] = at(x + dxs[d], y + dys[d]);
        jp[d] = at(px + dxs[d], py + dys[d]);
    }
    const int li = ld_wg(&L[i]), lp = ld_wg(&L[ip]);
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        lj[d] = ld_wg(&L[j[d]]);
        ljp[d] = ld_wg(&L[jp[d]]);
    }
    // the values the equal-value joins compare, loaded with the labels (inside the checks below
    // each would be one more dependent round trip)
    using V = typename std::conditional<MODE == FG_FEQ, float, uint8_t>::type;
    V vi = 0, vp = 0, vj[ND], vjp[ND];
    if (MODE != FG_NEQ) {
        auto val = [&](size_t q) -> V {
            if constexpr (MODE == FG_FEQ) return P.fv[q];
            else return P.img[q];
        };
        vi = val(i);
        vp = val(ip);
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            vj[d] = val(j[d]);
            vjp[d] = val(jp[d]);
        }
    }
    if (li < 0) return;
    auto cross = [&](int qx, int qy) { return (qx / CT_W != tx) || (qy / CT_H != ty); };
    auto valid = [&](int qx, int qy) { return qx >= 0 && qy >= 0 && qx < w && qy < h && cross(qx, qy); };
    const int lprev = has_prev ? lp : -1;
    // Along a seam, consecutive pixels usually link the same two tile-local components: a pixel
    // skips its union when the previous seam pixel joins, in the same direction, two pixels whose
    // labels are the ones it sees (labels read are ancestors of those pixels, so the previous
    // pixel's union joins the same two sets).
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const int qx = x + dxs[d], qy = y + dys[d];
        if (!valid(qx, qy) || lj[d] < 0 || (MODE != FG_NEQ && !(vi == vj[d]))) continue;
        if (has_prev && lprev == li) {
            const int rx = px + dxs[d], ry = py + dys[d];
            if (valid(rx, ry) && ljp[d] == lj[d] && (MODE == FG_NEQ || vp == vjp[d])) continue;  // same union
        }
        gunite(L, li, lj[d]);  // labels are ancestors of their pixels
    }
}

// Four pixels per thread, i, i + S, i + 2S, i + 3S (S = threads in the grid): each of the four
// loads is coalesced over the wave, and the four labels and the four first hops are each issued
// together instead of one dependent round trip after another.
// `zero_at_roots` / `zero8_at_roots` (optional): entry i cleared for every root i, so a per-root
// table (areas, selection flags) needs no image-sized fill before it is filled in.
// `flag_set` (optional): flag_set[root] = 1 for every component holding a pixel with
// flag_src[i] != 0 (flag_src null: a pixel on the border of the fw x fh image); the entries must
// have been cleared at the roots before (ccl_local's clear8)
__global__ void ccl_flatten(int32_t* L, size_t n, int32_t bgval, int32_t* __restrict__ zero_at_roots,
                            uint8_t* __restrict__ zero8_at_roots, const uint8_t* __restrict__ flag_src,
                            uint8_t* flag_set, int fw, int fh) {
    const size_t S = (size_t)gridDim.x * blockDim.x;
    for (size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += 4 * S) {
        int p[4], q[4];
        bool sv[4] = {false, false, false, false};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = i0 + j * S;
            p[j] = i < n ? ld_wg(&L[i]) : -1;
            if (flag_set && i < n) {
                if (flag_src) {
                    sv[j] = flag_src[i] != 0;
                } else {
                    const int x = (int)(i % fw), y = (int)(i / fw);
                    sv[j] = x == 0 || y == 0 || x == fw - 1 || y == fh - 1;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = p[j] >= 0 ? ld_wg(&L[p[j]]) : p[j];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = i0 + j * S;
            if (i >= n) continue;
            if (p[j] < 0) {
                if (bgval != -1) L[i] = bgval;
                continue;
            }
            // read-only walk: a path-jumping find here could re-point a pixel whose final root
            // another thread has already written back to an intermediate ancestor.  Roots are final
            // in this kernel and every value an entry takes is an ancestor, so cached loads are safe.
            if (zero_at_roots && p[j] == (int)i) zero_at_roots[i] = 0;
            if (zero8_at_roots && p[j] == (int)i) zero8_at_roots[i] = 0;
            int r = p[j], nx = q[j];
            while (nx != r) {
                r = nx;
                nx = ld_wg(&L[r]);
            }
            if (r != p[j]) L[i] = r;
            if (sv[j]) flag_set[r] = 1;
        }
    }
}

template <int MODE>
int ccl_run(const Px& P, int32_t* L, int w, int h, int conn, int32_t bgval, hipStream_t s,
            int32_t* zero_at_roots = nullptr, uint8_t* zero8_at_roots = nullptr, uint8_t* flag_set = nullptr,
            const uint8_t* flag_src = nullptr) {
    const int tiles_x = (int)cdiv(w, CT_W), tiles_y = (int)cdiv(h, CT_H);
    const int nt = tiles_x * tiles_y;
    const int nseam8 = nt * (CT_W + 2 * CT_H), nseam4 = nt * (CT_W + CT_H);
    const size_t n = (size_t)w * h;
    if (conn == 8) {
        hipLaunchKernelGGL((ccl_local<MODE, 8>), dim3(nt), dim3(256), 0, s, P, L, w, h, tiles_x, flag_set);
        hipLaunchKernelGGL((ccl_seams<MODE, 8>), dim3(cdiv(nseam8, 256)), dim3(256), 0, s, P, L, w, h, tiles_x, nt);
    } else {
        hipLaunchKernelGGL((ccl_local<MODE, 4>), dim3(nt), dim3(256), 0, s, P, L, w, h, tiles_x, flag_set);
        hipLaunchKernelGGL((ccl_seams<MODE, 4>), dim3(cdiv(nseam4, 256)), dim3(256), 0, s, P, L, w, h, tiles_x, nt);
    }
    hipLaunchKernelGGL(ccl_flatten, dim3(std::min<unsigned>(cdiv(n, 1024), 65535u)), dim3(256), 0, s, L, n, bgval,
                       zero_at_roots, zero8_at_roots, flag_src, flag_set, w, h);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

// ------------------------------------------------------------------------------------------
// operators on top of the labels
// ------------------------------------------------------------------------------------------
// Component areas with one atomic per run of equal labels in each wave's 64 consecutive pixels
// (coalesced loads; the background component of a fill would otherwise take one atomic per pixel
// on a single address).  A lane starts a run when its label differs from the previous lane's;
// the run ends where the next one starts.
__global__ void k_area(const int32_t* __restrict__ L, int32_t* __restrict__ area, size_t n) {
    const int lane = threadIdx.x & 63;
    const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const size_t nwave = ((size_t)gridDim.x * blockDim.x) >> 6;
    for (size_t base = wave * 64; base < n; base += nwave * 64) {
        const size_t i = base + lane;
        const int r = i < n ? L[i] : -2;
        const int prev = __shfl_up(r, 1);
        const bool head = (lane == 0 || prev != r) && r != -2;
        const unsigned long long heads = __ballot(head);
        if (head && r >= 0) {
            // next run start after this lane (or the end of the window / of the image)
            const unsigned long long later = lane == 63 ? 0ull : (heads >> (lane + 1));
            int end = later ? lane + 1 + __ffsll((long long)later) - 1 : 64;
            const size_t lim = n - base;  // valid lanes in this window
            if ((size_t)end > lim) end = (int)lim;
            atomicAdd(&area[r], end - lane);
        }
    }
}

// The per-pixel label consumers below take four pixels per thread and iteration, i, i + S, i + 2S,
// i + 3S (S = threads in the grid): the four labels, then the four per-root gathers, are issued
// together instead of one dependent label -> gather round trip per pixel.
#define NSC_FOR4(n)                                                                                  \
    for (size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x, S = (size_t)gridDim.x * blockDim.x; \
         i0 < (n); i0 += 4 * S)

// bwareaopen2 flatten output: 255 where lo <= area(root) < hi; count kept roots
__global__ void k_area_select(const int32_t* __restrict__ L, const int32_t* __restrict__ area, uint8_t* __restrict__ out,
                              size_t n, int lo, int hi, int* count) {
    int local = 0;
    NSC_FOR4(n) {
        int r[4], a[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) r[j] = i0 + j * S < n ? L[i0 + j * S] : -1;
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] = r[j] >= 0 ? area[r[j]] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = i0 + j * S;
            if (i >= n) continue;
            const bool keep = r[j] >= 0 && a[j] >= lo && a[j] < hi;
            out[i] = keep ? 255 : 0;
            local += (keep && (size_t)r[j] == i) ? 1 : 0;
        }
    }
    if (count) block_count_add(local, count);  // `count` is the same for every thread
}

// imreconstructBinary output: marker where marker != 0, else 255 where the mask component holds a
// source, else 0.  imfill additionally ORs with the image (mode 1) using mask = 255 - img.
__global__ void k_binrecon_out(const int32_t* __restrict__ L, const uint8_t* __restrict__ flag,
                               const uint8_t* __restrict__ marker, uint8_t* __restrict__ out, size_t n) {
    NSC_FOR4(n) {
        int r[4];
        uint8_t m[4], f[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = i0 + j * S;
            r[j] = i < n ? L[i] : -1;
            m[j] = i < n ? marker[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) f[j] = r[j] >= 0 ? flag[r[j]] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (i0 + j * S < n) out[i0 + j * S] = m[j] ? m[j] : (f[j] ? 255 : 0);
    }
}

// imfill(img, seeds): img | rec, rec = (255-img) at seeds inside passable (img != 255), 255 at other
// reached pixels, 0 elsewhere
__global__ void k_imfill_out(const int32_t* __restrict__ L, const uint8_t* __restrict__ flag, const uint8_t* __restrict__ img,
                             const uint8_t* __restrict__ seeds, uint8_t* __restrict__ out, size_t n) {
    NSC_FOR4(n) {
        int r[4];
        uint8_t v[4], sd[4], f[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = i0 + j * S;
            r[j] = i < n ? L[i] : -1;
            v[j] = i < n ? img[i] : 0;
            sd[j] = i < n ? seeds[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) f[j] = r[j] >= 0 ? flag[r[j]] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint8_t rec = sd[j] ? (uint8_t)(255 - v[j]) : (f[j] ? 255 : 0);
            if (i0 + j * S < n) out[i0 + j * S] = v[j] | rec;
        }
    }
}

// imfillHoles(binary): 255 unless the pixel is passable (!= 255) and its component touches the border
__global__ void k_fillholes_out(const int32_t* __restrict__ L, const uint8_t* __restrict__ flag, uint8_t* __restrict__ out,
                                size_t n) {
    NSC_FOR4(n) {
        int r[4];
        uint8_t f[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) r[j] = i0 + j * S < n ? L[i0 + j * S] : -1;
#pragma unroll
        for (int j = 0; j < 4; ++j) f[j] = r[j] >= 0 ? flag[r[j]] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (i0 + j * S < n) out[i0 + j * S] = f[j] ? 0 : 255;
    }
}

// bwselect: bw where the component holds a seed pixel
__global__ void k_bwselect_out(const int32_t* __restrict__ L, const uint8_t* __restrict__ flag, const uint8_t* __restrict__ bw,
                               uint8_t* __restrict__ out, size_t n) {
    NSC_FOR4(n) {
        int r[4];
        uint8_t b[4], f[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const size_t i = i0 + j * S;
            r[j] = i < n ? L[i] : -1;
            b[j] = i < n ? bw[i] : 0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) f[j] = r[j] >= 0 ? flag[r[j]] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (i0 + j * S < n) out[i0 + j * S] = f[j] ? b[j] : 0;
    }
}

__global__ void k_and_nonzero(const uint8_t* a, const uint8_t* b, uint8_t* out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint8_t av = a[i], bv = b[i];  // both loads unconditional (no branch around the second)
        out[i] = (av != 0 && bv != 0) ? 255 : 0;
    }
}

// ---- canonical relabel: rank of each root in raster order ----
// A workgroup of 1024 threads covers SCAN_B = 4096 consecutive pixels, four per thread (contiguous,
// so a thread's roots are in raster order after those of the threads before it); few workgroups
// keep the scan of their counts short.
constexpr int SCAN_B = 4096;

// exclusive prefix sum of v over the workgroup (at most 16 waves); *total = the sum of all v
__device__ __forceinline__ int block_scan_excl(int v, int* total) {
    __shared__ int wsum[16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    int x = v;  // inclusive scan within the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    __syncthreads();  // wsum may still be read by a previous call
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    int before = 0, tot = 0;
    for (int k = 0; k < nw; ++k) {
        const int t = wsum[k];
        if (k < wv) before += t;
        tot += t;
    }
    if (total) *total = tot;
    return before + x - v;
}

__global__ __launch_bounds__(1024) void k_root_block_count(const int32_t* __restrict__ L, size_t n, int32_t* __restrict__ bsum) {
    const size_t i0 = (size_t)blockIdx.x * SCAN_B + 4 * threadIdx.x;
    int c = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) c += (i0 + j < n && L[i0 + j] == (int32_t)(i0 + j)) ? 1 : 0;
    const int t = block_sum(c);
    if (threadIdx.x == 0) bsum[blockIdx.x] = t;
}
// exclusive scan of nb block sums in place (one workgroup of 1024 threads, a contiguous chunk each)
__global__ __launch_bounds__(1024) void k_scan_small(int32_t* bsum, int nb, int* total) {
    const int per = (nb + 1023) / 1024, start = threadIdx.x * per;
    const int end = min(start + per, nb);
    int acc = 0;
    for (int k = start; k < end; k += 4) {  // four loads in flight
        int v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = k + j < end ? bsum[k + j] : 0;
        acc += v[0] + v[1] + v[2] + v[3];
    }
    int tot = 0;
    int run = block_scan_excl(acc, &tot);
    if (threadIdx.x == 0 && total) *total = tot;
    for (int k = start; k < end; k += 4) {
        int v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = k + j < end ? bsum[k + j] : 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (k + j < end) bsum[k + j] = run;
            run += v[j];
        }
    }
}
// rank[i] (stored in place of area scratch) for roots: 1 + number of roots before i
__global__ __launch_bounds__(1024) void k_root_rank(const int32_t* __restrict__ L, size_t n, const int32_t* __restrict__ boff,
                                                   int32_t* __restrict__ rank) {
    const size_t i0 = (size_t)blockIdx.x * SCAN_B + 4 * threadIdx.x;
    bool root[4];
    int c = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        root[j] = i0 + j < n && L[i0 + j] == (int32_t)(i0 + j);
        c += root[j] ? 1 : 0;
    }
    int k = boff[blockIdx.x] + block_scan_excl(c, nullptr);
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (root[j]) rank[i0 + j] = ++k;
}
__global__ void k_apply_rank(int32_t* L, const int32_t* __restrict__ rank, size_t n, int32_t bgval) {
    NSC_FOR4(n) {
        int r[4], v[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) r[j] = i0 + j * S < n ? L[i0 + j * S] : -1;
#pragma unroll
        for (int j = 0; j < 4; ++j) v[j] = (r[j] == bgval || r[j] < 0) ? 0 : rank[r[j]];
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (i0 + j * S < n) L[i0 + j * S] = v[j];
    }
}

__global__ void k_bbox_init(int32_t* out5, int32_t* area, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out5[k] = k + 1;
    out5[n + k] = INT_MAX;
    out5[2 * n + k] = -1;
    out5[3 * n + k] = INT_MAX;
    out5[4 * n + k] = -1;
    if (area) area[k] = 0;
}
// A wave takes 64 consecutive pixels (coalesced loads of the window and of the rows above and
// below it) and visits each run of equal labels in its window once; a row start also starts a
// run.  A run issues an extent atomic only where no pixel of its component is known to beat it:
//   y0: no pixel of the run has its label directly above;      y1: likewise below;
//   x0: the pixels above its first pixel, and left of it in the rows y-1, y and y+1, lack its label;
//   x1: likewise above its last pixel, and right of it in the rows y-1, y and y+1.
// Every skip names a pixel of the component at a smaller row (y0), larger row (y1), smaller
// column or the same column one row up (x0), larger column or the same column one row up (x1).
// The extreme row's runs have no such pixel, and the topmost pixel of the extreme column has none
// either, so each extreme is still issued; the skips drop the ~4 atomics per row of a component to
// the few rows that hold an extreme.  Area stays one atomic per run.  Labels outside 1..n are ignored.
__global__ void k_bbox(const int32_t* __restrict__ L, int w, int h, int n, int32_t* out5, int32_t* area) {
    const size_t total = (size_t)w * h;
    const int lane = threadIdx.x & 63;
    const size_t wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const size_t nwave = ((size_t)gridDim.x * blockDim.x) >> 6;
    constexpr int NONE = INT_MIN;  // never a label in 1..n
    for (size_t base = wave * 64; base < total; base += nwave * 64) {
        const size_t i = base + lane;
        const bool valid = i < total;
        // position: the window's first pixel once per wave, then at most a few row wraps
        int y = (int)(base / w), x = (int)(base - (size_t)y * w) + lane;
        while (x >= w) {
            x -= w;
            ++y;
        }
        // the pixel and its neighbours above and below, and for the window's first / last lane the
        // same three one column further out (same row only; other lanes re-read their own): clamped
        // (always valid) addresses, all six loads issued together, out-of-image values replaced
        // afterwards
        const bool has_l = x > 0, has_r = x < w - 1;
        const bool has_u = valid && y > 0, has_d = valid && y < h - 1;
        const size_t ic = valid ? i : total - 1;
        const bool edge = valid && ((lane == 0 && has_l) || (lane == 63 && has_r));
        const size_t ie = edge ? (lane == 0 ? ic - 1 : ic + 1) : ic;
        const int r0 = L[ic], u0 = L[has_u ? ic - w : ic], d0 = L[has_d ? ic + w : ic];
        const int e0 = L[ie], eu0 = L[has_u ? ie - w : ie], ed0 = L[has_d ? ie + w : ie];
        const int r = valid ? r0 : -1;
        const int ru = has_u ? u0 : NONE, rd = has_d ? d0 : NONE;
        const int ev = edge ? e0 : NONE, evu = (edge && has_u) ? eu0 : NONE, evd = (edge && has_d) ? ed0 : NONE;
        // left and right neighbours: from the adjacent lanes, or the edge lanes' own loads
        int lr = __shfl_up(r, 1), lu = __shfl_up(ru, 1), ld = __shfl_up(rd, 1);
        int rr = __shfl_down(r, 1), ruu = __shfl_down(ru, 1), rdd = __shfl_down(rd, 1);
        if (lane == 0) {
            lr = ev; lu = evu; ld = evd;
        }
        if (lane == 63) {
            rr = ev; ruu = evu; rdd = evd;
        }
        const bool up = valid && ru == r, down = valid && rd == r;
        const bool x1_beaten = valid && (up || (has_r && (rr == r || ruu == r || rdd == r)));
        const unsigned long long m_up = __ballot(up), m_down = __ballot(down), m_x1 = __ballot(x1_beaten);
        const bool head = (lane == 0 || lr != r || x == 0) && valid;
        const unsigned long long heads = __ballot(head);
        if (head && r >= 1 && r <= n) {
            const unsigned long long later = lane == 63 ? 0ull : (heads >> (lane + 1));
            int end = later ? lane + __ffsll((long long)later) : 64;  // first lane after the run
            end = (int)std::min<size_t>((size_t)end, total - base);
            const int k = r - 1, len = end - lane;
            const unsigned long long run = (len == 64 ? ~0ull : ((1ull << len) - 1)) << lane;
            const bool x0_beaten = up || (has_l && (lr == r || lu == r || ld == r));
            if (!x0_beaten) atomicMin(&out5[n + k], x);
            if (!((m_x1 >> (end - 1)) & 1ull)) atomicMax(&out5[2 * n + k], x + len - 1);
            if (!(m_up & run)) atomicMin(&out5[3 * n + k], y);
            if (!(m_down & run)) atomicMax(&out5[4 * n + k], y);
            if (area) atomicAdd(&area[k], len);
        }
    }
}

inline dim3 grid1(size_t n) { return dim3(std::min<unsigned>(cdiv(n, 256), 8192u)); }

// ---- ConnComponents::boundingBox, literally (ConnComponents.cpp:241-284) -----------------------
// The reference scans labels in raster order, skips bgval, stops at the first -1 or 0 label, and
// lists the labels it met in first-seen order.  Here: the stop index, each label's first position
// (atomicMin), a rank of the first positions (the relabel kernels), then the box kernel on the
// ranks.  Labels met before the stop must lie in 1 .. w*h-1 (as ConnComponents::label makes them).
__global__ void k_bb_stop(const int32_t* __restrict__ L, int n, int32_t bgval, int* __restrict__ stop) {
    int best = INT_MAX;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t v = L[i];
        if (v != bgval && (v == -1 || v == 0)) {  // a thread's first hit in its stride is its smallest
            best = i;
            break;
        }
    }
    for (int off = 32; off > 0; off >>= 1) best = min(best, __shfl_down(best, off));
    __shared__ int part[16];
    const int wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) part[wv] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = INT_MAX;
        for (int k = 0; k < nw; ++k) b = min(b, part[k]);
        if (b != INT_MAX) atomicMin(stop, b);
    }
}
__global__ void k_bb_first(const int32_t* __restrict__ L, int n, int32_t bgval, const int* __restrict__ stop,
                           int* __restrict__ firstpos, int* __restrict__ bad) {
    const int i0 = min(*stop, n);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < i0; i += gridDim.x * blockDim.x) {
        const int32_t v = L[i];
        if (v == bgval) continue;
        if (v < 1 || v >= n) atomicOr(bad, 1);
        else atomicMin(&firstpos[v], i);
    }
}
// R[i] = i at first positions (the roots the rank kernels count), -1 elsewhere
__global__ void k_bb_roots(const int32_t* __restrict__ L, int n, int32_t bgval, const int* __restrict__ stop,
                           const int* __restrict__ firstpos, int32_t* __restrict__ R) {
    const int i0 = *stop;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t v = L[i];
        const bool ok = i < i0 && v != bgval && v >= 1 && v < n;
        const int f = ok ? firstpos[v] : -1;
        R[i] = (ok && f == i) ? i : -1;
    }
}
// canon[i] = rank of the first position of L[i] (1..count) over the scanned pixels, -1 elsewhere
__global__ void k_bb_canon(const int32_t* __restrict__ L, int n, int32_t bgval, const int* __restrict__ stop,
                           const int* __restrict__ firstpos, const int32_t* __restrict__ rank, int32_t* __restrict__ canon) {
    const int i0 = *stop;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t v = L[i];
        const bool ok = i < i0 && v != bgval && v >= 1 && v < n;
        canon[i] = ok ? rank[firstpos[v]] : -1;
    }
}
// label column: the box of rank k holds the label first seen at the k-th first position
// (rk: the rank at first positions, 0 elsewhere)
__global__ void k_bb_labels(const int32_t* __restrict__ L, int n, const int32_t* __restrict__ rk, int32_t* __restrict__ out5) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int32_t k = rk[i];
        if (k > 0) out5[k - 1] = L[i];
    }
}
__global__ void k_fill_i32(int32_t* p, int32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

// ------------------------------------------------------------------------------------------
int ccl_u8(const uint8_t* img, int32_t* labels, int w, int h, int conn, bool equal_values, int32_t bgval,
           hipStream_t s) {
    Px P{img, nullptr, 0};
    if (equal_values) return ccl_run<FG_EQ>(P, labels, w, h, conn, bgval, s);
    return ccl_run<FG_NEQ>(P, labels, w, h, conn, bgval, s);
}

int ccl_neq(const uint8_t* img, uint8_t key, int32_t* labels, int w, int h, int conn, hipStream_t s) {
    Px P{img, nullptr, key};
    return ccl_run<FG_NEQ>(P, labels, w, h, conn, -1, s);
}

int ccl_mask(const uint8_t* fg, const float* vals, int32_t* labels, int w, int h, int conn, hipStream_t s) {
    Px P{fg, vals, 0};
    return ccl_run<FG_FEQ>(P, labels, w, h, conn, -1, s);
}

// ccl_mask whose flatten also clears flags[root] for every root (a per-root flag table then
// needs no image-sized fill)
int ccl_mask_clear_root_flags(const uint8_t* fg, const float* vals, int32_t* labels, int w, int h, int conn,
                              uint8_t* flags, hipStream_t s) {
    Px P{fg, vals, 0};
    return ccl_run<FG_FEQ>(P, labels, w, h, conn, -1, s, nullptr, flags);
}

int bwareaopen_impl(const uint8_t* img, uint8_t* out, int w, int h, int lo, int hi, int conn, int* dcount,
                    int32_t* labels, int32_t* area, hipStream_t s) {
    const size_t n = (size_t)w * h;
    // equal-value labelling whose flatten clears the roots' area entries (the only ones read)
    NSC_CHECK(ccl_run<FG_EQ>(Px{img, nullptr, 0}, labels, w, h, conn, -1, s, area));
    hipLaunchKernelGGL(k_area, grid1(n), dim3(256), 0, s, labels, area, n);
    if (dcount) NSC_HIP(hipMemsetAsync(dcount, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_area_select, grid_counted(n), dim3(256), 0, s, labels, area, out, n, lo, hi, dcount);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

// component selection: passable = (img != key), sources = src != 0 (or the border when src null)
static int select_components(const uint8_t* img, uint8_t key, const uint8_t* src, int32_t* labels, uint8_t* flag,
                             int w, int h, int conn, hipStream_t s) {
    // labelling that clears the roots' flags (the only entries read or set) in its tile pass and
    // sets them while flattening: flag[root] = 1 for components holding a source
    return ccl_run<FG_NEQ>(Px{img, nullptr, key}, labels, w, h, conn, -1, s, nullptr, nullptr, flag, src);
}

int fill_holes_impl(const uint8_t* img, uint8_t* out, int w, int h, int conn, int32_t* labels, uint8_t* flag,
                    hipStream_t s) {
    NSC_CHECK(select_components(img, 255, nullptr, labels, flag, w, h, conn, s));
    hipLaunchKernelGGL(k_fillholes_out, grid1((size_t)w * h), dim3(256), 0, s, labels, flag, out, (size_t)w * h);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

int bwselect_impl(const uint8_t* bw, const uint8_t* seeds, uint8_t* out, int w, int h, int conn, int32_t* labels,
                  uint8_t* flag, hipStream_t s) {
    NSC_CHECK(select_components(bw, 0, seeds, labels, flag, w, h, conn, s));
    hipLaunchKernelGGL(k_bwselect_out, grid1((size_t)w * h), dim3(256), 0, s, labels, flag, bw, out, (size_t)w * h);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

int relabel_impl(int32_t* labels, int w, int h, int32_t bgval, int* dcount, int32_t* rank, hipStream_t s) {
    const size_t n = (size_t)w * h;
    const int nb = (int)cdiv(n, SCAN_B);
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(int32_t) * (nb + 1), s));
    int32_t* bsum = sc.as<int32_t>();
    hipLaunchKernelGGL(k_root_block_count, dim3(nb), dim3(SCAN_B / 4), 0, s, labels, n, bsum);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, bsum, nb, dcount);
    hipLaunchKernelGGL(k_root_rank, dim3(nb), dim3(SCAN_B / 4), 0, s, labels, n, bsum, rank);
    hipLaunchKernelGGL(k_apply_rank, grid1(n), dim3(256), 0, s, labels, rank, n, bgval);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

// the number of components of a labelling (its roots: label == own index), without relabelling
int root_count_impl(const int32_t* labels, int w, int h, int* dcount, hipStream_t s) {
    const size_t n = (size_t)w * h;
    const int nb = (int)cdiv(n, SCAN_B);
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(int32_t) * (nb + 1), s));
    int32_t* bsum = sc.as<int32_t>();
    hipLaunchKernelGGL(k_root_block_count, dim3(nb), dim3(SCAN_B / 4), 0, s, labels, n, bsum);
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, bsum, nb, dcount);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

int bbox_impl(const int32_t* canon, int w, int h, int n, int32_t* out5, int32_t* area, hipStream_t s);

// ConnComponents::boundingBox(w, h, labels, bgval, compcount) with the reference's result block
// (5 x compcount, column stride compcount) in out5 when compcount <= cap.  Scratch: firstpos, R,
// rank (w*h ints each) and cnt (4 ints).  Synchronises the stream (compcount is a host value).
int bbox_literal_impl(const int32_t* L, int w, int h, int32_t bgval, int32_t* out5, int cap, int* compcount,
                      int32_t* firstpos, int32_t* R, int32_t* rank, int* cnt, hipStream_t s) {
    const int n = w * h;
    hipLaunchKernelGGL(k_fill_i32, grid1(n), dim3(256), 0, s, firstpos, INT_MAX, (size_t)n);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, cnt, 0, (size_t)4);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, cnt, n, (size_t)1);  // stop = n: no stop
    hipLaunchKernelGGL(k_bb_stop, grid1(n), dim3(256), 0, s, L, n, bgval, &cnt[0]);
    hipLaunchKernelGGL(k_bb_first, grid1(n), dim3(256), 0, s, L, n, bgval, &cnt[0], firstpos, &cnt[2]);
    hipLaunchKernelGGL(k_bb_roots, grid1(n), dim3(256), 0, s, L, n, bgval, &cnt[0], firstpos, R);
    // ranks of the first positions (relabel's root rank: 1 + first positions before it)
    NSC_CHECK(relabel_impl(R, w, h, -1, &cnt[1], rank, s));
    int hc[4] = {0, 0, 0, 0};
    NSC_HIP(hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    if (hc[2]) {
        set_error("boundingBox: a label before the first background pixel lies outside 1..w*h-1");
        return NSC_ERR_ARG;
    }
    *compcount = hc[1];
    if (hc[1] == 0 || !out5 || hc[1] > cap) return NSC_OK;
    // R now holds the rank at each first position (0 elsewhere); rank[] becomes the rank image
    hipLaunchKernelGGL(k_bb_canon, grid1(n), dim3(256), 0, s, L, n, bgval, &cnt[0], firstpos, R, rank);
    NSC_CHECK(bbox_impl(rank, w, h, hc[1], out5, nullptr, s));
    hipLaunchKernelGGL(k_bb_labels, grid1(n), dim3(256), 0, s, L, n, R, out5);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

int bbox_impl(const int32_t* canon, int w, int h, int n, int32_t* out5, int32_t* area, hipStream_t s) {
    if (n <= 0) return NSC_OK;
    hipLaunchKernelGGL(k_bbox_init, dim3(cdiv(n, 256)), dim3(256), 0, s, out5, area, n);
    hipLaunchKernelGGL(k_bbox, grid1((size_t)w * h), dim3(256), 0, s, canon, w, h, n, out5, area);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

}  // namespace nsc

using namespace nsc;

extern "C" {

int nsc_cc_label_u8(const uint8_t* img, int32_t* labels, int w, int h, int32_t bgval, int conn, void* stream) {
    NSC_ARG(img && labels && w > 0 && h > 0 && valid_conn(conn), "cc_label_u8");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    return ccl_u8(img, labels, w, h, conn, true, bgval, as_stream(stream));
}

int nsc_bwlabel2_u8(const uint8_t* img, int32_t* labels, int w, int h, int conn, void* stream) {
    return nsc_cc_label_u8(img, labels, w, h, -1, conn, stream);
}

int nsc_cc_relabel(int32_t* labels, int w, int h, int32_t bgval, int* count, void* stream) {
    NSC_ARG(labels && w > 0 && h > 0, "cc_relabel");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(int32_t) * ((size_t)w * h + 1), s));
    int32_t* rank = sc.as<int32_t>();
    int* dcount = reinterpret_cast<int*>(rank + (size_t)w * h);
    NSC_CHECK(relabel_impl(labels, w, h, bgval, dcount, rank, s));
    if (count) {
        NSC_HIP(hipMemcpyAsync(count, dcount, sizeof(int), hipMemcpyDeviceToHost, s));
        NSC_HIP(stream_sync(s));
    }
    return NSC_OK;
}

int nsc_bwareaopen2_u8(const uint8_t* img, uint8_t* out, int w, int h, int minSize, int maxSize, int conn, int* count,
                       void* stream) {
    NSC_ARG(img && out && w > 0 && h > 0 && valid_conn(conn), "bwareaopen2_u8");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    const size_t n = (size_t)w * h;
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(int32_t) * (2 * n + 1), s));
    int32_t* labels = sc.as<int32_t>();
    int32_t* area = labels + n;
    int* dcount = reinterpret_cast<int*>(area + n);
    NSC_CHECK(bwareaopen_impl(img, out, w, h, minSize, maxSize, conn, dcount, labels, area, s));
    if (count) {
        NSC_HIP(hipMemcpyAsync(count, dcount, sizeof(int), hipMemcpyDeviceToHost, s));
        NSC_HIP(stream_sync(s));
    }
    return NSC_OK;
}

int nsc_imreconstruct_binary_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h, int conn,
                                void* stream) {
    NSC_ARG(marker && mask && out && w > 0 && h > 0 && valid_conn(conn), "imreconstruct_binary_u8");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    const size_t n = (size_t)w * h;
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(int32_t) * n + 2 * n, s));
    int32_t* labels = sc.as<int32_t>();
    uint8_t* flag = reinterpret_cast<uint8_t*>(labels + n);
    uint8_t* src = flag + n;
    hipLaunchKernelGGL(k_and_nonzero, grid1(n), dim3(256), 0, s, marker, mask, src, n);
    NSC_CHECK(select_components(mask, 0, src, labels, flag, w, h, conn, s));
    hipLaunchKernelGGL(k_binrecon_out, grid1(n), dim3(256), 0, s, labels, flag, marker, out, n);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

int nsc_imfill_u8(const uint8_t* img, const uint8_t* seeds, uint8_t* out, int w, int h, int conn, void* stream) {
    NSC_ARG(img && seeds && out && w > 0 && h > 0 && valid_conn(conn), "imfill_u8");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    const size_t n = (size_t)w * h;
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(int32_t) * n + n, s));
    int32_t* labels = sc.as<int32_t>();
    uint8_t* flag = reinterpret_cast<uint8_t*>(labels + n);
    NSC_CHECK(select_components(img, 255, seeds, labels, flag, w, h, conn, s));
    hipLaunchKernelGGL(k_imfill_out, grid1(n), dim3(256), 0, s, labels, flag, img, seeds, out, n);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

int nsc_imfill_holes_u8(const uint8_t* img, uint8_t* out, int w, int h, int conn, void* stream) {
    NSC_ARG(img && out && w > 0 && h > 0 && valid_conn(conn), "imfill_holes_u8");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    const size_t n = (size_t)w * h;
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(int32_t) * n + n, s));
    int32_t* labels = sc.as<int32_t>();
    return fill_holes_impl(img, out, w, h, conn, labels, reinterpret_cast<uint8_t*>(labels + n), s);
}

int nsc_bwselect_u8(const uint8_t* bw, const uint8_t* seeds, uint8_t* out, int w, int h, int conn, void* stream) {
    NSC_ARG(bw && seeds && out && w > 0 && h > 0 && valid_conn(conn), "bwselect_u8");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    const size_t n = (size_t)w * h;
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(int32_t) * n + n, s));
    int32_t* labels = sc.as<int32_t>();
    return bwselect_impl(bw, seeds, out, w, h, conn, labels, reinterpret_cast<uint8_t*>(labels + n), s);
}

int nsc_cc_bounding_box(const int32_t* labels, int w, int h, int32_t bgval, int32_t* out5, int cap, int* compcount,
                        void* stream) {
    NSC_ARG(labels && compcount && w > 0 && h > 0 && cap >= 0, "cc_bounding_box");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    const size_t n = (size_t)w * h;
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(int32_t) * (3 * n + 4), s));
    int32_t* firstpos = sc.as<int32_t>();
    return bbox_literal_impl(labels, w, h, bgval, out5, cap, compcount, firstpos, firstpos + n, firstpos + 2 * n,
                             reinterpret_cast<int*>(firstpos + 3 * n), s);
}

int nsc_bounding_boxes(const int32_t* canon_labels, int w, int h, int n, int32_t* out5, int32_t* area, void* stream) {
    NSC_ARG(canon_labels && out5 && w > 0 && h > 0 && n >= 0, "bounding_boxes");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    return bbox_impl(canon_labels, w, h, n, out5, area, as_stream(stream));
}

}  // extern "C"
