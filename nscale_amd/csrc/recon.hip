// Grayscale morphological reconstruction by dilation (nscale::imreconstruct<T>,
// MorphologicOperations.cpp:1730-1903) for uint8, int32 and float, 4- or 8-connectivity.
//
// The result is the unique reconstruction of min(marker, mask) under mask on the image padded
// with a constant-0 frame (the reference pads both inputs with 0, :1736-1739), so any schedule
// that reaches the fixpoint gives identical values.  Schedule used here:
//
//  * The image is cut into 62x62 owned tiles; one wave64 processes a 64x64 window = the tile plus
//    a 1-pixel frozen halo.  Frozen pixels (halo, and anything outside the image, which holds the
//    frame value 0) get mask = value, so the sweep arithmetic leaves them unchanged without
//    special cases.
//  * Inside the window the wave alternates direction sweeps until nothing changes:
//      - column state (lane = column, 64 values in VGPRs): top-down and bottom-up sweeps,
//        out = min(max(out, up-left, up, up-right), mask); the diagonal neighbours come from
//        the adjacent lanes by full-wave DPP shifts;
//      - row state (lane = row, after a transpose through LDS): left-right and right-left
//        sweeps, out = min(max(out, left, up-left, down-left), mask) (the diagonal neighbours only
//        with 8-connectivity; right instead of left for the right-left sweep).
//    The two sweeps of each state run as two interleaved chains updating the values in place.
//    A fixpoint of these four sweeps satisfies every 8- (or 4-) neighbour constraint.
//  * Pass 0 runs every tile once, reading marker and mask (halo = min(marker, mask)) and writing
//    the output: 3 bytes/px for uint8, the algorithmic minimum.  A tile whose owned border
//    changed marks the neighbours whose halo it feeds; later passes re-run only marked tiles
//    (reading out + mask) until no tile is marked.  The host reads the last pass's list length
//    every few passes; outside the server a short bridge of passes is queued behind each read so
//    the GPU keeps working during the round trip.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace nsc {
namespace {

// minimum waves per SIMD the window kernels are compiled for (bounds their registers)
#ifndef RECON_WAVES
#define RECON_WAVES 2
#endif

constexpr int TW = 64;  // window width (= wave size)
constexpr int TH = 64;  // window height
constexpr int OW = TW - 2;
constexpr int OH = TH - 2;

template <typename T> struct RegOf { using type = T; };
template <> struct RegOf<uint8_t> { using type = uint32_t; };

enum { CH_TOP = 1, CH_BOT = 2, CH_LEFT = 4, CH_RIGHT = 8 };

__device__ __forceinline__ void mark_tile(int t, unsigned next_pass, unsigned* stamp, int* next_list, int* next_count) {
    unsigned old = atomicMax(&stamp[t], next_pass);
    if (old < next_pass) {
        int k = atomicAdd(next_count, 1);
        next_list[k] = t;
    }
}

// Threads 0..7 of the calling workgroup mark one neighbour each (all 8, or the 4 edge neighbours
// when edges_only) whose halo the changed borders in b feed; the returning atomics of the
// different neighbours overlap instead of running one after another.
__device__ __forceinline__ void mark_neighbours(int t, unsigned b, int cx, int cy, int nx_max, int ny_max, int base,
                                                bool edges_only, unsigned next_pass, unsigned* stamp, int* next_list,
                                                int* next_count) {
    if (t >= 8) return;
    const int k = t < 4 ? t : t + 1;  // 3 x 3 neighbourhood without its centre
    const int dx = k % 3 - 1, dy = k / 3 - 1;
    if (edges_only && dx != 0 && dy != 0) return;
    const bool need = (dy == -1 && (b & CH_TOP)) || (dy == 1 && (b & CH_BOT)) || (dx == -1 && (b & CH_LEFT)) ||
                      (dx == 1 && (b & CH_RIGHT));
    const int nx = cx + dx, ny = cy + dy;
    if (need && nx >= 0 && nx < nx_max && ny >= 0 && ny < ny_max)
        mark_tile(base + ny * nx_max + nx, next_pass, stamp, next_list, next_count);
}

// Bitmap form of mark_neighbours for the work-queue drain: sets the dirty bit of each neighbour
// whose halo the changed borders in b feed (global tile id = base + tile index).
__device__ __forceinline__ void mark_neighbours_bits(int t, unsigned b, int cx, int cy, int nx_max, int ny_max, int base,
                                                     bool edges_only, unsigned long long* dirty) {
    if (t >= 8) return;
    const int k = t < 4 ? t : t + 1;
    const int dx = k % 3 - 1, dy = k / 3 - 1;
    if (edges_only && dx != 0 && dy != 0) return;
    const bool need = (dy == -1 && (b & CH_TOP)) || (dy == 1 && (b & CH_BOT)) || (dx == -1 && (b & CH_LEFT)) ||
                      (dx == 1 && (b & CH_RIGHT));
    const int nx = cx + dx, ny = cy + dy;
    if (need && nx >= 0 && nx < nx_max && ny >= 0 && ny < ny_max) {
        const int g = base + ny * nx_max + nx;
        atomicOr(&dirty[g >> 6], 1ull << (g & 63));
    }
}

// one image of a reconstruction launch: its planes, size, connectivity and first global tile id
template <typename T>
struct RImg {
    const T* marker;
    const T* mask;
    T* out;
    int w, h, tiles_x, tiles_y, conn, base;
    int rgx, rgy;  // region grid (G x G windows per region) of the launch; base counts regions
};

template <typename T>
struct RShared {
    using R = typename RegOf<T>::type;
    // the window's mask (frozen pixels: mask = value) is staged here only to be read back
    // transposed into registers; the sweeps' transposes then reuse the same LDS (one wave per
    // workgroup, so the staging reads precede the first transpose write in program order)
    union {
        T lm[TH][TW + 4];
        T tr[TH][TW + 1];
    };
    R init_l[TH], init_r[TH];
    unsigned chg_bits;
};

// ---- per-step arithmetic of the window sweeps ------------------------------------------------
// Values only rise (out starts at min(marker, mask) <= mask), so min(max(v, n), mk) = med3(v, n, mk)
// for integer types.  Frozen pixels have mk = v, hence med3(v, n, v) = v whatever n holds; that is
// why the neighbour maxima may read zero-filled DPP lanes (lanes 0 and 63 are always frozen).
template <typename R>
__device__ __forceinline__ R clamp_up(R v, R n, R mk) { return smin(smax(v, n), mk); }
template <>
__device__ __forceinline__ int32_t clamp_up(int32_t v, int32_t n, int32_t mk) {
    int32_t r;
    asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(n), "v"(mk));
    return r;
}
// max of a value and its left and right lane neighbours (8-connectivity); lanes whose DPP source
// is outside the wave read 0 (only frozen lanes, see above)
template <typename R>
__device__ __forceinline__ R nbr3(R pc) {
    const R l = __builtin_bit_cast(R, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, pc), 0x138, 0xF, 0xF, true));
    const R r = __builtin_bit_cast(R, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, pc), 0x130, 0xF, 0xF, true));
    return smax(pc, smax(l, r));
}
template <>
__device__ __forceinline__ uint32_t nbr3(uint32_t pc) {
    const uint32_t l = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pc, 0x138, 0xF, 0xF, true);
    const uint32_t r = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pc, 0x130, 0xF, 0xF, true);
    return max(max(l, pc), r);
}
template <>
__device__ __forceinline__ int32_t nbr3(int32_t pc) {
    const int32_t l = __builtin_amdgcn_update_dpp(0, pc, 0x138, 0xF, 0xF, true);
    const int32_t r = __builtin_amdgcn_update_dpp(0, pc, 0x130, 0xF, 0xF, true);
    return max(max(l, pc), r);
}

// Mask of a window in both orientations, in registers, so the sweeps never touch LDS for it:
// column state (mc: the lane's column, filled while the window loads) and row state (mr: the
// lane's row, read back transposed from the LDS staging copy).  uint8: packed four to a register
// (rows / columns 4j..4j+3).  Four-byte types: one value per register (192 registers with the
// window itself, which keeps two waves per SIMD with the staging buffer shared by the transposes).
template <typename T>
struct WinMask {
    using R = typename RegOf<T>::type;
    const T (*lm)[TW + 4];
    int lane;
    R mc[TH], mr[TW];
    __device__ __forceinline__ void build() {
#pragma unroll
        for (int x = 0; x < TW; ++x) mr[x] = (R)lm[lane][x];
    }
    __device__ void opaque() {}
    __device__ __forceinline__ R clamp_col(R v, R n, int y) const { return clamp_up(v, n, mc[y]); }
    __device__ __forceinline__ R clamp_row(R v, R n, int x) const { return clamp_up(v, n, mr[x]); }
    static __device__ __forceinline__ void fence(int) {}
};
template <>
struct WinMask<uint8_t> {
    const uint8_t (*lm)[TW + 4];
    int lane;
    uint32_t mc[TH / 4], mr[TW / 4];
    // mc is filled while the window loads; mr comes from the LDS copy (row `lane`, 4-byte words)
    __device__ __forceinline__ void build() {
#pragma unroll
        for (int j = 0; j < TW / 4; ++j) mr[j] = *reinterpret_cast<const uint32_t*>(&lm[lane][4 * j]);
    }
    // called at the top of every round: hides the packed words from loop-invariant code motion,
    // which would otherwise unpack all 124 masks once before the loop and spill them
    __device__ __forceinline__ void opaque() {
#pragma unroll
        for (int j = 0; j < TH / 4; ++j) asm volatile("" : "+v"(mc[j]), "+v"(mr[j]));
    }
    // min(max(v, n), mask byte): the byte extraction folds into the min as an SDWA operand select
    __device__ __forceinline__ uint32_t clamp_col(uint32_t v, uint32_t n, int y) const {
        return min(max(v, n), (mc[y >> 2] >> (8 * (y & 3))) & 0xFFu);
    }
    __device__ __forceinline__ uint32_t clamp_row(uint32_t v, uint32_t n, int x) const {
        return min(max(v, n), (mr[x >> 2] >> (8 * (x & 3))) & 0xFFu);
    }
    static __device__ __forceinline__ void fence(int) {}
};

// one round of the four direction sweeps over a window held in v (column state on entry and exit).
// The down and up sweeps run as two interleaved chains, each updating v in place: step k moves the
// down chain to row k and the up chain to row 63 - k, so the two dependency chains overlap in the
// pipeline instead of one waiting on the other.  Every update is min(max(v, neighbours), mask) of
// current values, a valid relaxation whatever the other chain has written, and past the middle each
// chain reads what the other has already raised.  Rows likewise with the left and right sweeps.
// (uint8: whether anything changed is decided by the caller from column sums)
template <typename T, int CONN>
__device__ __forceinline__ void sweep_round(typename RegOf<T>::type (&v)[TH], WinMask<T>& M, T (*tr)[TW + 1],
                                            int lane) {
    using R = typename RegOf<T>::type;
    M.opaque();
    {
        R pd = v[0], pu = v[TH - 1];
#pragma unroll
        for (int k = 1; k < TH - 1; ++k) {
            const int yd = k, yu = TH - 1 - k;
            const R nd = M.clamp_col(v[yd], (CONN == 8) ? nbr3(pd) : pd, yd);
            v[yd] = nd;
            pd = nd;
            const R nu = M.clamp_col(v[yu], (CONN == 8) ? nbr3(pu) : pu, yu);
            v[yu] = nu;
            pu = nu;
            M.fence(yd);
        }
    }
    // ---- transpose to row state (lane = row) ----
#pragma unroll
    for (int y = 0; y < TH; ++y) tr[y][lane] = (T)v[y];
    __syncthreads();
#pragma unroll
    for (int x = 0; x < TW; ++x) v[x] = (R)tr[lane][x];
    M.opaque();
    {
        // with 8-connectivity the row sweeps take the diagonal neighbours too (the rows above and
        // below are the adjacent lanes), so a path that wanders by one row at a time while it runs
        // along a row is carried in the same sweep instead of waiting for the next round
        R pl = v[0], pr = v[TW - 1];
#pragma unroll
        for (int k = 1; k < TW - 1; ++k) {
            const int xl = k, xr = TW - 1 - k;
            const R nl = M.clamp_row(v[xl], (CONN == 8) ? nbr3(pl) : pl, xl);
            v[xl] = nl;
            pl = nl;
            const R nr = M.clamp_row(v[xr], (CONN == 8) ? nbr3(pr) : pr, xr);
            v[xr] = nr;
            pr = nr;
            M.fence(xl);
        }
    }
#pragma unroll
    for (int x = 0; x < TW; ++x) tr[lane][x] = (T)v[x];
    __syncthreads();
#pragma unroll
    for (int y = 0; y < TH; ++y) v[y] = (R)tr[y][lane];
    __syncthreads();
}

// Change detection: values only rise, so a column changed in a round iff the sum of an
// order-preserving unsigned key of its values rose.  The sweeps carry no per-step compares (the
// compiler would keep all 248 compare results in scalar registers and OR them as a tree).
__device__ __forceinline__ uint32_t order_key(uint32_t v) { return v; }
__device__ __forceinline__ uint32_t order_key(int32_t v) { return (uint32_t)v ^ 0x80000000u; }
__device__ __forceinline__ uint32_t order_key(float v) {
    const uint32_t b = __builtin_bit_cast(uint32_t, v);
    return b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);
}
template <typename R>
__device__ __forceinline__ auto col_sum(const R (&v)[TH]) {
    // uint8 values: 64 x 255 fits 32 bits; four-byte keys need 64
    using S = typename std::conditional<std::is_same<R, uint32_t>::value, uint32_t, uint64_t>::type;
    S s = 0;
#pragma unroll
    for (int y = 0; y < TH; ++y) s += (S)order_key(v[y]);
    return s;
}

// one tile of one image: load window, sweep to stability, write owned pixels, mark neighbours
template <typename T, int CONN>
__device__ __forceinline__ void recon_tile(const RImg<T>& im, int tile, int pass, RShared<T>& sh, int* __restrict__ next_list,
                                           int* __restrict__ next_count, unsigned* __restrict__ stamp,
                                           unsigned long long* dirty = nullptr) {
    using R = typename RegOf<T>::type;
    constexpr bool U8 = std::is_same<T, uint8_t>::value;
    auto& lm = sh.lm;
    auto& init_l = sh.init_l;
    auto& init_r = sh.init_r;
    auto& chg_bits = sh.chg_bits;
    const T* marker = im.marker;
    const T* __restrict__ mask = im.mask;
    T* out = im.out;
    const int w = im.w, h = im.h, tiles_x = im.tiles_x, tiles_y = im.tiles_y;
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int lane = threadIdx.x;
    const int gx = tx * OW - 1 + lane;
    const bool col_in = (gx >= 0 && gx < w);
    const bool col_frozen = (lane == 0 || lane == TW - 1 || !col_in);
    if (lane == 0) chg_bits = 0;

    WinMask<T> M;
    M.lm = lm;
    M.lane = lane;
    R v[TH];
    // every lane loads from a clamped (always valid) address so that the 128 loads of the window
    // issue back to back; pixels outside the image are replaced by the frame value afterwards
    const auto gmask = as_global(mask);
    const auto gsrc = as_global(pass == 0 ? marker : (const T*)out);
    const size_t gxc = (size_t)clampi(gx, 0, w - 1);
    // 16 rows at a time: all 32 loads of a chunk are issued before any of them is used (the
    // scheduling fences stop the compiler from sinking loads to their uses)
#pragma unroll
    for (int c = 0; c < TH; c += 16) {
        R m_ld[16], s_ld[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const size_t i = (size_t)clampi(ty * OH - 1 + c + j, 0, h - 1) * w + gxc;
            m_ld[j] = (R)gmask[i];  // unconditional: the address is valid
            s_ld[j] = (R)gsrc[i];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int y = c + j, gy = ty * OH - 1 + y;
            const bool inb = col_in && gy >= 0 && gy < h;
            const uint32_t keep = inb ? ~0u : 0u;
            const R mk = keep_bits(m_ld[j], keep);
            const R val = keep_bits(smin(s_ld[j], m_ld[j]), keep);
            const bool frozen = col_frozen || y == 0 || y == TH - 1 || !inb;
            v[y] = val;
            const R mw = frozen ? val : mk;
            lm[y][lane] = (T)mw;
            if constexpr (U8) {
                if ((y & 3) == 0) M.mc[y >> 2] = 0;
                M.mc[y >> 2] |= (uint32_t)mw << (8 * (y & 3));
            } else {
                M.mc[y] = mw;
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if (lane == 1) {
#pragma unroll
        for (int y = 0; y < TH; ++y) init_l[y] = v[y];
    }
    if (lane == TW - 2) {
#pragma unroll
        for (int y = 0; y < TH; ++y) init_r[y] = v[y];
    }
    const R init_top = v[1], init_bot = v[TH - 2];
    __syncthreads();
    M.build();

    const auto s0 = col_sum(v);
    auto sv = s0;
    for (int round = 0; round < 4096; ++round) {
        sweep_round<T, CONN>(v, M, sh.tr, lane);
        const auto s1 = col_sum(v);
        const bool wave_changed = __any(s1 != sv);
        sv = s1;
        if (!wave_changed) break;
    }
    const bool any_change = __any(sv != s0);

    // pass 0 writes every owned pixel (the output starts as min(marker, mask)); later passes
    // only rewrite tiles that changed
    if (pass != 0 && !any_change) return;

    // ---- write the owned pixels ----
    if (!col_frozen) {
        // the column pointer is made opaque so that its per-row addresses are formed here and not
        // hoisted above the sweep loop, where they would hold 128 registers
        auto gcol = as_global(out) + gx;
        asm volatile("" : "+v"(gcol));
#pragma unroll
        for (int y = 1; y < TH - 1; ++y) {
            const int gy = ty * OH - 1 + y;
            if (gy < h) gcol[(size_t)gy * w] = (T)v[y];
        }
    }
    // ---- owned-border changes -> mark neighbours ----
    unsigned bits = 0;
    if (lane >= 1 && lane <= TW - 2) {
        if (v[1] != init_top) bits |= CH_TOP;
        if (v[TH - 2] != init_bot) bits |= CH_BOT;
    }
    if (lane == 1) {
#pragma unroll
        for (int y = 1; y < TH - 1; ++y)
            if (v[y] != init_l[y]) bits |= CH_LEFT;
    }
    if (lane == TW - 2) {
#pragma unroll
        for (int y = 1; y < TH - 1; ++y)
            if (v[y] != init_r[y]) bits |= CH_RIGHT;
    }
    if (dirty) __threadfence();  // the owned pixels are visible before any neighbour is marked
    if (bits) atomicOr(&chg_bits, bits);
    __syncthreads();
    if (dirty)
        mark_neighbours_bits(lane, chg_bits, tx, ty, tiles_x, tiles_y, im.base, false, dirty);
    else
        mark_neighbours(lane, chg_bits, tx, ty, tiles_x, tiles_y, im.base, false, (unsigned)pass + 1, stamp, next_list,
                        next_count);
}

// Pass 0 runs every tile of every image; pass k > 0 runs the tiles listed by pass k - 1.  Blocks
// stride over their work; a batch's images are found through tile_img (global tile -> image).
// CONN 0 = per-image connectivity.
template <typename T, int CONN>
__device__ __forceinline__ void recon_tiles_body(RShared<T>& sh, const RImg<T>* __restrict__ imgs, const RImg<T>& one,
                                                 const int* __restrict__ tile_img, int ntiles, const int* __restrict__ list,
                                                 const int* __restrict__ list_count, int* __restrict__ next_list,
                                                 int* __restrict__ next_count, int* __restrict__ clear_count,
                                                 unsigned* __restrict__ stamp, int pass,
                                                 unsigned long long* __restrict__ dirty) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && clear_count) *clear_count = 0;  // the counter of pass + 2 (ring)
    const int n = (pass == 0) ? ntiles : *list_count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int g = (pass == 0) ? k : list[k];
        const RImg<T> im = imgs ? imgs[tile_img[g]] : one;
        const int local = g - im.base;
        if (CONN == 8 || (CONN == 0 && im.conn == 8))
            recon_tile<T, 8>(im, local, pass, sh, next_list, next_count, stamp, dirty);
        else
            recon_tile<T, 4>(im, local, pass, sh, next_list, next_count, stamp, dirty);
        __syncthreads();  // the shared window is reused by the next tile
    }
}

template <typename T, int CONN>
__global__ __launch_bounds__(64, RECON_WAVES) void recon_tiles(const RImg<T>* __restrict__ imgs, RImg<T> one, const int* __restrict__ tile_img,
                                                  int ntiles, const int* __restrict__ list, const int* __restrict__ list_count,
                                                  int* __restrict__ next_list, int* __restrict__ next_count,
                                                  int* __restrict__ clear_count, unsigned* __restrict__ stamp, int pass,
                                                  unsigned long long* __restrict__ dirty) {
    __shared__ RShared<T> sh;
    recon_tiles_body<T, CONN>(sh, imgs, one, tile_img, ntiles, list, list_count, next_list, next_count, clear_count, stamp,
                              pass, dirty);
}

// ---- the residual wavefront as a device-side work queue ----------------------------------------
// After pass 0 (every tile once, marking the neighbours whose halo changed in the `dirty` bitmap),
// one launch drains the bitmap: a wave claims a dirty tile that no other wave is solving (the
// `busy` bitmap makes a tile's solves exclusive, so a tile's pixels only ever rise), clears its
// dirty bit, re-solves it from the current output and marks the neighbours it changed.  No pass
// structure, no host round trip, no grid-wide barrier: a tile re-runs as soon as a neighbour
// feeds it.  The claim is an acquire (the fence after it discards stale L1 lines), the marking a
// release (recon_tile fences its stores first).  A wave leaves when a scan finds no dirty bit
// while no tile is being solved (a wave that solved a tile always scans again before leaving, so
// the last marks are drained), or after `spin_limit` empty scans (harmless: others drain).  The
// host then checks that the bitmap is empty.
// One wave looks for a dirty tile that no other wave is solving and claims it: returns its id
// (the same in every lane) or -1; `seen` tells whether any dirty bit was found.  The running
// count is raised before a claim attempt, so a claimed tile is always counted.
__device__ __forceinline__ int drain_claim(int lane, int nwords, int start, unsigned long long* dirty,
                                           unsigned long long* busy, int* running, bool& seen) {
    int got = -1;
    seen = false;
    for (int k0 = 0; k0 < nwords && got < 0; k0 += 64) {
        const int k = k0 + lane;
        int wi = start + k;
        if (wi >= nwords) wi -= nwords;
        unsigned long long d = 0;
        if (k < nwords) d = __hip_atomic_load(&dirty[wi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned long long cand = __ballot(d != 0);
        seen |= cand != 0;
        while (cand && got < 0) {
            const int src = __ffsll((long long)cand) - 1;
            cand &= cand - 1;
            int mine = -1;
            if (lane == src) {
                atomicAdd(running, 1);
                unsigned long long bits = d;
                while (bits) {
                    const int b = __ffsll((long long)bits) - 1;
                    bits &= bits - 1;
                    const unsigned long long m = 1ull << b;
                    if (atomicOr(&busy[wi], m) & m) continue;  // another wave is solving it
                    if (atomicAnd(&dirty[wi], ~m) & m) {
                        mine = wi * 64 + b;
                        break;
                    }
                    atomicAnd(&busy[wi], ~m);
                }
                if (mine < 0) atomicSub(running, 1);
            }
            got = __shfl(mine, src);
        }
    }
    return got;
}

// end of a claimed tile's solve (one lane): its pixels and marks are out, it may be claimed again
__device__ __forceinline__ void drain_release(int got, unsigned long long* busy, int* running) {
    __threadfence();
    atomicAnd(&busy[got >> 6], ~(1ull << (got & 63)));
    atomicSub(running, 1);
}

template <typename T, int CONN>
__global__ __launch_bounds__(64, RECON_WAVES) void recon_drain(const RImg<T>* __restrict__ imgs, RImg<T> one,
                                                               const int* __restrict__ tile_img, int nwords,
                                                               unsigned long long* dirty, unsigned long long* busy,
                                                               int* running, int spin_limit) {
    __shared__ RShared<T> sh;
    const int lane = threadIdx.x;
    int idle = 0;
    const int start = (int)((blockIdx.x * 97u) % (unsigned)nwords);
    for (;;) {
        bool seen = false;
        const int got = drain_claim(lane, nwords, start, dirty, busy, running, seen);
        if (got < 0) {
            const int r = __hip_atomic_load(running, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((!seen && r == 0) || ++idle > spin_limit) return;
            __builtin_amdgcn_s_sleep(4);
            continue;
        }
        idle = 0;
        __threadfence();  // acquire: the neighbours' latest pixels, not stale L1 lines
        const RImg<T> im = imgs ? imgs[tile_img[got]] : one;
        const int local = got - im.base;
        if (CONN == 8 || (CONN == 0 && im.conn == 8))
            recon_tile<T, 8>(im, local, 1, sh, nullptr, nullptr, nullptr, dirty);
        else
            recon_tile<T, 4>(im, local, 1, sh, nullptr, nullptr, nullptr, dirty);
        __syncthreads();  // the shared window is reused by the next tile
        if (lane == 0) drain_release(got, busy, running);
    }
}

// 1 in *flag if any bit of the bitmap is set
__global__ void k_any_bits(const unsigned long long* __restrict__ bits, int nwords, int* flag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += gridDim.x * blockDim.x)
        if (bits[i]) atomicOr(flag, 1);
}

// ---- regions of G x G windows per workgroup --------------------------------------------
// One wave per window as above; the windows of a region exchange their shared halos through LDS
// after every round (so information crosses G windows per pass instead of one), only the region's
// outer halo is frozen for the pass.  Halo refreshes take the max of the old and the published
// value (reconstruction by dilation only raises values), so a value published a round late is
// still a valid bound; the region stops when a whole round changes nothing anywhere.
template <typename T, int G>
struct RRegion {
    using R = typename RegOf<T>::type;
    T lm[G * G][TH][TW + 4];
    T tr[G * G][TH][TW + 1];
    R init_l[G * G][TH], init_r[G * G][TH];
    T xtop[G * G][TW], xbot[G * G][TW], xleft[G * G][TH], xright[G * G][TH];
    unsigned chg_bits;
};

template <typename T, int CONN, int G>
__device__ __forceinline__ void recon_region(const RImg<T>& im, int region, int pass, RRegion<T, G>& sh,
                                             int* __restrict__ next_list, int* __restrict__ next_count,
                                             unsigned* __restrict__ stamp) {
    using R = typename RegOf<T>::type;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wx = wv % G, wy = wv / G;
    auto& lm = sh.lm[wv];
    auto& tr = sh.tr[wv];
    const T* marker = im.marker;
    const T* __restrict__ mask = im.mask;
    T* out = im.out;
    const int w = im.w, h = im.h;
    const int rx = region % im.rgx, ry = region / im.rgx;
    const int tx = rx * G + wx, ty = ry * G + wy;
    const bool valid = tx < im.tiles_x && ty < im.tiles_y;
    const bool has_l = valid && wx > 0, has_u = valid && wy > 0;
    const bool has_r = valid && wx < G - 1 && tx + 1 < im.tiles_x, has_d = valid && wy < G - 1 && ty + 1 < im.tiles_y;
    const int gx = tx * OW - 1 + lane;
    const bool col_in = (gx >= 0 && gx < w);
    const bool col_frozen = (lane == 0 || lane == TW - 1 || !col_in);
    if (threadIdx.x == 0) sh.chg_bits = 0;

    R v[TH];
    R init_top = 0, init_bot = 0;
    if (valid) {
        // 16 rows at a time from clamped (always valid) addresses, all loads of a chunk before any
        // use (see recon_tile)
        const auto gmask = as_global(mask);
        const auto gsrc = as_global(pass == 0 ? marker : (const T*)out);
        const size_t gxc = (size_t)clampi(gx, 0, w - 1);
#pragma unroll
        for (int c = 0; c < TH; c += 16) {
            R m_ld[16], s_ld[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const size_t i = (size_t)clampi(ty * OH - 1 + c + j, 0, h - 1) * w + gxc;
                m_ld[j] = (R)gmask[i];
                s_ld[j] = (R)gsrc[i];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int y = c + j, gy = ty * OH - 1 + y;
                const bool inb = col_in && gy >= 0 && gy < h;
                const uint32_t keep = inb ? ~0u : 0u;
                const R mk = keep_bits(m_ld[j], keep);
                const R val = keep_bits(smin(s_ld[j], m_ld[j]), keep);
                const bool frozen = col_frozen || y == 0 || y == TH - 1 || !inb;
                v[y] = val;
                lm[y][lane] = (T)(frozen ? val : mk);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (lane == 1) {
#pragma unroll
            for (int y = 0; y < TH; ++y) { sh.init_l[wv][y] = v[y]; sh.xleft[wv][y] = (T)v[y]; }
        }
        if (lane == TW - 2) {
#pragma unroll
            for (int y = 0; y < TH; ++y) { sh.init_r[wv][y] = v[y]; sh.xright[wv][y] = (T)v[y]; }
        }
        init_top = v[1];
        init_bot = v[TH - 2];
        sh.xtop[wv][lane] = (T)v[1];
        sh.xbot[wv][lane] = (T)v[TH - 2];
    }
    __syncthreads();

    bool owned_change = false;
    for (int round = 0; round < 65536; ++round) {
        bool changed = false, swept = false;
        if (valid) {
            // shared top/bottom halos from the siblings' last published rows (column layout)
            if (has_u) {
                const R n = (R)sh.xbot[wv - G][lane];
                if (v[0] < n) { v[0] = n; lm[0][lane] = (T)n; changed = true; }
            }
            if (has_d) {
                const R n = (R)sh.xtop[wv + G][lane];
                if (v[TH - 1] < n) { v[TH - 1] = n; lm[TH - 1][lane] = (T)n; changed = true; }
            }
            {
                R pc = v[0];
                R pl = from_left(pc, pc), pr = from_right(pc, pc);
#pragma unroll
                for (int y = 1; y < TH - 1; ++y) {
                    R up = pc;
                    if (CONN == 8) up = smax(up, smax(pl, pr));
                    const R mk = (R)lm[y][lane];
                    const R nv = smin(smax(v[y], up), mk);
                    swept |= (nv != v[y]);
                    v[y] = nv;
                    pc = nv;
                    if (CONN == 8) { pl = from_left(pc, pc); pr = from_right(pc, pc); }
                }
            }
            {
                R pc = v[TH - 1];
                R pl = from_left(pc, pc), pr = from_right(pc, pc);
#pragma unroll
                for (int y = TH - 2; y >= 1; --y) {
                    R dn = pc;
                    if (CONN == 8) dn = smax(dn, smax(pl, pr));
                    const R mk = (R)lm[y][lane];
                    const R nv = smin(smax(v[y], dn), mk);
                    swept |= (nv != v[y]);
                    v[y] = nv;
                    pc = nv;
                    if (CONN == 8) { pl = from_left(pc, pc); pr = from_right(pc, pc); }
                }
            }
#pragma unroll
            for (int y = 0; y < TH; ++y) tr[y][lane] = (T)v[y];
        }
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int x = 0; x < TW; ++x) v[x] = (R)tr[lane][x];
            // shared left/right halos (row layout: lane = row)
            if (has_l) {
                const R n = (R)sh.xright[wv - 1][lane];
                if (v[0] < n) { v[0] = n; lm[lane][0] = (T)n; changed = true; }
            }
            if (has_r) {
                const R n = (R)sh.xleft[wv + 1][lane];
                if (v[TW - 1] < n) { v[TW - 1] = n; lm[lane][TW - 1] = (T)n; changed = true; }
            }
            // row sweeps, with the diagonal neighbours (adjacent lanes) under 8-connectivity
            R prev = v[0];
            R ql = from_left(prev, prev), qr = from_right(prev, prev);
#pragma unroll
            for (int x = 1; x < TW - 1; ++x) {
                R lt = prev;
                if (CONN == 8) lt = smax(lt, smax(ql, qr));
                const R mk = (R)lm[lane][x];
                const R nv = smin(smax(v[x], lt), mk);
                swept |= (nv != v[x]);
                v[x] = nv;
                prev = nv;
                if (CONN == 8) { ql = from_left(prev, prev); qr = from_right(prev, prev); }
            }
            prev = v[TW - 1];
            ql = from_left(prev, prev);
            qr = from_right(prev, prev);
#pragma unroll
            for (int x = TW - 2; x >= 1; --x) {
                R rt = prev;
                if (CONN == 8) rt = smax(rt, smax(ql, qr));
                const R mk = (R)lm[lane][x];
                const R nv = smin(smax(v[x], rt), mk);
                swept |= (nv != v[x]);
                v[x] = nv;
                prev = nv;
                if (CONN == 8) { ql = from_left(prev, prev); qr = from_right(prev, prev); }
            }
            sh.xleft[wv][lane] = (T)v[1];
            sh.xright[wv][lane] = (T)v[TW - 2];
#pragma unroll
            for (int x = 0; x < TW; ++x) tr[lane][x] = (T)v[x];
        }
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int y = 0; y < TH; ++y) v[y] = (R)tr[y][lane];
            sh.xtop[wv][lane] = (T)v[1];
            sh.xbot[wv][lane] = (T)v[TH - 2];
        }
        owned_change |= swept;
        if (!__syncthreads_or((changed || swept) ? 1 : 0)) break;
    }

    const bool wave_owned = __any(owned_change);
    if (valid && (pass == 0 || wave_owned) && !col_frozen) {
#pragma unroll
        for (int y = 1; y < TH - 1; ++y) {
            const int gy = ty * OH - 1 + y;
            if (gy < h) out[(size_t)gy * w + gx] = (T)v[y];
        }
    }
    // region-border changes -> mark neighbouring regions
    unsigned bits = 0;
    if (valid) {
        const bool edge_l = wx == 0, edge_r = wx == G - 1 || tx == im.tiles_x - 1;
        const bool edge_u = wy == 0, edge_d = wy == G - 1 || ty == im.tiles_y - 1;
        if (lane >= 1 && lane <= TW - 2) {
            if (edge_u && v[1] != init_top) bits |= CH_TOP;
            if (edge_d && v[TH - 2] != init_bot) bits |= CH_BOT;
        }
        if (lane == 1 && edge_l) {
#pragma unroll
            for (int y = 1; y < TH - 1; ++y)
                if (v[y] != sh.init_l[wv][y]) bits |= CH_LEFT;
        }
        if (lane == TW - 2 && edge_r) {
#pragma unroll
            for (int y = 1; y < TH - 1; ++y)
                if (v[y] != sh.init_r[wv][y]) bits |= CH_RIGHT;
        }
    }
    if (bits) atomicOr(&sh.chg_bits, bits);
    __syncthreads();
    mark_neighbours(threadIdx.x, sh.chg_bits, rx, ry, im.rgx, im.rgy, im.base, false, (unsigned)pass + 1, stamp,
                    next_list, next_count);
}

template <typename T, int CONN, int G>
__global__ __launch_bounds__(64 * G * G) void recon_regions(const RImg<T>* __restrict__ imgs, RImg<T> one,
                                                            const int* __restrict__ tile_img, int ntiles,
                                                            const int* __restrict__ list, const int* __restrict__ list_count,
                                                            int* __restrict__ next_list, int* __restrict__ next_count,
                                                            int* __restrict__ clear_count, unsigned* __restrict__ stamp,
                                                            int pass) {
    __shared__ RRegion<T, G> sh;
    if (blockIdx.x == 0 && threadIdx.x == 0) *clear_count = 0;  // the counter of pass + 2 (ring)
    const int n = (pass == 0) ? ntiles : *list_count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int g = (pass == 0) ? k : list[k];
        const RImg<T> im = imgs ? imgs[tile_img[g]] : one;
        const int local = g - im.base;
        if (CONN == 8 || (CONN == 0 && im.conn == 8))
            recon_region<T, 8, G>(im, local, pass, sh, next_list, next_count, stamp);
        else
            recon_region<T, 4, G>(im, local, pass, sh, next_list, next_count, stamp);
        __syncthreads();
    }
}

}  // namespace

// Passes over a set of images until no tile is marked.  imgs_dev = nullptr: the single image `one`.
template <typename T, int CONN, int G>
static void launch_pass(int grid, hipStream_t s, const RImg<T>* imgs_dev, const RImg<T>& one, const int* tile_img, int nt,
                        const int* lst, const int* cnt, int* nl, int* nc, int* clr, unsigned* stamp, int pass,
                        unsigned long long* dirty = nullptr) {
    if (G == 1)
        hipLaunchKernelGGL((recon_tiles<T, CONN>), dim3(grid), dim3(64), 0, s, imgs_dev, one, tile_img, nt, lst, cnt, nl, nc,
                           clr, stamp, pass, dirty);
    else
        hipLaunchKernelGGL((recon_regions<T, CONN, G>), dim3(grid), dim3(64 * G * G), 0, s, imgs_dev, one, tile_img, nt, lst,
                           cnt, nl, nc, clr, stamp, pass);
}

template <typename T, int G>
static void launch_conn(int conn, int grid, hipStream_t s, const RImg<T>* imgs_dev, const RImg<T>& one, const int* tile_img,
                        int nt, const int* lst, const int* cnt, int* nl, int* nc, int* clr, unsigned* stamp, int pass) {
    if (conn == 8) launch_pass<T, 8, G>(grid, s, imgs_dev, one, tile_img, nt, lst, cnt, nl, nc, clr, stamp, pass);
    else if (conn == 4) launch_pass<T, 4, G>(grid, s, imgs_dev, one, tile_img, nt, lst, cnt, nl, nc, clr, stamp, pass);
    else launch_pass<T, 0, G>(grid, s, imgs_dev, one, tile_img, nt, lst, cnt, nl, nc, clr, stamp, pass);
}

// largest region side the workgroup's LDS holds for the element type
template <typename T> constexpr int max_region() { return sizeof(T) == 1 ? 4 : 2; }

// Schedule of the per-pass driver, per element size (0: u8, 1: 4-byte): region side (0 = the
// measured default) and the number of passes launched between host checks.  Per calling thread:
// the autotuner (nsc_recon_autotune) sets the calling thread's schedule only, so concurrent callers
// never change each other's.
struct ReconTune {
    int region = 0, batch = 4;
    int drain = -1;  // 1: device-side work queue for one-window regions, 0: passes, -1: NSC_RECON_DRAIN
};
static thread_local ReconTune tl_tune[2];
template <typename T> static ReconTune& tune() { return tl_tune[sizeof(T) == 1 ? 0 : 1]; }

// Per-pass counters live in a ring of kRing entries: pass p reads entry (p-1) % kRing (its list
// length), adds to entry p % kRing, and its block 0 clears entry (p+1) % kRing for pass p + 1,
// so no buffer grows with the number of passes.
constexpr int kRing = 4;

// Upper bound on the passes of a call: a pass that leaves a tile marked moved the front by at
// least one pixel along some propagation path, and every path is shorter than the pixel count.
static long long pass_limit(long long pixels) { return pixels + 8; }

// passes queued behind a host check's count read (about the round trip's length in empty passes)
constexpr int kPassBridge = 4;

// Device buffers of a reconstruction call for nt tiles (regions), carved from one block: the drain's
// dirty and busy bitmaps and its counters, the passes' stamps and counter ring (all of these start
// at zero: one clear covers them), then the passes' two tile lists.
struct ReconBufs {
    unsigned long long *dirty, *busy;
    int* flags;  // [0] tiles being solved by the drain, [1] bits left after the drain
    unsigned* stamp;
    int *counts, *listA, *listB;
    size_t zero_bytes;  // leading bytes to clear
};
static size_t recon_words(int nt) { return ((size_t)nt + 63) / 64; }
static size_t recon_scratch_bytes(int nt) {
    return 2 * sizeof(unsigned long long) * recon_words(nt) + 4 * sizeof(int) + sizeof(unsigned) * (size_t)nt +
           sizeof(int) * (kRing + 2 * (size_t)nt);
}
static ReconBufs recon_carve(void* base, int nt) {
    ReconBufs B;
    const size_t nw = recon_words(nt);
    B.dirty = static_cast<unsigned long long*>(base);
    B.busy = B.dirty + nw;
    B.flags = reinterpret_cast<int*>(B.busy + nw);
    B.stamp = reinterpret_cast<unsigned*>(B.flags + 4);
    B.counts = reinterpret_cast<int*>(B.stamp + nt);
    B.listA = B.counts + kRing;
    B.listB = B.listA + nt;
    B.zero_bytes = (size_t)(reinterpret_cast<uint8_t*>(B.listA) - static_cast<uint8_t*>(base));
    return B;
}

// clears the leading words of the call's buffers (a kernel, not a runtime fill, so a captured graph
// of a call replays it faithfully)
__global__ void k_zero_words(uint32_t* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}
static void recon_clear(const ReconBufs& B, hipStream_t s) {
    const size_t words = B.zero_bytes / 4;
    hipLaunchKernelGGL(k_zero_words, dim3(std::min<unsigned>(cdiv(words, 256), 1024u)), dim3(256), 0, s,
                       reinterpret_cast<uint32_t*>(B.dirty), words);
}

// waves of the drain kernel that fit on the device at once (the grid of the drain launch)
template <typename T>
static int drain_grid() {
    static int g = [] {
        int dev = 0, per_cu = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, recon_drain<T, 8>, 64, 0) != hipSuccess) {
            (void)hipGetLastError();
            return 2048;
        }
        return std::max(1, per_cu) * std::max(1, cus);
    }();
    return g;
}

// Schedule of the residual wavefront: host-launched passes (default) or the device-side work
// queue (NSC_RECON_DRAIN=1; one-window regions only).  The queue is exact but measured far slower
// on MI355X (imreconstruct<uchar> of the 4096^2 ellipse tile: 15.5 ms against 0.44 ms with passes):
// every claim, mark and bitmap scan is a device-scope atomic or coherent load, which the eight
// XCDs' separate L2 caches cannot serve locally, and thousands of waves contend for a few bitmap
// lines, while a pass boundary makes the same hand-off once per launch.
static bool drain_env() {
    static bool v = [] {
        const char* e = getenv("NSC_RECON_DRAIN");
        return e && e[0] == '1';
    }();
    return v;
}
template <typename T>
static bool drain_enabled() {
    const int d = tune<T>().drain;
    return d < 0 ? drain_env() : d != 0;
}

template <typename T>
static int recon_passes(const RImg<T>* imgs_dev, const RImg<T>& one, const int* tile_img, int nt, long long pixels, int conn,
                        int G, const ReconBufs& B, hipStream_t s) {
    const long long max_pass = pass_limit(pixels);
    const int grid0 = std::min(nt, 1 << 16);
    unsigned* stamp = B.stamp;
    int* counts = B.counts;
    auto launch = [&](int pass, const int* lst, const int* cnt, int* nl, int* nc, int* clr) {
        // a serving thread's later passes run on at most 1024 workgroups that loop over the list:
        // beside the other slots' kernels, a pass's idle workgroups only wait for CUs
        const int grid = pass == 0 || !thread_serving() ? grid0 : std::min(grid0, 1024);
        switch (G) {
            case 2: launch_conn<T, 2>(conn, grid, s, imgs_dev, one, tile_img, nt, lst, cnt, nl, nc, clr, stamp, pass); break;
            case 3: launch_conn<T, (max_region<T>() >= 3 ? 3 : 2)>(conn, grid, s, imgs_dev, one, tile_img, nt, lst, cnt, nl,
                                                                    nc, clr, stamp, pass); break;
            case 4: launch_conn<T, max_region<T>()>(conn, grid, s, imgs_dev, one, tile_img, nt, lst, cnt, nl, nc, clr, stamp,
                                                    pass);
                    break;
            default: launch_conn<T, 1>(conn, grid, s, imgs_dev, one, tile_img, nt, lst, cnt, nl, nc, clr, stamp, pass);
        }
    };
    launch(0, nullptr, nullptr, B.listA, &counts[0], &counts[1]);
    NSC_HIP(hipGetLastError());
    long long pass = 1;  // next pass to queue
    auto queue_to = [&](long long last) {
        for (; pass <= last && pass <= max_pass; ++pass) {
            int* cur = (pass & 1) ? B.listA : B.listB;
            int* nxt = (pass & 1) ? B.listB : B.listA;
            launch((int)std::min<long long>(pass, INT_MAX - 1), cur, &counts[(pass - 1) % kRing], nxt, &counts[pass % kRing],
                   &counts[(pass + 1) % kRing]);
        }
    };
    // passes between host checks: the schedule's batch b, then 2b, then 4b from there on (a host
    // round trip costs about as much as ten empty passes, and calls that run past 3b passes
    // usually run far past it); a bridge of kPassBridge passes is queued behind each count read
    int batch = std::max(1, tune<T>().batch);
    for (int check = 0;; ++check, batch = check <= 2 ? 2 * batch : batch) {
        queue_to(pass - 1 + batch);
        NSC_HIP(hipGetLastError());
        const long long last = pass - 1;  // the check reads the list length pass `last` produced
        int c = 0;
        if (const int* hc = count_read_async(&counts[last % kRing], s)) {
            queue_to(last + kPassBridge);
            NSC_HIP(hipGetLastError());
            NSC_HIP(count_wait());
            c = *hc;
        } else {
            NSC_HIP(hipMemcpyAsync(&c, &counts[last % kRing], sizeof(int), hipMemcpyDeviceToHost, s));
            NSC_HIP(stream_sync(s));
        }
        if (c == 0) break;
        if (last >= max_pass) {
            set_error("imreconstruct: no convergence after %lld passes", max_pass);
            return NSC_ERR_LIMIT;
        }
    }
    return NSC_OK;
}

// pass 0 marking into the dirty bitmap, then one drain launch; the bitmap must end empty
template <typename T>
static int recon_drain_run(const RImg<T>* imgs_dev, const RImg<T>& one, const int* tile_img, int nt, int conn,
                           const ReconBufs& B, hipStream_t s, bool* leftover) {
    const int grid0 = std::min(nt, 1 << 16);
    const int nw = (int)recon_words(nt);
    if (conn == 8) launch_pass<T, 8, 1>(grid0, s, imgs_dev, one, tile_img, nt, nullptr, nullptr, nullptr, nullptr, nullptr,
                                        nullptr, 0, B.dirty);
    else if (conn == 4) launch_pass<T, 4, 1>(grid0, s, imgs_dev, one, tile_img, nt, nullptr, nullptr, nullptr, nullptr,
                                             nullptr, nullptr, 0, B.dirty);
    else launch_pass<T, 0, 1>(grid0, s, imgs_dev, one, tile_img, nt, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                              0, B.dirty);
    const int grid = std::min(nt, drain_grid<T>());
    constexpr int kSpin = 4096;
    if (conn == 8)
        hipLaunchKernelGGL((recon_drain<T, 8>), dim3(grid), dim3(64), 0, s, imgs_dev, one, tile_img, nw, B.dirty, B.busy,
                           &B.flags[0], kSpin);
    else if (conn == 4)
        hipLaunchKernelGGL((recon_drain<T, 4>), dim3(grid), dim3(64), 0, s, imgs_dev, one, tile_img, nw, B.dirty, B.busy,
                           &B.flags[0], kSpin);
    else
        hipLaunchKernelGGL((recon_drain<T, 0>), dim3(grid), dim3(64), 0, s, imgs_dev, one, tile_img, nw, B.dirty, B.busy,
                           &B.flags[0], kSpin);
    hipLaunchKernelGGL(k_any_bits, dim3(std::min<unsigned>(cdiv(nw, 256), 256u)), dim3(256), 0, s, B.dirty, nw, &B.flags[1]);
    NSC_HIP(hipGetLastError());
    int left = 0;
    NSC_HIP(hipMemcpyAsync(&left, &B.flags[1], sizeof(int), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    *leftover = left != 0;
    return NSC_OK;
}

static std::atomic<long long> g_drain_redo{0};  // drains that left marked tiles (reporting)

// the whole reconstruction on cleared buffers B
template <typename T>
static int recon_run(const RImg<T>* imgs_dev, const RImg<T>& one, const int* tile_img, int nt, long long pixels, int conn,
                     int G, const ReconBufs& B, hipStream_t s) {
    if (G == 1 && drain_enabled<T>()) {
        bool leftover = false;
        NSC_CHECK(recon_drain_run<T>(imgs_dev, one, tile_img, nt, conn, B, s, &leftover));
        if (!leftover) return NSC_OK;
        // a drain that ended with marked tiles (its waves gave up): the passes finish from scratch
        ++g_drain_redo;
        recon_clear(B, s);
    }
    return recon_passes<T>(imgs_dev, one, tile_img, nt, pixels, conn, G, B, s);
}

template <typename T>
static int recon_run_scratch(const RImg<T>* imgs_dev, const RImg<T>& one, const int* tile_img, int nt, long long pixels,
                             int conn, int G, hipStream_t s) {
    Scratch sc;
    NSC_CHECK(sc.alloc(recon_scratch_bytes(nt), s));
    const ReconBufs B = recon_carve(sc.p, nt);
    recon_clear(B, s);
    return recon_run<T>(imgs_dev, one, tile_img, nt, pixels, conn, G, B, s);
}

// region side (windows per workgroup side) for element type T: NSC_RECON_REGION overrides
template <typename T>
static int region_side() {
    static int env = [] {
        const char* e = getenv("NSC_RECON_REGION");
        return e ? std::max(1, std::min(atoi(e), max_region<T>())) : 0;
    }();
    if (env) return env;
    // measured default 1: 2 ties with it on the 4096^2 workloads, 3 and 4 are slower (a region's
    // rounds wait on its slowest window)
    const int t = tune<T>().region;
    return t ? std::min(t, max_region<T>()) : 1;
}

template <typename T>
static void set_regions(RImg<T>& r, int G) {
    r.rgx = (int)cdiv(r.tiles_x, G);
    r.rgy = (int)cdiv(r.tiles_y, G);
}

template <typename T>
int recon_impl(const T* marker, const T* mask, T* out, int w, int h, int conn, hipStream_t s) {
    RImg<T> one{marker, mask, out, w, h, (int)cdiv(w, OW), (int)cdiv(h, OH), conn, 0, 0, 0};
    const int G = region_side<T>();
    set_regions(one, G);
    return recon_run_scratch<T>(nullptr, one, nullptr, one.rgx * one.rgy, (long long)w * h, conn, G, s);
}

// A batch of images in one launch per pass: the image table and the tile -> image map live in
// stream-ordered scratch.  conn: 4 or 8 for the whole batch, 0 = per image.
template <typename T>
int recon_batched_impl(const nsc_recon_desc* d, int n, hipStream_t s) {
    std::vector<RImg<T>> imgs(n);
    long long nt = 0, pixels = 0;
    int conn_all = d[0].conn;
    const int G = region_side<T>();
    for (int i = 0; i < n; ++i) {
        NSC_ARG(d[i].marker && d[i].mask && d[i].out && d[i].w > 0 && d[i].h > 0 && (d[i].conn == 4 || d[i].conn == 8),
                "imreconstruct batched: descriptor");
        RImg<T>& r = imgs[i];
        r.marker = static_cast<const T*>(d[i].marker);
        r.mask = static_cast<const T*>(d[i].mask);
        r.out = static_cast<T*>(d[i].out);
        r.w = d[i].w;
        r.h = d[i].h;
        r.tiles_x = (int)cdiv(r.w, OW);
        r.tiles_y = (int)cdiv(r.h, OH);
        r.conn = d[i].conn;
        r.base = (int)nt;
        set_regions(r, G);
        nt += (long long)r.rgx * r.rgy;
        pixels += (long long)r.w * r.h;
        if (d[i].conn != conn_all) conn_all = 0;
    }
    NSC_ARG(nt < (1ll << 30), "imreconstruct batched: too many tiles");
    std::vector<int> tile_img((size_t)nt);
    for (int i = 0; i < n; ++i)
        std::fill(tile_img.begin() + imgs[i].base, tile_img.begin() + imgs[i].base + imgs[i].rgx * imgs[i].rgy, i);
    Scratch tab;
    const size_t ib = ((sizeof(RImg<T>) * n + 255) / 256) * 256;
    NSC_CHECK(tab.alloc(ib + sizeof(int) * (size_t)nt, s));
    RImg<T>* imgs_dev = tab.as<RImg<T>>();
    int* tile_dev = reinterpret_cast<int*>(static_cast<char*>(tab.p) + ib);
    NSC_HIP(hipMemcpyAsync(imgs_dev, imgs.data(), sizeof(RImg<T>) * n, hipMemcpyHostToDevice, s));
    NSC_HIP(hipMemcpyAsync(tile_dev, tile_img.data(), sizeof(int) * (size_t)nt, hipMemcpyHostToDevice, s));
    const int rc = recon_run_scratch<T>(imgs_dev, imgs[0], tile_dev, (int)nt, pixels, conn_all, G, s);
    // the host tables must outlive the copies (pageable copies complete before returning; be explicit)
    NSC_HIP(stream_sync(s));
    return rc;
}

template int recon_impl<uint8_t>(const uint8_t*, const uint8_t*, uint8_t*, int, int, int, hipStream_t);
template int recon_impl<int32_t>(const int32_t*, const int32_t*, int32_t*, int, int, int, hipStream_t);
template int recon_impl<float>(const float*, const float*, float*, int, int, int, hipStream_t);

// Times every (region side, passes per host check) candidate on the caller's batch -- the
// outputs are written each time, identically -- and keeps the fastest for later calls of the
// same element type.  timings (optional, host) receives one entry per candidate in ms.
template <typename T>
static int recon_autotune(const nsc_recon_desc* d, int n, hipStream_t s, int* region, int* batch, float* timings) {
    // candidates: (region side, passes per host check, work queue)
    const int cand[5][3] = {{1, 4, 0}, {1, 8, 0}, {2, 4, 0}, {2, 8, 0}, {1, 4, 1}};
    ReconTune& t = tune<T>();
    const ReconTune t0 = t;
    hipEvent_t e0, e1;
    NSC_HIP(hipEventCreate(&e0));
    NSC_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int bi = 0;
    for (int k = 0; k < 5; ++k) {
        t.region = cand[k][0];
        t.batch = cand[k][1];
        t.drain = cand[k][2];
        int rc = recon_batched_impl<T>(d, n, s);  // warm
        float ms = 0.f;
        if (rc == NSC_OK) {
            (void)hipEventRecord(e0, s);
            rc = recon_batched_impl<T>(d, n, s);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
        }
        if (rc != NSC_OK) {
            t = t0;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
        if (timings) timings[k] = ms;
        if (ms < best) { best = ms; bi = k; }
    }
    t.region = cand[bi][0];
    t.batch = cand[bi][1];
    t.drain = cand[bi][2];
    if (region) *region = t.region;
    if (batch) *batch = t.drain ? 0 : t.batch;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return NSC_OK;
}

template int recon_batched_impl<uint8_t>(const nsc_recon_desc*, int, hipStream_t);
template int recon_batched_impl<float>(const nsc_recon_desc*, int, hipStream_t);

namespace {

// ---- a batched reconstruction as a replayable hipGraph ----------------------------------------
// The pass structure of recon_passes with the pass number read from a device counter: one captured
// sequence of K window-pass launches and the counter's advance replays as passes base .. base+K-1,
// base+K .. base+2K-1, and so on.  A pass whose list is empty only clears its ring counter, so
// passes past convergence cost a near-empty launch; the host checks the last pass's count once per
// replay.  Every clear is a kernel (nothing a graph would record as a runtime fill).
template <typename T, int CONN>
__global__ __launch_bounds__(64, RECON_WAVES) void recon_tiles_graph(const RImg<T>* __restrict__ imgs,
                                                                     const int* __restrict__ tile_img, int ntiles,
                                                                     int* __restrict__ listA, int* __restrict__ listB,
                                                                     int* __restrict__ counts, unsigned* __restrict__ stamp,
                                                                     const int* __restrict__ base, int i) {
    __shared__ RShared<T> sh;
    const int pass = *base + i;
    const int* list = (pass & 1) ? listA : listB;
    int* next_list = (pass & 1) ? listB : listA;
    const int* list_count = pass ? &counts[(pass - 1) % kRing] : nullptr;
    recon_tiles_body<T, CONN>(sh, imgs, imgs[0], tile_img, ntiles, list, list_count, next_list, &counts[pass % kRing],
                              &counts[(pass + 1) % kRing], stamp, pass, nullptr);
}

__global__ void k_pass_advance(int* base, int k) {
    if (threadIdx.x == 0) *base += k;
}

__global__ void k_pass_reset(int* base) {
    if (threadIdx.x == 0) *base = 0;
}

}  // namespace
}  // namespace nsc

struct nsc_recon_graph {
    int is_float = 0, K = 0, nt = 0;
    long long pixels = 0;
    void* mem = nullptr;  // image table, tile -> image map, pass scratch, pass counter
    nsc::ReconBufs B{};
    int* base = nullptr;
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace nsc {
namespace {

template <typename T>
int recon_graph_build(const nsc_recon_desc* d, int n, int K, nsc_recon_graph* g) {
    std::vector<RImg<T>> imgs(n);
    long long nt = 0, pixels = 0;
    int conn_all = d[0].conn;
    for (int i = 0; i < n; ++i) {
        NSC_ARG(d[i].marker && d[i].mask && d[i].out && d[i].w > 0 && d[i].h > 0 && (d[i].conn == 4 || d[i].conn == 8),
                "recon_graph: descriptor");
        RImg<T>& r = imgs[i];
        r.marker = static_cast<const T*>(d[i].marker);
        r.mask = static_cast<const T*>(d[i].mask);
        r.out = static_cast<T*>(d[i].out);
        r.w = d[i].w;
        r.h = d[i].h;
        r.tiles_x = (int)cdiv(r.w, OW);
        r.tiles_y = (int)cdiv(r.h, OH);
        r.conn = d[i].conn;
        r.base = (int)nt;
        set_regions(r, 1);
        nt += (long long)r.rgx * r.rgy;
        pixels += (long long)r.w * r.h;
        if (d[i].conn != conn_all) conn_all = 0;
    }
    NSC_ARG(nt < (1ll << 30), "recon_graph: too many tiles");
    std::vector<int> tile_img((size_t)nt);
    for (int i = 0; i < n; ++i)
        std::fill(tile_img.begin() + imgs[i].base, tile_img.begin() + imgs[i].base + imgs[i].rgx * imgs[i].rgy, i);
    Carver m(nullptr);
    m.take<RImg<T>>(n);
    m.take<int>((size_t)nt);
    m.take<uint8_t>(recon_scratch_bytes((int)nt));
    m.take<int>(1);
    NSC_HIP(hipMalloc(&g->mem, m.size()));
    Carver c(g->mem);
    RImg<T>* imgs_dev = c.take<RImg<T>>(n);
    int* tile_dev = c.take<int>((size_t)nt);
    g->B = recon_carve(c.take<uint8_t>(recon_scratch_bytes((int)nt)), (int)nt);
    g->base = c.take<int>(1);
    NSC_HIP(hipMemcpy(imgs_dev, imgs.data(), sizeof(RImg<T>) * n, hipMemcpyHostToDevice));
    NSC_HIP(hipMemcpy(tile_dev, tile_img.data(), sizeof(int) * (size_t)nt, hipMemcpyHostToDevice));
    g->nt = (int)nt;
    g->pixels = pixels;
    g->K = K;
    NSC_HIP(hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking));
    NSC_HIP(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
    const int grid = (int)std::min<long long>(nt, 1 << 16);
    for (int i = 0; i < K; ++i) {
        if (conn_all == 8)
            hipLaunchKernelGGL((recon_tiles_graph<T, 8>), dim3(grid), dim3(TW), 0, g->cap, imgs_dev, tile_dev, (int)nt,
                               g->B.listA, g->B.listB, g->B.counts, g->B.stamp, g->base, i);
        else if (conn_all == 4)
            hipLaunchKernelGGL((recon_tiles_graph<T, 4>), dim3(grid), dim3(TW), 0, g->cap, imgs_dev, tile_dev, (int)nt,
                               g->B.listA, g->B.listB, g->B.counts, g->B.stamp, g->base, i);
        else
            hipLaunchKernelGGL((recon_tiles_graph<T, 0>), dim3(grid), dim3(TW), 0, g->cap, imgs_dev, tile_dev, (int)nt,
                               g->B.listA, g->B.listB, g->B.counts, g->B.stamp, g->base, i);
    }
    hipLaunchKernelGGL(k_pass_advance, dim3(1), dim3(64), 0, g->cap, g->base, K);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(g->cap, &g->graph);
    NSC_HIP(le);
    NSC_HIP(ce);
    NSC_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    return NSC_OK;
}

}  // namespace
}  // namespace nsc

using nsc::set_error;

extern "C" int nsc_recon_graph_destroy(nsc_recon_graph* g) {
    if (!g) return NSC_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->cap) (void)hipStreamDestroy(g->cap);
    if (g->mem) (void)hipFree(g->mem);
    delete g;
    return NSC_OK;
}

extern "C" int nsc_recon_graph_create(const nsc_recon_desc* descs, int n, int is_float, int passes_per_replay,
                                      nsc_recon_graph** out) {
    NSC_CHECK(nsc::ensure_init());
    NSC_ARG(descs && n > 0 && out, "recon_graph_create");
    NSC_ARG(passes_per_replay >= 1 && passes_per_replay <= 256, "recon_graph_create: passes per replay 1..256");
    *out = nullptr;
    auto* g = new nsc_recon_graph();
    g->is_float = is_float ? 1 : 0;
    const int rc = is_float ? nsc::recon_graph_build<float>(descs, n, passes_per_replay, g)
                            : nsc::recon_graph_build<uint8_t>(descs, n, passes_per_replay, g);
    if (rc != NSC_OK) {
        nsc_recon_graph_destroy(g);
        return rc;
    }
    *out = g;
    return NSC_OK;
}

extern "C" int nsc_recon_graph_run(nsc_recon_graph* g, int* passes, void* stream) {
    NSC_ARG(g && g->exec, "recon_graph_run");
    const hipStream_t s = nsc::as_stream(stream);
    nsc::recon_clear(g->B, s);
    hipLaunchKernelGGL(nsc::k_pass_reset, dim3(1), dim3(64), 0, s, g->base);
    NSC_HIP(hipGetLastError());
    const long long max_pass = nsc::pass_limit(g->pixels);
    for (long long done = 0;;) {
        NSC_HIP(hipGraphLaunch(g->exec, s));
        done += g->K;
        int c = 0;
        NSC_HIP(hipMemcpyAsync(&c, &g->B.counts[(done - 1) % nsc::kRing], sizeof(int), hipMemcpyDeviceToHost, s));
        NSC_HIP(nsc::stream_sync(s));
        if (c == 0) {
            if (passes) *passes = (int)std::min<long long>(done, INT_MAX);
            return NSC_OK;
        }
        if (done > max_pass) {
            set_error("recon_graph_run: no convergence after %lld passes", max_pass);
            return NSC_ERR_LIMIT;
        }
    }
}

// ---- reconstruction without allocation ---------------------------------------------------
// The pass driver's scratch (tile stamps, the counter ring, two tile lists: 12 B per tile plus the
// ring) for callers that cannot allocate lives in one file-scope device array, `scratch`, which
// every GPU holds at its own address for the life of the process.  Host code never passes the
// array's name to a kernel (in host code it is a stand-in, not a device address): the device
// address comes from hipGetSymbolAddress, once per device, with that device current.
namespace nsc {
namespace {

constexpr int kNoAllocMaxTiles = 1 << 16;  // 62 x 62 tiles: e.g. 15872 x 15872 pixels
constexpr size_t kNoAllocBytes = 2 * sizeof(unsigned long long) * (kNoAllocMaxTiles / 64) + 4 * sizeof(int) +
                                sizeof(unsigned) * kNoAllocMaxTiles + sizeof(int) * (kRing + 2 * (size_t)kNoAllocMaxTiles);
static_assert(kNoAllocMaxTiles % 64 == 0, "whole bitmap words");

__device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};

// the calling device's copy of `scratch` (looked up once per device)
int noalloc_scratch(void** out) {
    int dev = 0;
    NSC_HIP(hipGetDevice(&dev));
    NSC_ARG(dev >= 0 && dev < kMaxDevices, "imreconstruct_u8_noalloc: device index");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        NSC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return NSC_OK;
}

}  // namespace
}  // namespace nsc

extern "C" int nsc_imreconstruct_u8_noalloc(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h,
                                            int conn, void* stream) {
    using namespace nsc;
    NSC_ARG(marker && mask && out && w > 0 && h > 0 && valid_conn(conn), "imreconstruct_u8_noalloc");
    NSC_CHECK(ensure_init());
    RImg<uint8_t> one{marker, mask, out, w, h, (int)cdiv(w, OW), (int)cdiv(h, OH), conn, 0, 0, 0};
    const int G = region_side<uint8_t>();
    set_regions(one, G);
    const long long nt = (long long)one.rgx * one.rgy;
    if (nt > kNoAllocMaxTiles) {
        set_error("imreconstruct_u8_noalloc: %lld tiles of 62 x 62 pixels exceed the fixed scratch of %d tiles", nt,
                  kNoAllocMaxTiles);
        return NSC_ERR_LIMIT;
    }
    void* base = nullptr;
    NSC_CHECK(noalloc_scratch(&base));
    hipStream_t s = as_stream(stream);
    const ReconBufs B = recon_carve(base, (int)nt);  // fits: nt <= kNoAllocMaxTiles
    recon_clear(B, s);
    // the host checks read their counts synchronously here: the bridged reads would allocate a
    // pinned word on this thread's first use
    const bool bridge = set_thread_pass_bridge(false);
    const int rc = recon_run<uint8_t>(nullptr, one, nullptr, (int)nt, (long long)w * h, conn, G, B, s);
    (void)set_thread_pass_bridge(bridge);
    return rc;
}

extern "C" int nsc_recon_set_schedule(int is_float, int region, int passes, int queue) {
    NSC_ARG(region >= 0 && region <= 4 && passes >= 0 && queue >= -1 && queue <= 1, "recon_set_schedule");
    nsc::ReconTune& t = is_float ? nsc::tune<float>() : nsc::tune<uint8_t>();
    t.region = region;
    t.batch = passes ? passes : 4;
    t.drain = queue;
    return NSC_OK;
}

extern "C" int nsc_recon_autotune(const nsc_recon_desc* descs, int n, int is_float, void* stream, int* region, int* passes,
                                  float* timings) {
    NSC_ARG(descs && n > 0, "recon_autotune");
    NSC_CHECK(nsc::ensure_init());
    hipStream_t s = nsc::as_stream(stream);
    return is_float ? nsc::recon_autotune<float>(descs, n, s, region, passes, timings)
                    : nsc::recon_autotune<uint8_t>(descs, n, s, region, passes, timings);
}

// ------------------------------------------------------------------------------------------
// Geodesic (4-connected, unit step) distance inside a passable mask from source pixels with
// dist = 0: the same tile/sweep/pass schedule as the reconstruction above in the (min, +1)
// algebra.  dist holds the initial values (0 at sources, UINT_MAX elsewhere) and the result.
// ------------------------------------------------------------------------------------------
namespace nsc {
namespace {

constexpr unsigned GINF = 0xFFFFFFFFu;
// Frozen pixels of a window (window halo, outside the image, not passable) as bits, in both
// orientations: column state fc (bit y & 31 of fc[y >> 5] = pixel (y, lane)), row state fr (bit
// x & 31 of fr[x >> 5] = pixel (lane, x)).  blk = all ones on frozen pixels (value kept) and 0 on
// free ones feeds each step (geo_step).
struct GeoFlags {
    uint32_t fc[2], fr[2];
    // called at the top of every round: keeps the 124 unpacked flags out of loop-invariant code
    // motion (they would be computed once before the loop and spilled)
    __device__ __forceinline__ void opaque() { asm volatile("" : "+v"(fc[0]), "+v"(fc[1]), "+v"(fr[0]), "+v"(fr[1])); }
    __device__ __forceinline__ unsigned blk_col(int y) const { return 0u - ((fc[y >> 5] >> (y & 31)) & 1u); }
    __device__ __forceinline__ unsigned blk_row(int x) const { return 0u - ((fr[x >> 5] >> (x & 31)) & 1u); }
};

// One step along a line: min(v, prev + 1) for a passable pixel (blk = 0); a frozen pixel inside a
// line always holds GINF (outside the image, or not passable: never lowered) and keeps it, since
// med3(GINF, x, GINF) = GINF, while med3(v, x, 0) = min(v, x).  The halo lanes (lanes 0 and TW-1,
// frozen but holding finite neighbour distances) are kept out of the sweeps by the callers' lane
// test instead.
__device__ __forceinline__ unsigned geo_step(unsigned v, unsigned prev, unsigned blk) {
    unsigned r;
    asm("v_med3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(v), "v"(__builtin_elementwise_add_sat(prev, 1u)), "v"(blk));
    return r;
}

// Frozen-pixel bits of every window (the passable set is fixed for a whole geodesic call, so
// they are computed once): one uint4 per lane, {fc[0], fc[1], fr[0], fr[1]}.
__global__ __launch_bounds__(64) void geo_flags(const uint8_t* __restrict__ pass, uint4* __restrict__ flags, int w, int h,
                                                int tiles_x) {
    __shared__ uint8_t lp[TH][TW + 4];  // frozen flags, row-major (staging for the row-state bits)
    const int tile = blockIdx.x, tx = tile % tiles_x, ty = tile / tiles_x;
    const int lane = threadIdx.x;
    const int gx = tx * OW - 1 + lane;
    const bool col_in = (gx >= 0 && gx < w);
    const bool col_frozen = (lane == 0 || lane == TW - 1 || !col_in);
    const auto gpass = as_global(pass);
    const size_t gxc = (size_t)clampi(gx, 0, w - 1);
    uint32_t fc[2] = {0, 0}, fr[2] = {0, 0};
#pragma unroll
    for (int c = 0; c < TH; c += 16) {
        uint32_t p_ld[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) p_ld[j] = gpass[(size_t)clampi(ty * OH - 1 + c + j, 0, h - 1) * w + gxc];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int y = c + j, gy = ty * OH - 1 + y;
            const bool inb = col_in && gy >= 0 && gy < h;
            const bool frozen = col_frozen || y == 0 || y == TH - 1 || !inb || p_ld[j] == 0;
            fc[y >> 5] |= (frozen ? 1u : 0u) << (y & 31);
            lp[y][lane] = frozen ? 1 : 0;
        }
    }
    __syncthreads();
    // row state: the lane's row of flag bytes, four at a time; (w * 0x10204080) >> 28 gathers the
    // low bits of the four bytes (partial products never overlap, so nothing carries into them)
#pragma unroll
    for (int j = 0; j < TW / 4; ++j) {
        const uint32_t wd = *reinterpret_cast<const uint32_t*>(&lp[lane][4 * j]);
        fr[j >> 3] |= ((wd * 0x10204080u) >> 28) << (4 * (j & 7));
    }
    flags[(size_t)tile * TW + lane] = make_uint4(fc[0], fc[1], fr[0], fr[1]);
}

// Load a window's distances into v (column state) and its frozen bits.  Every lane reads clamped,
// always valid addresses, and all 64 rows are issued before any is used (the scheduling fence
// stops the compiler from sinking loads to their uses, which would wait on every row in turn).
__device__ __forceinline__ void geo_load(const uint4* __restrict__ flags, const unsigned* dist, int w, int h, int tx, int ty,
                                         int tiles_x, int lane, unsigned (&v)[TH], GeoFlags& F) {
    const int gx = tx * OW - 1 + lane;
    const bool col_in = (gx >= 0 && gx < w);
    const auto gdist = as_global(dist);
    const size_t gxc = (size_t)clampi(gx, 0, w - 1);
    const uint4 fw = flags[((size_t)ty * tiles_x + tx) * TW + lane];
#pragma unroll
    for (int y = 0; y < TH; ++y) v[y] = gdist[(size_t)clampi(ty * OH - 1 + y, 0, h - 1) * w + gxc];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int y = 0; y < TH; ++y) {
        const int gy = ty * OH - 1 + y;
        const bool inb = col_in && gy >= 0 && gy < h;
        v[y] |= inb ? 0u : GINF;
    }
    F.fc[0] = fw.x;
    F.fc[1] = fw.y;
    F.fr[0] = fw.z;
    F.fr[1] = fw.w;
}

// The sweeps do not track changes step by step (a compare per step, and the compiler keeps the
// compare results in scalar registers until it ORs them as a tree); distances only fall, so a
// column changed iff the 64-bit sum of ~v over it rose (geo_key_sum).
// Along one line the shortest path to a pixel comes straight from above or straight from below
// (a path that turns back covers pixels twice), so the downward and the upward sweep can run on
// the same input as two independent chains, interleaved for instruction-level parallelism, and
// their minimum is exactly the sequential down-then-up result.
__device__ __forceinline__ void geo_col_sweeps(unsigned (&v)[TH], GeoFlags& F) {
    F.opaque();
    const int lane = threadIdx.x & 63;
    if (lane < 1 || lane > TW - 2) return;  // halo columns: sources, never stepped
    unsigned dn[TH], up[TH];
    unsigned pd = v[0], pu = v[TH - 1];
#pragma unroll
    for (int k = 1; k < TH - 1; ++k) {
        const int yd = k, yu = TH - 1 - k;
        dn[yd] = geo_step(v[yd], pd, F.blk_col(yd));
        pd = dn[yd];
        up[yu] = geo_step(v[yu], pu, F.blk_col(yu));
        pu = up[yu];
    }
#pragma unroll
    for (int y = 1; y < TH - 1; ++y) v[y] = min(dn[y], up[y]);
}

__device__ __forceinline__ void geo_row_sweeps(unsigned (&v)[TW], GeoFlags& F) {
    F.opaque();
    const int lane = threadIdx.x & 63;
    if (lane < 1 || lane > TH - 2) return;  // halo rows: sources, never stepped
    unsigned lt[TW], rt[TW];
    unsigned pl = v[0], pr = v[TW - 1];
#pragma unroll
    for (int k = 1; k < TW - 1; ++k) {
        const int xl = k, xr = TW - 1 - k;
        lt[xl] = geo_step(v[xl], pl, F.blk_row(xl));
        pl = lt[xl];
        rt[xr] = geo_step(v[xr], pr, F.blk_row(xr));
        pr = rt[xr];
    }
#pragma unroll
    for (int x = 1; x < TW - 1; ++x) v[x] = min(lt[x], rt[x]);
}

// sum of ~v over rows [y0, y1) of the lane's column (monotone in every value as distances fall)
__device__ __forceinline__ uint64_t geo_key_sum(const unsigned (&v)[TH], int y0, int y1) {
    uint64_t s = 0;
#pragma unroll
    for (int y = 0; y < TH; ++y)
        if (y >= y0 && y < y1) s += (uint64_t)(~v[y]);
    return s;
}

// owned pixels of a window back to dist (the column pointer is opaque so that the per-row
// addresses are formed here, not hoisted above the sweep loop)
__device__ __forceinline__ void geo_store(unsigned* dist, int w, int h, int tx, int ty, int lane, const unsigned (&v)[TH]) {
    const int gx = tx * OW - 1 + lane;
    if (lane == 0 || lane == TW - 1 || gx >= w) return;
    auto gcol = as_global(dist) + gx;
    asm volatile("" : "+v"(gcol));
#pragma unroll
    for (int y = 1; y < TH - 1; ++y) {
        const int gy = ty * OH - 1 + y;
        if (gy < h) gcol[(size_t)gy * w] = v[y];
    }
}

__global__ __launch_bounds__(64, 2) void geo_tiles(const uint4* __restrict__ flags, unsigned* dist, int w, int h,
                                                             int tiles_x, int tiles_y, const int* __restrict__ list,
                                                             const int* __restrict__ list_count, int* __restrict__ next_list,
                                                             int* __restrict__ next_count, int* __restrict__ clear_count,
                                                             unsigned* __restrict__ stamp, int pass_no) {
    __shared__ unsigned tr[TH][TW + 1];
    __shared__ unsigned init_l[TH], init_r[TH];
    __shared__ unsigned chg_bits;
    if (blockIdx.x == 0 && threadIdx.x == 0) *clear_count = 0;  // the counter of pass + 2 (ring)
    int tile;
    if (pass_no == 0) tile = blockIdx.x;
    else {
        if ((int)blockIdx.x >= *list_count) return;
        tile = list[blockIdx.x];
    }
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int lane = threadIdx.x;
    if (lane == 0) chg_bits = 0;
    unsigned v[TH];
    GeoFlags F;
    geo_load(flags, dist, w, h, tx, ty, tiles_x, lane, v, F);
    if (lane == 1) {
#pragma unroll
        for (int y = 0; y < TH; ++y) init_l[y] = v[y];
    }
    if (lane == TW - 2) {
#pragma unroll
        for (int y = 0; y < TH; ++y) init_r[y] = v[y];
    }
    const unsigned init_top = v[1], init_bot = v[TH - 2];
    __syncthreads();
    const uint64_t s0 = geo_key_sum(v, 0, TH);
    uint64_t s = s0;
    for (int round = 0; round < 1 << 20; ++round) {
        geo_col_sweeps(v, F);
#pragma unroll
        for (int y = 0; y < TH; ++y) tr[y][lane] = v[y];
        __syncthreads();
#pragma unroll
        for (int x = 0; x < TW; ++x) v[x] = tr[lane][x];
        geo_row_sweeps(v, F);
#pragma unroll
        for (int x = 0; x < TW; ++x) tr[lane][x] = v[x];
        __syncthreads();
#pragma unroll
        for (int y = 0; y < TH; ++y) v[y] = tr[y][lane];
        __syncthreads();
        const uint64_t s1 = geo_key_sum(v, 0, TH);
        const bool wave_changed = __any(s1 != s);
        s = s1;
        if (!wave_changed) break;
    }
    if (!__any(s != s0)) return;
    geo_store(dist, w, h, tx, ty, lane, v);
    unsigned bits = 0;
    if (lane >= 1 && lane <= TW - 2) {
        if (v[1] != init_top) bits |= CH_TOP;
        if (v[TH - 2] != init_bot) bits |= CH_BOT;
    }
    if (lane == 1) {
#pragma unroll
        for (int y = 1; y < TH - 1; ++y)
            if (v[y] != init_l[y]) bits |= CH_LEFT;
    }
    if (lane == TW - 2) {
#pragma unroll
        for (int y = 1; y < TH - 1; ++y)
            if (v[y] != init_r[y]) bits |= CH_RIGHT;
    }
    if (bits) atomicOr(&chg_bits, bits);
    __syncthreads();
    // 4-connected: edge neighbours only
    mark_neighbours(lane, chg_bits, tx, ty, tiles_x, tiles_y, 0, true, (unsigned)pass_no + 1, stamp, next_list,
                    next_count);
}

// GX x GY windows per workgroup, shared halos exchanged through LDS after every round (as the
// reconstruction regions above, in the (min, +1) algebra: halo refreshes take the min, since
// distances only fall).
template <int GX, int GY>
struct GeoRegion {
    unsigned tr[GX * GY][TH][TW + 1];
    unsigned xtop[GX * GY][TW], xbot[GX * GY][TW], xleft[GX * GY][TH], xright[GX * GY][TH];
    unsigned chg_bits;
};

// one region of GX x GY windows: load, solve to stability with halo exchanges through LDS, store,
// mark the neighbouring regions whose halo changed (in the pass lists, or the drain's bitmap)
template <int GX, int GY>
__device__ __forceinline__ void geo_region_solve(const uint4* __restrict__ flags, unsigned* dist, int w, int h, int tiles_x,
                                                 int tiles_y, int rgx, int rgy, int region, GeoRegion<GX, GY>& sh,
                                                 int* __restrict__ next_list, int* __restrict__ next_count,
                                                 unsigned* __restrict__ stamp, int pass_no, unsigned long long* dirty) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wx = wv % GX, wy = wv / GX;
    auto& tr = sh.tr[wv];
    const int rx = region % rgx, ry = region / rgx;
    const int tx = rx * GX + wx, ty = ry * GY + wy;
    const bool valid = tx < tiles_x && ty < tiles_y;
    const bool has_l = valid && wx > 0, has_u = valid && wy > 0;
    const bool has_r = valid && wx < GX - 1 && tx + 1 < tiles_x, has_d = valid && wy < GY - 1 && ty + 1 < tiles_y;
    if (threadIdx.x == 0) sh.chg_bits = 0;
    unsigned v[TH];
    GeoFlags F;
    unsigned init_top = 0, init_bot = 0;
    if (valid) {
        geo_load(flags, dist, w, h, tx, ty, tiles_x, lane, v, F);
        if (lane == 1) {
#pragma unroll
            for (int y = 0; y < TH; ++y) sh.xleft[wv][y] = v[y];
        }
        if (lane == TW - 2) {
#pragma unroll
            for (int y = 0; y < TH; ++y) sh.xright[wv][y] = v[y];
        }
        init_top = v[1];
        init_bot = v[TH - 2];
        sh.xtop[wv][lane] = v[1];
        sh.xbot[wv][lane] = v[TH - 2];
    }
    __syncthreads();
    // owned pixels: rows 1..62 of lanes 1..62; the halo rows and lanes change only through the
    // exchanges below, which the full-column sums also see
    const uint64_t own0 = geo_key_sum(v, 1, TH - 1);
    uint64_t full = own0 + (uint64_t)(~v[0]) + (uint64_t)(~v[TH - 1]), own = own0;
    for (int round = 0; round < (1 << 20); ++round) {
        bool changed = false;
        if (valid) {
            if (has_u) {
                const unsigned nb = sh.xbot[wv - GX][lane];
                if (nb < v[0]) v[0] = nb;
            }
            if (has_d) {
                const unsigned nb = sh.xtop[wv + GX][lane];
                if (nb < v[TH - 1]) v[TH - 1] = nb;
            }
            geo_col_sweeps(v, F);
#pragma unroll
            for (int y = 0; y < TH; ++y) tr[y][lane] = v[y];
        }
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int x = 0; x < TW; ++x) v[x] = tr[lane][x];
            if (has_l) {
                const unsigned nb = sh.xright[wv - 1][lane];
                if (nb < v[0]) v[0] = nb;
            }
            if (has_r) {
                const unsigned nb = sh.xleft[wv + 1][lane];
                if (nb < v[TW - 1]) v[TW - 1] = nb;
            }
            geo_row_sweeps(v, F);
            sh.xleft[wv][lane] = v[1];
            sh.xright[wv][lane] = v[TW - 2];
#pragma unroll
            for (int x = 0; x < TW; ++x) tr[lane][x] = v[x];
        }
        __syncthreads();
        if (valid) {
#pragma unroll
            for (int y = 0; y < TH; ++y) v[y] = tr[y][lane];
            sh.xtop[wv][lane] = v[1];
            sh.xbot[wv][lane] = v[TH - 2];
            own = geo_key_sum(v, 1, TH - 1);
            const uint64_t f1 = own + (uint64_t)(~v[0]) + (uint64_t)(~v[TH - 1]);
            changed = f1 != full;
            full = f1;
        }
        if (!__syncthreads_or(changed ? 1 : 0)) break;
    }
    const bool lane_owned = lane >= 1 && lane <= TW - 2;
    if (valid && __any(lane_owned && own != own0)) geo_store(dist, w, h, tx, ty, lane, v);
    unsigned bits = 0;
    if (valid) {
        const bool edge_l = wx == 0, edge_r = wx == GX - 1 || tx == tiles_x - 1;
        const bool edge_u = wy == 0, edge_d = wy == GY - 1 || ty == tiles_y - 1;
        if (lane >= 1 && lane <= TW - 2) {
            if (edge_u && v[1] != init_top) bits |= CH_TOP;
            if (edge_d && v[TH - 2] != init_bot) bits |= CH_BOT;
        }
        // an owned edge column changed iff its key sum rose (distances only fall)
        if (lane == 1 && edge_l && own != own0) bits |= CH_LEFT;
        if (lane == TW - 2 && edge_r && own != own0) bits |= CH_RIGHT;
    }
    if (dirty) __threadfence();  // the region's pixels are visible before any neighbour is marked
    if (bits) atomicOr(&sh.chg_bits, bits);
    __syncthreads();
    // 4-connected: edge neighbours only
    if (dirty)
        mark_neighbours_bits(threadIdx.x, sh.chg_bits, rx, ry, rgx, rgy, 0, true, dirty);
    else
        mark_neighbours(threadIdx.x, sh.chg_bits, rx, ry, rgx, rgy, 0, true, (unsigned)pass_no + 1, stamp, next_list,
                        next_count);
    __syncthreads();
}

template <int GX, int GY>
__global__ __launch_bounds__(64 * GX * GY) void geo_regions(const uint4* __restrict__ flags, unsigned* dist, int w, int h,
                                                          int tiles_x, int tiles_y, int rgx, int rgy,
                                                          const int* __restrict__ list, const int* __restrict__ list_count,
                                                          int* __restrict__ next_list, int* __restrict__ next_count,
                                                          int* __restrict__ clear_count, unsigned* __restrict__ stamp,
                                                          int pass_no) {
    __shared__ GeoRegion<GX, GY> sh;
    if (blockIdx.x == 0 && threadIdx.x == 0) *clear_count = 0;  // the counter of pass + 2 (ring)
    const int n = pass_no == 0 ? rgx * rgy : *list_count;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int region = pass_no == 0 ? k : list[k];
        geo_region_solve<GX, GY>(flags, dist, w, h, tiles_x, tiles_y, rgx, rgy, region, sh, next_list, next_count, stamp,
                            pass_no, nullptr);
    }
}

// The geodesic distance as a device-side work queue (as recon_drain): every region starts dirty;
// the first wave of the workgroup claims a region and the workgroup solves it.
template <int GX, int GY>
__global__ __launch_bounds__(64 * GX * GY) void geo_drain(const uint4* __restrict__ flags, unsigned* dist, int w, int h,
                                                        int tiles_x, int tiles_y, int rgx, int rgy, int nwords,
                                                        unsigned long long* dirty, unsigned long long* busy, int* running,
                                                        int spin_limit) {
    __shared__ GeoRegion<GX, GY> sh;
    __shared__ int s_got;
    int idle = 0;
    const int start = (int)((blockIdx.x * 97u) % (unsigned)nwords);
    for (;;) {
        if (threadIdx.x < 64) {
            bool seen = false;
            int got = drain_claim(threadIdx.x, nwords, start, dirty, busy, running, seen);
            if (got < 0) {
                const int r = __hip_atomic_load(running, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                got = ((!seen && r == 0) || ++idle > spin_limit) ? -2 : -1;
            } else {
                idle = 0;
            }
            if (threadIdx.x == 0) s_got = got;
        }
        __syncthreads();
        const int got = s_got;
        __syncthreads();  // s_got is rewritten only after every thread has read it
        if (got == -2) return;
        if (got == -1) {
            __builtin_amdgcn_s_sleep(4);
            continue;
        }
        __threadfence();  // acquire: the neighbours' latest distances, not stale L1 lines
        geo_region_solve<GX, GY>(flags, dist, w, h, tiles_x, tiles_y, rgx, rgy, got, sh, nullptr, nullptr, nullptr, 1, dirty);
        if (threadIdx.x == 0) drain_release(got, busy, running);
    }
}

// The first pass's list when the only finite distance is one seed pixel (every other pixel starts
// at GINF, so no other region can change): the seed's region, or an empty list without a seed.
__global__ void k_geo_seed_list(const int* __restrict__ seedp, int w, int h, int rgx, int gx, int gy, int* list,
                                int* count, unsigned* stamp) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int sd = *seedp;
    if (sd >= 0 && (long long)sd < (long long)w * h) {
        const int r = ((sd / w) / OH / gy) * rgx + (sd % w) / OW / gx;
        list[0] = r;
        stamp[r] = 1;  // listed for pass 1, as pass 0 would mark it
        *count = 1;
    } else {
        *count = 0;
    }
}

// sets the first n bits of a bitmap (every region dirty)
__global__ void k_set_bits(unsigned long long* bits, int n) {
    const int nw = (n + 63) / 64;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += gridDim.x * blockDim.x) {
        const int left = n - 64 * i;
        bits[i] = left >= 64 ? ~0ull : ((1ull << left) - 1);
    }
}

}  // namespace

int geodist_impl(const uint8_t* pass, unsigned* dist, int w, int h, hipStream_t s, const int* seed) {
    // region shape in windows: 2 x 2, or single windows with NSC_GEO_REGION=1 (4 x 2 and 2 x 4
    // regions were measured slower on the 4096^2 chain tile: 3.2-3.4 ms against 2.7 ms, the extra
    // rounds a region needs to carry its halos across more windows outweigh the fewer passes)
    static const int GXY = [] {
        const char* e = getenv("NSC_GEO_REGION");
        return e && !strcmp(e, "1") ? 11 : 22;
    }();
    const int GX = GXY / 10, GY = GXY % 10;
    const int tiles_x = (int)cdiv(w, OW), tiles_y = (int)cdiv(h, OH);
    const int rgx = (int)cdiv(tiles_x, GX), rgy = (int)cdiv(tiles_y, GY);
    const int nt = rgx * rgy;
    const long long max_pass = pass_limit((long long)w * h);
    const int nwin = tiles_x * tiles_y;
    Scratch sc;
    NSC_CHECK(sc.alloc(sizeof(uint4) * TW * (size_t)nwin + recon_scratch_bytes(nt), s));
    uint4* flags = sc.as<uint4>();
    const ReconBufs B = recon_carve(flags + TW * (size_t)nwin, nt);
    recon_clear(B, s);
    unsigned* stamp = B.stamp;
    int* counts = B.counts;
    int* listA = B.listA;
    int* listB = B.listB;
    hipLaunchKernelGGL(geo_flags, dim3(nwin), dim3(TW), 0, s, pass, flags, w, h, tiles_x);
    if (GXY == 22 && drain_env()) {
        // every region dirty, one drain launch, then the bitmap must be empty
        const int nw = (int)recon_words(nt);
        hipLaunchKernelGGL(k_set_bits, dim3(std::min<unsigned>(cdiv(nw, 256), 256u)), dim3(256), 0, s, B.dirty, nt);
        static int grid_d = [] {
            int dev = 0, per_cu = 0, cus = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, geo_drain<2, 2>, 256, 0) != hipSuccess) {
                (void)hipGetLastError();
                return 512;
            }
            return std::max(1, per_cu) * std::max(1, cus);
        }();
        hipLaunchKernelGGL((geo_drain<2, 2>), dim3(std::min(nt, grid_d)), dim3(256), 0, s, flags, dist, w, h, tiles_x, tiles_y,
                           rgx, rgy, nw, B.dirty, B.busy, &B.flags[0], 4096);
        hipLaunchKernelGGL(k_any_bits, dim3(std::min<unsigned>(cdiv(nw, 256), 256u)), dim3(256), 0, s, B.dirty, nw,
                           &B.flags[1]);
        NSC_HIP(hipGetLastError());
        int left = 0;
        NSC_HIP(hipMemcpyAsync(&left, &B.flags[1], sizeof(int), hipMemcpyDeviceToHost, s));
        NSC_HIP(stream_sync(s));
        if (!left) return NSC_OK;
        // waves gave up with regions still marked: the passes finish from the current distances
        ++g_drain_redo;
    }
    // a serving thread's passes (lists of a few tens of regions) run on a small grid that loops
    // over longer lists: idle region workgroups would each wait for 71 KB of free LDS beside the
    // other slots' kernels
    const unsigned grid = (unsigned)std::min(nt, thread_serving() ? 128 : 1 << 16);
    auto launch = [&](int p, const int* lst, const int* cnt, int* nl, int* nc, int* clr) {
        if (GXY == 22)
            hipLaunchKernelGGL((geo_regions<2, 2>), dim3(grid), dim3(256), 0, s, flags, dist, w, h, tiles_x, tiles_y, rgx, rgy,
                               lst, cnt, nl, nc, clr, stamp, p);
        else
            hipLaunchKernelGGL(geo_tiles, dim3(nt), dim3(64), 0, s, flags, dist, w, h, tiles_x, tiles_y, lst, cnt, nl, nc,
                               clr, stamp, p);
    };
    if (seed) {
        // dist holds GINF everywhere but at *seed: pass 0 would change the seed's region only
        hipLaunchKernelGGL(k_geo_seed_list, dim3(1), dim3(64), 0, s, seed, w, h, rgx, GX, GY, listA, &counts[0], stamp);
    } else {
        launch(0, nullptr, nullptr, listA, &counts[0], &counts[1]);
    }
    long long p = 1;  // next pass to queue
    auto queue_to = [&](long long last) {
        for (; p <= last && p <= max_pass; ++p) {
            int* cur = (p & 1) ? listA : listB;
            int* nxt = (p & 1) ? listB : listA;
            launch((int)std::min<long long>(p, INT_MAX - 1), cur, &counts[(p - 1) % kRing], nxt, &counts[p % kRing],
                   &counts[(p + 1) % kRing]);
        }
    };
    // host checks after 8, 24 and 56 passes, then every 8: the front needs tens of passes to cross
    // a tile (about 60 on a 4096^2 nuclei tile), an empty pass costs a few microseconds and a host
    // round trip a few tens; a bridge of kPassBridge passes is queued behind each count read
    for (long long target = 8, batch = 8;; batch = target >= 56 ? 8 : 2 * batch, target += batch) {
        queue_to(target);
        NSC_HIP(hipGetLastError());
        const long long last = p - 1;  // the check reads the list length pass `last` produced
        int c = 0;
        if (const int* hc = count_read_async(&counts[last % kRing], s)) {
            queue_to(last + kPassBridge);
            NSC_HIP(hipGetLastError());
            NSC_HIP(count_wait());
            c = *hc;
        } else {
            NSC_HIP(hipMemcpyAsync(&c, &counts[last % kRing], sizeof(int), hipMemcpyDeviceToHost, s));
            NSC_HIP(stream_sync(s));
        }
        if (c == 0) break;
        if (last >= max_pass) {
            set_error("geodesic distance: no convergence after %lld passes", max_pass);
            return NSC_ERR_LIMIT;
        }
    }
    return NSC_OK;
}

}  // namespace nsc
