// Grey erosion / dilation with a k x k structuring element (cv::erode / cv::dilate as used by
// nscale::morphOpen, MorphologicOperations.cpp:3848-3882, and HistologicalEntities.cpp:330),
// anchor at the centre, pixels beyond the image taking a given constant.
//
// Each kernel row is a run of ones [a, b] (true for disks, crosses, rectangles); a row with holes
// is split into several runs.  Per 64x64 output tile (256 threads):
//   1. the (64+2r)^2 input window is staged in LDS (outside pixels = the given constant);
//   2. for every distinct run, the horizontal min/max over [x+a, x+b] is computed for all window
//      rows; runs are processed shortest first and a run that contains the previous one extends
//      its result by the extra taps only (disk19: 19 taps per pixel for all 5 distinct runs);
//   3. the output is the min/max over kernel rows of the run result of the matching window row.
// Traffic: 1 byte in + 1 byte out per pixel (plus the halo re-read, r/32 of the tile).
// The window and run results are held as 16-bit pixel pairs where that fits in LDS (morph2_kernel),
// else as bytes (morph_kernel).  3 x 3 elements take a direct kernel instead (morph3_kernel).
#include <algorithm>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace nsc {
namespace {

constexpr int MT = 64;       // output tile side
constexpr int MAXK = 31;     // largest kernel side
constexpr int MAXRUN = 64;   // runs in total (rows with holes contribute several)

struct MorphPlan {
    int k, r;
    int nrun;                 // distinct runs
    int ra[MAXRUN], rb[MAXRUN];
    int base[MAXRUN];         // index of the run this one extends, -1 = from scratch
    int nrow;                 // (kernel row, run) pairs
    int row_i[MAXRUN], row_run[MAXRUN];
};

template <bool ERODE>
__device__ __forceinline__ uint32_t pick(uint32_t a, uint32_t b) { return ERODE ? min(a, b) : max(a, b); }

template <bool ERODE>
__global__ __launch_bounds__(256) void morph_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, int w, int h,
                                                    MorphPlan P, uint8_t outside, int tiles_x) {
    extern __shared__ uint8_t sm[];
    const int r = P.r;
    const int WW = MT + 2 * r;           // window side
    const int WS = WW + 4;               // padded stride
    uint8_t* win = sm;                   // WW x WS
    uint8_t* H = sm + WW * WS;           // nrun x WW x MT
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int x0 = tx * MT - r, y0 = ty * MT - r;
    // window load, eight loads per thread in flight: every lane reads a clamped (always valid)
    // address and pixels beyond the image take `outside` afterwards (a load per iteration with its
    // use in the same iteration compiles to one memory round trip after another)
    const auto gin = as_global(in);
    const int WN = WW * WW;
    // k / WW by multiply-high (exact for k < 2^16; the window holds at most 94^2 pixels)
    const uint32_t wmagic = (uint32_t)((0x100000000ull + (unsigned)WW - 1) / (unsigned)WW);
    for (int k0 = 0; k0 < WN; k0 += 8 * 256) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = min(k0 + (int)threadIdx.x + 256 * j, WN - 1);
            const int ly = (int)__umulhi((uint32_t)k, wmagic), lx = k - ly * WW;
            v[j] = gin[(size_t)clampi(y0 + ly, 0, h - 1) * w + clampi(x0 + lx, 0, w - 1)];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + (int)threadIdx.x + 256 * j;
            if (k >= WN) continue;
            const int ly = (int)__umulhi((uint32_t)k, wmagic), lx = k - ly * WW;
            const int gx = x0 + lx, gy = y0 + ly;
            const uint32_t keep = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? 0xFFu : 0u;
            win[ly * WS + lx] = (uint8_t)((v[j] & keep) | ((uint32_t)outside & ~keep));
        }
    }
    __syncthreads();
    for (int d = 0; d < P.nrun; ++d) {
        const int a = P.ra[d], b = P.rb[d], bs = P.base[d];
        uint8_t* Hd = H + d * WW * MT;
        for (int k = threadIdx.x; k < WW * MT; k += blockDim.x) {
            const int x = k % MT, row = k / MT;
            const uint8_t* src = win + row * WS + x + r;
            uint32_t v;
            if (bs >= 0) {
                const int pa = P.ra[bs], pb = P.rb[bs];
                v = H[bs * WW * MT + k];
                for (int j = a; j < pa; ++j) v = pick<ERODE>(v, src[j]);
                for (int j = pb + 1; j <= b; ++j) v = pick<ERODE>(v, src[j]);
            } else {
                v = src[a];
                for (int j = a + 1; j <= b; ++j) v = pick<ERODE>(v, src[j]);
            }
            Hd[k] = (uint8_t)v;
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < MT * MT; k += blockDim.x) {
        const int x = k % MT, y = k / MT;
        const int gx = tx * MT + x, gy = ty * MT + y;
        if (gx >= w || gy >= h) continue;
        uint32_t v = ERODE ? 255u : 0u;
        for (int t = 0; t < P.nrow; ++t) {
            const int i = P.row_i[t];
            v = pick<ERODE>(v, H[P.row_run[t] * WW * MT + (y + i) * MT + x]);
        }
        out[(size_t)gy * w + gx] = (uint8_t)v;
    }
}

// Two pixels per instruction: the same three steps with the window and the run results held as
// 16-bit values, so one aligned 32-bit LDS word is a pair of neighbouring pixels and one packed
// u16 min/max serves both.  A tap at an odd offset takes the pair straddling two aligned words
// (funnel shift of the words read for the neighbouring even taps), so a run of n taps costs
// about n/2 LDS reads and n packed ops per two outputs.  Used where its LDS (twice the byte
// kernel's) fits.
__device__ __forceinline__ uint32_t lds_u32(const uint16_t* p) { return *reinterpret_cast<const uint32_t*>(p); }

template <bool ERODE>
__device__ __forceinline__ uint32_t pick2(uint32_t a, uint32_t b) {
    typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
    const u16x2 va = __builtin_bit_cast(u16x2, a), vb = __builtin_bit_cast(u16x2, b);
    return __builtin_bit_cast(uint32_t, ERODE ? __builtin_elementwise_min(va, vb) : __builtin_elementwise_max(va, vb));
}

// taps q0..q1 (pair offsets from an even pair base `p`, q0 >= 0) folded into `acc`; reads the
// aligned words at even offsets q0 & ~1 .. (q1 + 1) & ~1 only
template <bool ERODE>
__device__ __forceinline__ uint32_t taps2(uint32_t acc, const uint16_t* p, int q0, int q1) {
    int q = q0 & ~1;
    uint32_t cur = lds_u32(p + q);
    for (; q <= q1; q += 2) {
        if (q >= q0) acc = pick2<ERODE>(acc, cur);          // pixels q, q+1
        if (q + 1 <= q1) {
            const uint32_t nxt = lds_u32(p + q + 2);
            acc = pick2<ERODE>(acc, (cur >> 16) | (nxt << 16));  // pixels q+1, q+2
            cur = nxt;
        }
    }
    return acc;
}

constexpr int M2_THREADS = 512;

template <bool ERODE>
__global__ __launch_bounds__(M2_THREADS) void morph2_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, int w,
                                                            int h, MorphPlan P, uint8_t outside, int tiles_x, bool words) {
    extern __shared__ uint16_t sm16[];
    const int r = P.r;
    const int WW = MT + 2 * r;  // window side (even, so rows start on 4-byte boundaries)
    const int WS = WW;          // row stride in pixels; the last pair read is (WW-2, WW-1)
    uint16_t* win = sm16;       // WW x WS
    uint16_t* H = sm16 + WW * WS;  // nrun x WW x MT
    const int tx = blockIdx.x % tiles_x, ty = blockIdx.x / tiles_x;
    const int x0 = tx * MT - r, y0 = ty * MT - r;
    const auto gin = as_global(in);
    const int WN = WW * WW;
    // k / WW by multiply-high (exact for k < 2^16; the window holds at most 94^2 pixels)
    const uint32_t wmagic = (uint32_t)((0x100000000ull + (unsigned)WW - 1) / (unsigned)WW);
    for (int k0 = 0; k0 < WN; k0 += 8 * M2_THREADS) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = min(k0 + (int)threadIdx.x + M2_THREADS * j, WN - 1);
            const int ly = (int)__umulhi((uint32_t)k, wmagic), lx = k - ly * WW;
            v[j] = gin[(size_t)clampi(y0 + ly, 0, h - 1) * w + clampi(x0 + lx, 0, w - 1)];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + (int)threadIdx.x + M2_THREADS * j;
            if (k >= WN) continue;
            const int ly = (int)__umulhi((uint32_t)k, wmagic), lx = k - ly * WW;
            const int gx = x0 + lx, gy = y0 + ly;
            const bool in_img = gx >= 0 && gx < w && gy >= 0 && gy < h;
            win[ly * WS + lx] = (uint16_t)(in_img ? v[j] : (uint32_t)outside);
        }
    }
    __syncthreads();
    const uint32_t ident = ERODE ? 0x00FF00FFu : 0u;
    constexpr int NP = MT / 2;  // output pairs per row
    // nested runs (each extends the one before it, as for disks): a thread carries its running
    // result from run to run in a register and writes every run's table once, one barrier in all
    bool chain = true;
    for (int d = 0; d < P.nrun; ++d) chain = chain && P.base[d] == d - 1;
    if (chain) {
        for (int k = threadIdx.x; k < WW * NP; k += M2_THREADS) {
            const int x = 2 * (k % NP), row = k / NP;
            const uint16_t* src = win + row * WS + x;
            uint32_t v = taps2<ERODE>(ident, src, r + P.ra[0], r + P.rb[0]);
            *reinterpret_cast<uint32_t*>(H + row * MT + x) = v;
            for (int d = 1; d < P.nrun; ++d) {
                const int a = P.ra[d], b = P.rb[d], pa = P.ra[d - 1], pb = P.rb[d - 1];
                if (a < pa) v = taps2<ERODE>(v, src, r + a, r + pa - 1);
                if (b > pb) v = taps2<ERODE>(v, src, r + pb + 1, r + b);
                *reinterpret_cast<uint32_t*>(H + (d * WW + row) * MT + x) = v;
            }
        }
        __syncthreads();
    }
    for (int d = 0; d < (chain ? 0 : P.nrun); ++d) {
        const int a = P.ra[d], b = P.rb[d], bs = P.base[d];
        uint16_t* Hd = H + d * WW * MT;
        for (int k = threadIdx.x; k < WW * NP; k += M2_THREADS) {
            const int x = 2 * (k % NP), row = k / NP;
            const uint16_t* src = win + row * WS + x;  // pair offset q = r + tap
            uint32_t v;
            if (bs >= 0) {
                v = lds_u32(H + (bs * WW + row) * MT + x);
                if (a < P.ra[bs]) v = taps2<ERODE>(v, src, r + a, r + P.ra[bs] - 1);
                if (b > P.rb[bs]) v = taps2<ERODE>(v, src, r + P.rb[bs] + 1, r + b);
            } else {
                v = taps2<ERODE>(ident, src, r + a, r + b);
            }
            *reinterpret_cast<uint32_t*>(Hd + row * MT + x) = v;
        }
        __syncthreads();
    }
    // output: four pixels per thread (two pairs), one 8-byte LDS read per (kernel row, run)
    constexpr int NQ = MT / 4;
    for (int k = threadIdx.x; k < MT * NQ; k += M2_THREADS) {
        const int x = 4 * (k % NQ), y = k / NQ;
        uint32_t v0 = ident, v1 = ident;
        for (int t = 0; t < P.nrow; ++t) {
            const uint2 hv = *reinterpret_cast<const uint2*>(H + (P.row_run[t] * WW + y + P.row_i[t]) * MT + x);
            v0 = pick2<ERODE>(v0, hv.x);
            v1 = pick2<ERODE>(v1, hv.y);
        }
        const int gx = tx * MT + x, gy = ty * MT + y;
        if (gy >= h || gx >= w) continue;
        const uint32_t o = (v0 & 0xFFu) | ((v0 >> 8) & 0xFF00u) | ((v1 & 0xFFu) << 16) | ((v1 >> 16) << 24);
        uint8_t* orow = out + (size_t)gy * w;
        if (words && gx + 3 < w) {
            *reinterpret_cast<uint32_t*>(orow + gx) = o;
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (gx + j < w) orow[gx + j] = (uint8_t)(o >> (8 * j));
        }
    }
}

// 3 x 3 elements (the cross of HistologicalEntities.cpp:330 and :498): a thread makes four
// consecutive outputs of one row from the 3 x 6 input bytes around them (read as one aligned word
// per row plus its two edge bytes when `words`: both planes 4-byte aligned and the row width a
// multiple of four; else byte by byte).
// `mask` bit 3 * (dy + 1) + (dx + 1) selects the element's taps.
template <bool ERODE>
__global__ __launch_bounds__(256) void morph3_kernel(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, int w, int h,
                                                     unsigned mask, uint8_t outside, bool words) {
    const int qw = (w + 3) >> 2;  // four-pixel groups per row
    const size_t ngroups = (size_t)qw * h;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < ngroups; t += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(t / qw), x0 = (int)(t - (size_t)y * qw) * 4;
        uint32_t b[3][6];  // rows y-1..y+1, columns x0-1..x0+4
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int yy = y + r - 1;
            const bool row_in = yy >= 0 && yy < h;
            const uint8_t* row = in + (size_t)clampi(yy, 0, h - 1) * w;
            if (words && x0 + 3 < w) {
                const uint32_t wd = *reinterpret_cast<const uint32_t*>(row + x0);
#pragma unroll
                for (int j = 0; j < 4; ++j) b[r][1 + j] = (wd >> (8 * j)) & 255u;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) b[r][1 + j] = row[clampi(x0 + j, 0, w - 1)];
            }
            b[r][0] = row[clampi(x0 - 1, 0, w - 1)];
            b[r][5] = row[clampi(x0 + 4, 0, w - 1)];
#pragma unroll
            for (int j = 0; j < 6; ++j) {
                const int xx = x0 + j - 1;
                if (!row_in || xx < 0 || xx >= w) b[r][j] = outside;
            }
        }
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t v = ERODE ? 255u : 0u;
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if ((mask >> (3 * r + c)) & 1u) v = pick<ERODE>(v, b[r][j + c]);
            o[j] = v;
        }
        uint8_t* orow = out + (size_t)y * w;
        if (words && x0 + 3 < w) {
            *reinterpret_cast<uint32_t*>(orow + x0) = o[0] | (o[1] << 8) | (o[2] << 16) | (o[3] << 24);
        } else {
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (x0 + j < w) orow[x0 + j] = (uint8_t)o[j];
        }
    }
}

int make_plan(const uint8_t* kernel, int k, MorphPlan& P) {
    if (k < 1 || k > MAXK || (k % 2) == 0) {
        set_error("structuring element must be odd-sized and at most %d", MAXK);
        return NSC_ERR_ARG;
    }
    P.k = k;
    P.r = k / 2;
    std::vector<std::pair<int, int>> runs;
    std::vector<std::pair<int, std::pair<int, int>>> rows;
    for (int i = 0; i < k; ++i) {
        int j = 0;
        while (j < k) {
            while (j < k && !kernel[i * k + j]) ++j;
            if (j >= k) break;
            int s = j;
            while (j < k && kernel[i * k + j]) ++j;
            rows.push_back({i, {s - P.r, j - 1 - P.r}});
            runs.push_back({s - P.r, j - 1 - P.r});
        }
    }
    std::sort(runs.begin(), runs.end(), [](auto& x, auto& y) {
        int lx = x.second - x.first, ly = y.second - y.first;
        return lx != ly ? lx < ly : x < y;
    });
    runs.erase(std::unique(runs.begin(), runs.end()), runs.end());
    if ((int)runs.size() > MAXRUN || (int)rows.size() > MAXRUN) {
        set_error("structuring element has too many runs");
        return NSC_ERR_ARG;
    }
    P.nrun = (int)runs.size();
    for (int d = 0; d < P.nrun; ++d) {
        P.ra[d] = runs[d].first;
        P.rb[d] = runs[d].second;
        P.base[d] = -1;
        for (int e = d - 1; e >= 0; --e)
            if (runs[e].first >= runs[d].first && runs[e].second <= runs[d].second) { P.base[d] = e; break; }
    }
    P.nrow = (int)rows.size();
    for (int t = 0; t < P.nrow; ++t) {
        P.row_i[t] = rows[t].first;
        int d = 0;
        while (!(runs[d] == rows[t].second)) ++d;
        P.row_run[t] = d;
    }
    return NSC_OK;
}

}  // namespace

int morph_impl(const uint8_t* in, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside, bool erode,
               hipStream_t s) {
    MorphPlan P;
    NSC_CHECK(make_plan(kernel, k, P));
    const int WW = MT + 2 * P.r;
    const size_t lds = (size_t)WW * (WW + 4) + (size_t)P.nrun * WW * MT;
    if (lds > 160 * 1024) {
        set_error("structuring element needs %zu bytes of LDS", lds);
        return NSC_ERR_LIMIT;
    }
    const int tiles_x = (int)cdiv(w, MT), tiles_y = (int)cdiv(h, MT);
    if (P.nrow == 0) {  // empty element: erosion -> 255, dilation -> 0 everywhere
        NSC_HIP(hipMemsetAsync(out, erode ? 255 : 0, (size_t)w * h, s));
        return NSC_OK;
    }
    if (k == 3) {
        unsigned mask = 0;
        for (int i = 0; i < 9; ++i) mask |= (kernel[i] ? 1u : 0u) << i;
        const size_t groups = cdiv((size_t)w, 4) * (size_t)h;
        const dim3 grid((unsigned)std::min<size_t>(cdiv(groups, 256), 16384));
        const bool words = (w & 3) == 0 && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 3) == 0;
        if (erode)
            hipLaunchKernelGGL((morph3_kernel<true>), grid, dim3(256), 0, s, in, out, w, h, mask, (uint8_t)outside, words);
        else
            hipLaunchKernelGGL((morph3_kernel<false>), grid, dim3(256), 0, s, in, out, w, h, mask, (uint8_t)outside, words);
        NSC_HIP(hipGetLastError());
        return NSC_OK;
    }
    const size_t lds2 = 2 * ((size_t)WW * WW + (size_t)P.nrun * WW * MT);
    if (lds2 <= 160 * 1024) {
        const bool words = (w & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 3) == 0;
        const void* fn = erode ? (const void*)morph2_kernel<true> : (const void*)morph2_kernel<false>;
        if (lds2 > 64 * 1024) NSC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2));
        if (erode)
            hipLaunchKernelGGL((morph2_kernel<true>), dim3(tiles_x * tiles_y), dim3(M2_THREADS), lds2, s, in, out, w, h, P,
                               (uint8_t)outside, tiles_x, words);
        else
            hipLaunchKernelGGL((morph2_kernel<false>), dim3(tiles_x * tiles_y), dim3(M2_THREADS), lds2, s, in, out, w, h, P,
                               (uint8_t)outside, tiles_x, words);
        NSC_HIP(hipGetLastError());
        return NSC_OK;
    }
    if (lds > 64 * 1024) {
        const void* fn = erode ? (const void*)morph_kernel<true> : (const void*)morph_kernel<false>;
        NSC_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    if (erode)
        hipLaunchKernelGGL((morph_kernel<true>), dim3(tiles_x * tiles_y), dim3(256), lds, s, in, out, w, h, P,
                           (uint8_t)outside, tiles_x);
    else
        hipLaunchKernelGGL((morph_kernel<false>), dim3(tiles_x * tiles_y), dim3(256), lds, s, in, out, w, h, P,
                           (uint8_t)outside, tiles_x);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

int morph_open_impl(const uint8_t* in, uint8_t* out, uint8_t* tmp, int w, int h, const uint8_t* kernel, int k,
                    hipStream_t s) {
    NSC_CHECK(morph_impl(in, tmp, w, h, kernel, k, 255, true, s));
    return morph_impl(tmp, out, w, h, kernel, k, 0, false, s);
}

}  // namespace nsc

using namespace nsc;

extern "C" {

int nsc_erode_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside,
                 void* stream) {
    NSC_ARG(img && out && kernel && w > 0 && h > 0 && outside >= 0 && outside <= 255, "erode_u8");
    NSC_CHECK(ensure_init());
    return morph_impl(img, out, w, h, kernel, k, outside, true, as_stream(stream));
}

int nsc_dilate_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside,
                  void* stream) {
    NSC_ARG(img && out && kernel && w > 0 && h > 0 && outside >= 0 && outside <= 255, "dilate_u8");
    NSC_CHECK(ensure_init());
    return morph_impl(img, out, w, h, kernel, k, outside, false, as_stream(stream));
}

int nsc_morph_open_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, void* stream) {
    NSC_ARG(img && out && kernel && w > 0 && h > 0, "morph_open_u8");
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    Scratch sc;
    NSC_CHECK(sc.alloc((size_t)w * h, s));
    return morph_open_impl(img, out, sc.as<uint8_t>(), w, h, kernel, k, s);
}

}  // extern "C"
