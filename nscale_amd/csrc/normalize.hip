// Colour normalization of src/normalization/Normalization.cpp: Reinhard's transfer in the
// l-alpha-beta space of the foreground that a fixed linear discriminant separates from the
// background (normalization, Normalization.cpp:436-499), and the foreground statistics alone
// (targetParameters, Normalization.cpp:502-582).
//
// The reference builds seven full-size float planes; here every pixel's result depends on its
// own colour and six statistics only, so the work is two streaming passes over the BGR image:
//   1. k_norm_stats: discriminant, LAB, and per-workgroup double sums of the selected pixels'
//      values and float squares (one row chunk of 1024 pixels per work item);
//   2. k_norm_finish (one workgroup): sums the partials in a fixed order, mean / std in double,
//      the per-channel float transfer factors;
//   3. k_norm_apply: LAB again, the transfer for foreground pixels, back to BGR.
// Arithmetic follows the reference operation by operation: the discriminant is the double
// arithmetic of OpenCV 2.4's float addWeighted rounded to float twice, the RGB->LMS products are
// double (the constants are double literals) rounded to float, everything else is float.  log10f
// follows the fdlibm formula C libraries use (float operations around a correctly rounded logf);
// powf(10, x) is correctly rounded.  Both evaluate their core in double with relative error near
// 2^-48, so they differ from a C library only where that library's own logf / powf is not
// correctly rounded (under 1% / 0.1% of arguments for glibc 2.35).
#include <cfloat>
#include <cmath>

#include "common.hpp"

namespace nsc {
namespace {

constexpr int kNormThreads = 256;
constexpr int kNormChunk = 1024;        // pixels of one row per work item (4 per thread)
constexpr int kNormMaxBlocks = 2048;    // statistics partials: kNormMaxBlocks x 8 doubles

struct NormConsts {
    float k3, k6, k2;  // 1/sqrtf(3), 1/sqrtf(6), 1/sqrtf(2) (Normalization.cpp:112-114, 172-174)
    float s255;        // the float factor convertTo uses for "/= 255" (Normalization.cpp:331)
};

// Fisher discriminant coefficients (Normalization.cpp:451): rows background, foreground over R G B 1
struct NormDisc {
    float m[8];
};

struct NormParams {
    float meanI[3], ratio[3];  // image statistics and target_std / image_std (float division)
};

// segFG (Normalization.cpp:19-56): addWeighted(addWeighted(R, m0, G, m1, 0), 1, B, m2, m3) for
// each discriminant; each addWeighted evaluates in double and stores float.  With the decimal
// coefficients, d_fg - d_bg = (97 R - 852 G + 621 B - 4169) / 1000 exactly; the float coefficients
// and the float roundings of values below 300 move it by less than 1e-4, so wherever that integer
// is not zero its sign decides, and only exact ties (75 of the 2^24 colours) take the float path.
__device__ __forceinline__ bool norm_fg(const NormDisc& D, uint32_t b, uint32_t g, uint32_t r) {
    const int num = 97 * (int)r - 852 * (int)g + 621 * (int)b - 4169;
    if (num != 0) return num > 0;
    float d[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const float* m = D.m + 4 * k;
        const float t = (float)((double)r * (double)m[0] + (double)g * (double)m[1] + 0.0);
        d[k] = (float)((double)t * 1.0 + (double)b * (double)m[2] + (double)m[3]);
    }
    return d[0] < d[1];
}

// Natural logarithm of m in [0.5, 2) in double: m = 2^e u with u in [sqrt(1/2), sqrt(2)),
// log u = 2 atanh(t), t = (u - 1) / (u + 1), |t| <= 0.1716, series to t^17 (the first omitted
// term is below 2^-50 of the result).  Relative error about 2^-50, so the float rounding of the
// result is the correct rounding except within that distance of a rounding boundary.  The
// reciprocal is the hardware estimate refined by two Newton steps; polynomials use fused
// multiply-adds (this is not an operation of the reference, only its accuracy matters).
__device__ __forceinline__ double log_unit(double m) {
    double e = 0.0;
    if (m > 0x1.6a09e667f3bcdp+0) {
        m *= 0.5;
        e = 1.0;
    } else if (m < 0x1.6a09e667f3bcdp-1) {
        m *= 2.0;
        e = -1.0;
    }
    const double d = m + 1.0;  // m - 1 and m + 1 exact (m has 24 significant bits)
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    const double t = (m - 1.0) * r;
    const double t2 = t * t;
    double p = 1.0 / 17;
    p = fma(p, t2, 1.0 / 15);
    p = fma(p, t2, 1.0 / 13);
    p = fma(p, t2, 1.0 / 11);
    p = fma(p, t2, 1.0 / 9);
    p = fma(p, t2, 1.0 / 7);
    p = fma(p, t2, 1.0 / 5);
    p = fma(p, t2, 1.0 / 3);
    p = fma(p, t2, 1.0);
    return fma(e, 0x1.62e42fefa39efp-1, 2.0 * t * p);
}

// log10f as the fdlibm-derived C libraries compute it (glibc's e_log10f.c, the function the
// reference calls at Normalization.cpp:92-94): x = 2^y m with m in [1, 2) (in [0.5, 1) when the
// exponent is negative), log10 x = y lo + ivln10 logf(m) + y hi in float operations with fdlibm's
// constants; logf(m) is the correctly rounded natural logarithm.
__device__ __forceinline__ float log10f_lib(float x) {
    constexpr float kIvln10 = 0x1.bcb7b2p-2f;  // 4.3429449201e-01
    constexpr float kHi = 0x1.344100p-2f;      // 3.0102920532e-01
    constexpr float kLo = 0x1.a84fb6p-21f;     // 7.9034151668e-07
    if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : __builtin_nanf("");
    if (x == INFINITY) return x;
    uint32_t hx = __float_as_uint(x);
    int k = 0;
    if (hx < 0x00800000u) {  // subnormal: scale by 2^25
        x *= 33554432.0f;
        hx = __float_as_uint(x);
        k = -25;
    }
    k += (int)(hx >> 23) - 127;
    const int i = k < 0 ? 1 : 0;
    const float m = __uint_as_float((hx & 0x007fffffu) | ((uint32_t)(0x7f - i) << 23));
    const float y = (float)(k + i);
    const float lnm = (float)log_unit((double)m);
    const float z = y * kLo + kIvln10 * lnm;
    return z + y * kHi;
}

// 10^y rounded to float, evaluated in double: y log2(10) = k + f with the constant split so that
// y * hi is exact, 2^f = e^(f ln2) by its Taylor series to degree 11 (|f ln2| <= 0.347, truncation
// below 2^-48, fused multiply-adds); relative error about 2^-48 before the final rounding.
__device__ __forceinline__ float exp10f_rn(float y) {
    if (y != y) return y;
    if (y > 40.0f) return INFINITY;
    if (y < -46.0f) return 0.0f;
    const double z = (double)y * 0x1.a934f09p+1;  // log2(10) to 29 bits: exact product
    const double k = rint(z);
    const double f = (z - k) + (double)y * 0x1.e68dc57f24960p-29;
    const double g = f * 0x1.62e42fefa39efp-1;
    double p = 1.0 / 39916800;
    p = fma(p, g, 1.0 / 3628800);
    p = fma(p, g, 1.0 / 362880);
    p = fma(p, g, 1.0 / 40320);
    p = fma(p, g, 1.0 / 5040);
    p = fma(p, g, 1.0 / 720);
    p = fma(p, g, 1.0 / 120);
    p = fma(p, g, 1.0 / 24);
    p = fma(p, g, 1.0 / 6);
    p = fma(p, g, 0.5);
    p = fma(p, g, 1.0);
    p = fma(p, g, 1.0);
    return (float)ldexp(p, (int)k);
}

// bgr2Lab of one pixel (Normalization.cpp:64-146) after PixelClass's scaling (:328-333)
__device__ __forceinline__ void norm_lab(const NormConsts& C, uint32_t b8, uint32_t g8, uint32_t r8, float lab[3]) {
    const float b = (float)b8 * C.s255, g = (float)g8 * C.s255, r = (float)r8 * C.s255;
    const float L = (float)(0.3811 * r + 0.5783 * g + 0.0402 * b);
    const float M = (float)(0.1967 * r + 0.7244 * g + 0.0782 * b);
    const float S = (float)(0.0241 * r + 0.1288 * g + 0.8444 * b);
    const float l = log10f_lib(L), m = log10f_lib(M), s = log10f_lib(S);
    lab[0] = l * C.k3 + m * C.k3 + s * C.k3;
    lab[1] = l * C.k6 + m * C.k6 + s * (-2.0f * C.k6);
    lab[2] = l * C.k2 + m * (-C.k2) + s * 0.0f;  // s * 0 keeps the NaN of a black pixel
}

// lab2BGR of one pixel (Normalization.cpp:162-306): back to log LMS, 10^x (non-finite -> 0), the
// LMS->RGB matrix, x 255 in double, clamp, round half up; a NaN channel gives 0
__device__ __forceinline__ void norm_bgr(const NormConsts& C, const float lab[3], uint8_t out[3]) {
    const float ll = lab[0] * C.k3 + lab[1] * C.k6 + lab[2] * C.k2;
    const float lm = lab[0] * C.k3 + lab[1] * C.k6 + lab[2] * (-C.k2);
    const float ls = lab[0] * C.k3 + lab[1] * (-2.0f * C.k6) + lab[2] * (0.0f * C.k2);
    float L = exp10f_rn(ll), M = exp10f_rn(lm), S = exp10f_rn(ls);
    if (!(fabsf(L) <= FLT_MAX)) L = 0.0f;
    if (!(fabsf(M) <= FLT_MAX)) M = 0.0f;
    if (!(fabsf(S) <= FLT_MAX)) S = 0.0f;
    const float bgrf[3] = {L * 0.0585f + M * -0.2611f + S * 1.2057f,
                           L * -1.2197f + M * 2.3831f + S * -0.1626f,
                           L * 4.4687f + M * -3.5887f + S * 0.1196f};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float v = (float)((double)bgrf[k] * 255.0);
        v = v < 0.0f ? 0.0f : v;
        v = v > 255.0f ? 255.0f : v;
        const float f = floorf(v);
        int q = (int)f + ((v - f >= 0.5f) ? 1 : 0);
        out[k] = (v == v) ? (uint8_t)q : (uint8_t)0;
    }
}

struct NormImage {
    const uint8_t* bgr;
    size_t ld;
    int w, h, chunks_per_row;
    long long items;  // h * chunks_per_row
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Per-workgroup sums over the selected pixels: [sum L, sum a, sum b, sq L, sq a, sq b, count, 0]
__global__ __launch_bounds__(kNormThreads) void k_norm_stats(NormImage I, NormConsts C, NormDisc D,
                                                             double* __restrict__ part) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long long it = blockIdx.x; it < I.items; it += gridDim.x) {
        const int y = (int)(it / I.chunks_per_row);
        const int x0 = (int)(it - (long long)y * I.chunks_per_row) * kNormChunk;
        const uint8_t* row = I.bgr + (size_t)y * I.ld;
#pragma unroll
        for (int k = 0; k < kNormChunk / kNormThreads; ++k) {
            const int x = x0 + k * kNormThreads + (int)threadIdx.x;
            if (x >= I.w) break;
            const uint8_t* p = row + 3 * (size_t)x;
            const uint32_t b = p[0], g = p[1], r = p[2];
            if (!norm_fg(D, b, g, r)) continue;
            float lab[3];
            norm_lab(C, b, g, r, lab);
            if (!((lab[0] * lab[1]) * lab[2] > 0.0f)) continue;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                acc[c] += (double)lab[c];
                acc[3 + c] += (double)(lab[c] * lab[c]);
            }
            acc[6] += 1.0;
        }
    }
    __shared__ double red[kNormThreads / 64][8];
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const double v = wave_sum(acc[c]);
        if (ln == 0) red[wv][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        double v = 0.0;
        if (threadIdx.x < 7)
            for (int k = 0; k < kNormThreads / 64; ++k) v += red[k][threadIdx.x];
        part[(size_t)blockIdx.x * 8 + threadIdx.x] = v;
    }
}

// Sums the partials in a fixed order (deterministic), then TransferI's statistics
// (Normalization.cpp:383-394): mean = sum / count, var = sq / count - mean^2, std = sqrt(var),
// each stored as float; the transfer factor is target_std / std in float (Normalization.cpp:402).
// out: NormParams, then double mean[3], std[3], count.
__global__ __launch_bounds__(kNormThreads) void k_norm_finish(const double* __restrict__ part, int nblocks,
                                                              NormParams target, double* __restrict__ stats,
                                                              NormParams* __restrict__ params) {
    __shared__ double red[kNormThreads / 64][8];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += kNormThreads)
#pragma unroll
        for (int c = 0; c < 7; ++c) acc[c] += part[(size_t)b * 8 + c];
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const double v = wave_sum(acc[c]);
        if (ln == 0) red[wv][c] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot[7];
        for (int c = 0; c < 7; ++c) {
            tot[c] = 0.0;
            for (int k = 0; k < kNormThreads / 64; ++k) tot[c] += red[k][c];
        }
        const double n = tot[6];
        NormParams P;
        for (int c = 0; c < 3; ++c) {
            const double mean = tot[c] / n;
            const double var = tot[3 + c] / n - mean * mean;
            const double sd = sqrt(var);
            stats[c] = mean;
            stats[3 + c] = sd;
            P.meanI[c] = (float)mean;
            P.ratio[c] = target.ratio[c] / (float)sd;  // target.ratio holds target_std here
        }
        stats[6] = n;
        *params = P;
    }
}

// normalization's per-pixel part (Normalization.cpp:479-496): foreground LAB values become
// (x - meanI) * ratio + target_mean in float, background ones stay, then lab2BGR.
__global__ __launch_bounds__(kNormThreads) void k_norm_apply(NormImage I, NormConsts C, NormDisc D,
                                                             const NormParams* __restrict__ params,
                                                             NormParams target, uint8_t* out,
                                                             size_t out_ld) {
    const NormParams P = *params;
    for (long long it = blockIdx.x; it < I.items; it += gridDim.x) {
        const int y = (int)(it / I.chunks_per_row);
        const int x0 = (int)(it - (long long)y * I.chunks_per_row) * kNormChunk;
        const uint8_t* row = I.bgr + (size_t)y * I.ld;
        uint8_t* orow = out + (size_t)y * out_ld;
#pragma unroll
        for (int k = 0; k < kNormChunk / kNormThreads; ++k) {
            const int x = x0 + k * kNormThreads + (int)threadIdx.x;
            if (x >= I.w) break;
            const uint8_t* p = row + 3 * (size_t)x;
            const uint32_t b = p[0], g = p[1], r = p[2];
            float lab[3];
            norm_lab(C, b, g, r, lab);
            if (norm_fg(D, b, g, r)) {
#pragma unroll
                for (int c = 0; c < 3; ++c) lab[c] = (lab[c] - P.meanI[c]) * P.ratio[c] + target.meanI[c];
            }
            uint8_t o[3];
            norm_bgr(C, lab, o);
            uint8_t* q = orow + 3 * (size_t)x;
            q[0] = o[0];
            q[1] = o[1];
            q[2] = o[2];
        }
    }
}

NormConsts host_consts() {
    NormConsts C;
    C.k3 = 1 / sqrtf(3.0f);
    C.k6 = 1 / sqrtf(6.0f);
    C.k2 = 1 / sqrtf(2.0f);
    C.s255 = (float)(1. / 255);
    return C;
}

NormDisc host_disc() {
    return NormDisc{{-0.154f, 0.035f, 0.549f, -45.718f, -0.057f, -0.817f, 1.170f, -49.887f}};
}

NormImage image_of(const uint8_t* bgr, int w, int h, size_t ld) {
    NormImage I;
    I.bgr = bgr;
    I.ld = ld;
    I.w = w;
    I.h = h;
    I.chunks_per_row = (int)cdiv((size_t)w, kNormChunk);
    I.items = (long long)h * I.chunks_per_row;
    return I;
}

int stats_blocks(const NormImage& I) { return (int)std::min<long long>(I.items, kNormMaxBlocks); }

size_t norm_scratch_bytes(int nblocks) {
    Carver c(nullptr);
    c.take<double>((size_t)nblocks * 8);
    c.take<double>(8);
    c.take<NormParams>(1);
    return c.size();
}

// stats + finish on the stream; the device results are left in the carved scratch
int norm_statistics(const NormImage& I, const float* target_std3, void* scratch, int nblocks, double** stats,
                    NormParams** params, hipStream_t s) {
    Carver c(scratch);
    double* part = c.take<double>((size_t)nblocks * 8);
    *stats = c.take<double>(8);
    *params = c.take<NormParams>(1);
    NormParams T{};
    for (int k = 0; k < 3; ++k) T.ratio[k] = target_std3 ? target_std3[k] : 1.0f;
    hipLaunchKernelGGL(k_norm_stats, dim3(nblocks), dim3(kNormThreads), 0, s, I, host_consts(), host_disc(), part);
    hipLaunchKernelGGL(k_norm_finish, dim3(1), dim3(kNormThreads), 0, s, part, nblocks, T, *stats, *params);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

int check_image(const uint8_t* bgr, int w, int h, size_t ld) {
    NSC_ARG(bgr && w > 0 && h > 0, "normalization: image");
    NSC_ARG(ld >= (size_t)3 * w, "normalization: row stride below 3 * width");
    return NSC_OK;
}

}  // namespace
}  // namespace nsc

using namespace nsc;

extern "C" int nsc_normalization_target_parameters(const uint8_t* bgr, int w, int h, size_t ld, float* mean3,
                                                   float* std3, int64_t* count, void* stream) {
    NSC_CHECK(ensure_init());
    NSC_CHECK(check_image(bgr, w, h, ld));
    NSC_ARG(mean3 && std3, "normalization_target_parameters: outputs");
    const hipStream_t s = as_stream(stream);
    const NormImage I = image_of(bgr, w, h, ld);
    const int nb = stats_blocks(I);
    Scratch sc;
    NSC_CHECK(sc.alloc(norm_scratch_bytes(nb), s));
    double* stats;
    NormParams* params;
    NSC_CHECK(norm_statistics(I, nullptr, sc.p, nb, &stats, &params, s));
    double host[8];
    NSC_HIP(hipMemcpyAsync(host, stats, sizeof(host), hipMemcpyDeviceToHost, s));
    NSC_HIP(stream_sync(s));
    for (int k = 0; k < 3; ++k) {
        mean3[k] = (float)host[k];
        std3[k] = (float)host[3 + k];
    }
    if (count) *count = (int64_t)host[6];
    return NSC_OK;
}

extern "C" int nsc_normalization(const uint8_t* bgr, int w, int h, size_t ld, const float* target_mean3,
                                 const float* target_std3, uint8_t* out, size_t out_ld, void* stream) {
    NSC_CHECK(ensure_init());
    NSC_CHECK(check_image(bgr, w, h, ld));
    NSC_ARG(target_mean3 && target_std3 && out, "normalization: targets / output");
    NSC_ARG(out_ld >= (size_t)3 * w, "normalization: output row stride below 3 * width");
    // in place is fine (each pixel is read before it is written, by the same thread, after the
    // statistics pass); any other overlap is not
    const size_t in_bytes = ld * (size_t)(h - 1) + (size_t)3 * w, out_bytes = out_ld * (size_t)(h - 1) + (size_t)3 * w;
    NSC_ARG(out == bgr ? out_ld == ld
                       : ((const uint8_t*)out + out_bytes <= bgr || bgr + in_bytes <= (const uint8_t*)out),
            "normalization: output overlaps the input (only exact in-place use is allowed)");
    const hipStream_t s = as_stream(stream);
    const NormImage I = image_of(bgr, w, h, ld);
    const int nb = stats_blocks(I);
    Scratch sc;
    NSC_CHECK(sc.alloc(norm_scratch_bytes(nb), s));
    double* stats;
    NormParams* params;
    NSC_CHECK(norm_statistics(I, target_std3, sc.p, nb, &stats, &params, s));
    NormParams T{};
    for (int k = 0; k < 3; ++k) T.meanI[k] = target_mean3[k];
    const int nbx = (int)std::min<long long>(I.items, 8 * kNormMaxBlocks);
    hipLaunchKernelGGL(k_norm_apply, dim3(nbx), dim3(kNormThreads), 0, s, I, host_consts(), host_disc(),
                       (const NormParams*)params, T, out, out_ld);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}
