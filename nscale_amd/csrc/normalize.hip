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
// double (the constants are double literals) rounded to float, everything else is float.  The
// two transcendental steps (log10f, powf(10, x)) are evaluated in double and rounded once, so
// they are correctly rounded except where the double result lies within about 2^-50 of a float
// rounding boundary; a C library's float versions may differ from them by an ulp.
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
// each discriminant; each addWeighted evaluates in double and stores float.
__device__ __forceinline__ bool norm_fg(const NormDisc& D, uint32_t b, uint32_t g, uint32_t r) {
    float d[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const float* m = D.m + 4 * k;
        const float t = (float)((double)r * (double)m[0] + (double)g * (double)m[1] + 0.0);
        d[k] = (float)((double)t * 1.0 + (double)b * (double)m[2] + (double)m[3]);
    }
    return d[0] < d[1];
}

__device__ __forceinline__ float log10_rn(float x) { return (float)log10((double)x); }
__device__ __forceinline__ float exp10_rn(float x) { return (float)exp10((double)x); }

// bgr2Lab of one pixel (Normalization.cpp:64-146) after PixelClass's scaling (:328-333)
__device__ __forceinline__ void norm_lab(const NormConsts& C, uint32_t b8, uint32_t g8, uint32_t r8, float lab[3]) {
    const float b = (float)b8 * C.s255, g = (float)g8 * C.s255, r = (float)r8 * C.s255;
    const float L = (float)(0.3811 * r + 0.5783 * g + 0.0402 * b);
    const float M = (float)(0.1967 * r + 0.7244 * g + 0.0782 * b);
    const float S = (float)(0.0241 * r + 0.1288 * g + 0.8444 * b);
    const float l = log10_rn(L), m = log10_rn(M), s = log10_rn(S);
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
    float L = exp10_rn(ll), M = exp10_rn(lm), S = exp10_rn(ls);
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
