// Per-nucleus features of src/features/ObjFeatures.cpp (geometric: area, perimeter, ellipse axes
// and eccentricity, extent ratio, circularity; histogram statistics of an intensity plane and of
// its gradient magnitude, Operators.cpp) from a label image and the reference's box block
// (5 x n, column stride n: [label | minx | maxx | miny | maxy]).
//
// One workgroup per nucleus walks its box (the boxes hold a few hundred pixels): the moment sums
// of the ellipse fit are exact (offsets from the box centre are multiples of 1/2, their products
// of 1/4, summed in double), so they do not depend on the order of the additions; the perimeter
// counts the 2x2 cells of the box padded by one empty pixel per table class (identical to the
// reference's interior, edge and corner cells), the histograms are integer counts in LDS.  The
// statistics are then evaluated per nucleus by one thread with the reference's arithmetic.
#include <cfloat>
#include <cmath>

#include "common.hpp"

namespace nsc {
namespace {

constexpr double kHalfDiag = 0.70710678118;  // ObjFeatures.cpp:151-166 table values
constexpr double kDiag = 1.41421356237;
constexpr double kPi = 3.14159265359;         // ObjFeatures.cpp:10

// table class of a marching-squares cell code: 0 none, 1 half diagonal, 2 straight, 3 two diagonals
__device__ __forceinline__ int cell_class(int m) {
    // codes 1 2 4 8 7 11 13 14 -> 1; 3 6 9 12 -> 2; 5 10 -> 3; 0 15 -> 0
    constexpr uint32_t lo = 0x6DB6;  // bit c set: class is odd  (1 or 3)
    constexpr uint32_t hi = 0x1668;  // bit c set: class >= 2    (2 or 3)
    return (int)(((lo >> m) & 1u) | (((hi >> m) & 1u) << 1));
}

// Operators::gradient (Operators.cpp:568-640): central differences [-0.5 0 0.5] with replicated
// borders, one-sided on the first / last column and row, float magnitude, rounded half up,
// clamped to 0..255.  Every value is exact before the square root (multiples of 1/4 below 2^17),
// and the float square root is correctly rounded, so the plane is exact.
__global__ void k_gradient(const uint8_t* __restrict__ img, int stride, int w, int h, uint8_t* __restrict__ out) {
    const size_t n = (size_t)w * h;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(i / w), x = (int)(i - (size_t)y * w);
        const int xl = max(x - 1, 0), xr = min(x + 1, w - 1), yu = max(y - 1, 0), yd = min(y + 1, h - 1);
        const float l = img[((size_t)y * w + xl) * stride], r = img[((size_t)y * w + xr) * stride];
        const float u = img[((size_t)yu * w + x) * stride], d = img[((size_t)yd * w + x) * stride];
        float dx = 0.5f * r - 0.5f * l;
        if (x == 0 || x == w - 1) dx *= 2.0f;
        float dy = 0.5f * d - 0.5f * u;
        if (y == 0 || y == h - 1) dy *= 2.0f;
        const float mag = __fsqrt_rn(dx * dx + dy * dy);
        const float fl = floorf(mag);
        int v = (int)fl + ((mag - fl >= 0.5f) ? 1 : 0);
        out[i] = (uint8_t)min(v, 255);
    }
}

struct FeatOut {
    int32_t* area;
    double *perimeter, *major, *minor, *ecc, *extent, *circ;
    int* hist;  // n x 512: intensity then gradient counts (scratch)
};

// One workgroup per nucleus (grid-stride): area, moment sums, cell classes, histograms.
__global__ __launch_bounds__(256) void k_object_pass(const int32_t* __restrict__ L, int w, int h,
                                                     const int32_t* __restrict__ box, int n,
                                                     const uint8_t* __restrict__ img, int stride,
                                                     const uint8_t* __restrict__ grad, FeatOut F) {
    __shared__ int hist[512];
    __shared__ double red[5][4];
    __shared__ int ired[4][4];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int label = box[i], minx = box[n + i], maxx = box[2 * n + i], miny = box[3 * n + i], maxy = box[4 * n + i];
        const int bw = maxx - minx + 1, bh = maxy - miny + 1;
        const float midX = (float)(minx + maxx) / 2.0f, midY = (float)(miny + maxy) / 2.0f;
        for (int k = t; k < 512; k += 256) hist[k] = 0;
        __syncthreads();
        int area = 0;
        double sx = 0, sy = 0, sxy = 0, ssqx = 0, ssqy = 0;
        for (int k = t; k < bw * bh; k += 256) {
            const int yy = k / bw, xx = k - yy * bw;
            const int x = minx + xx, y = miny + yy;
            const size_t p = (size_t)y * w + x;
            if (L[p] == label) {
                ++area;
                const float cx = (float)x - midX, cy = (float)y - midY;
                sx += cx;
                sy += cy;
                sxy += cx * cy;
                ssqx += cx * cx;
                ssqy += cy * cy;
                if (img) atomicAdd(&hist[img[p * stride]], 1);
                if (grad) atomicAdd(&hist[256 + grad[p]], 1);
            }
        }
        // marching-squares cells of the box padded by one empty pixel: (bw + 1) x (bh + 1)
        int c1 = 0, c2 = 0, c3 = 0;
        const int cw = bw + 1, ch = bh + 1;
        for (int k = t; k < cw * ch; k += 256) {
            const int cy = k / cw, cx = k - cy * cw;
            const int x0 = minx + cx - 1, y0 = miny + cy - 1;  // the cell's lower-left pixel
            auto in = [&](int x, int y) {
                return (x >= minx && x <= maxx && y >= miny && y <= maxy && L[(size_t)y * w + x] == label) ? 1 : 0;
            };
            const int m = (in(x0, y0) << 3) | (in(x0 + 1, y0) << 2) | (in(x0 + 1, y0 + 1) << 1) | in(x0, y0 + 1);
            const int c = cell_class(m);
            c1 += c == 1;
            c2 += c == 2;
            c3 += c == 3;
        }
        // workgroup sums (double sums are exact: see the file comment)
        double dv[5] = {sx, sy, sxy, ssqx, ssqy};
        int iv[4] = {area, c1, c2, c3};
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int q = 0; q < 5; ++q) dv[q] += __shfl_down(dv[q], off);
#pragma unroll
            for (int q = 0; q < 4; ++q) iv[q] += __shfl_down(iv[q], off);
        }
        if (lane == 0) {
            for (int q = 0; q < 5; ++q) red[q][wv] = dv[q];
            for (int q = 0; q < 4; ++q) ired[q][wv] = iv[q];
        }
        __syncthreads();
        if (t == 0) {
            double S[5] = {0, 0, 0, 0, 0};
            int I[4] = {0, 0, 0, 0};
            for (int k = 0; k < 4; ++k) {
                for (int q = 0; q < 5; ++q) S[q] += red[q][k];
                for (int q = 0; q < 4; ++q) I[q] += ired[q][k];
            }
            const double a = I[0];
            F.area[i] = I[0];
            // ObjFeatures::ellipse (ObjFeatures.cpp:96-110)
            const double frac = 1.0 / 12.0, root = sqrt(8.0);
            const double xbar = S[0] / a, ybar = S[1] / a;
            const double mxx = ((S[3] + a * xbar * xbar - 2.0 * xbar * S[0]) / a) + frac;
            const double myy = ((S[4] + a * ybar * ybar - 2.0 * ybar * S[1]) / a) + frac;
            const double mxy = (S[2] - ybar * S[0] - xbar * S[1] + xbar * ybar * a) / a;
            const double delta = sqrt((mxx - myy) * (mxx - myy) + 4.0 * mxy * mxy);
            const double maj = root * sqrt(mxx + myy + delta), mnr = root * sqrt(mxx + myy - delta);
            F.major[i] = maj;
            F.minor[i] = mnr;
            F.ecc[i] = (2.0 * sqrt(fabs(maj * maj - mnr * mnr) / 4.0)) / maj;
            // perimeter (ObjFeatures.cpp:144-278) from the cell class counts, extent (:118-141),
            // circularity (:281-293)
            const double per = (double)I[1] * kHalfDiag + (double)I[2] * 1.0 + (double)I[3] * kDiag;
            F.perimeter[i] = per;
            F.extent[i] = a / (double)(bw * bh);
            F.circ[i] = (4.0 * kPi * a) / (per * per);
        }
        if (F.hist)
            for (int k = t; k < 512; k += 256) F.hist[(size_t)i * 512 + k] = hist[k];
        __syncthreads();  // hist and the reduction buffers are reused by the next nucleus
    }
}

// histogram statistics of Operators.cpp:281-505 with the reference's arithmetic: out[0..] = mean,
// (max, min when with_minmax), std, entropy, energy, skewness, kurtosis
__device__ void hist_stats(const int* hist, float* f, bool with_minmax) {
    double hm = 0.0;
    int ne = 0;
    for (int i = 0; i < 256; ++i) {
        hm += hist[i] * i;
        ne += hist[i];
    }
    const double mean = ne == 0 ? 0.0 : hm / ne;
    unsigned mx = 0, mn = 256;
    for (int i = 255; i >= 0; --i)
        if (hist[i] > 0) { mx = i; break; }
    for (int i = 0; i < 256; ++i)
        if (hist[i] > 0) { mn = i; break; }
    double sum = 0.0, nd = 0;
    for (int i = 0; i < 256; ++i) {
        sum += i * hist[i];
        nd += hist[i];
    }
    const double m2 = sum / nd;
    double sq = 0.0;
    for (int i = 0; i < 256; ++i) {
        const double d = i - m2;
        sq += d * d * hist[i];
    }
    const double sd = sqrt(sq / nd);
    const float nf = (float)ne;
    float ent = 0.0f, energy = 0.0f;
    for (int i = 0; i < 256; ++i) {
        const float p = (float)hist[i] / nf;
        if (hist[i] != 0) ent += p * log2f(p);
    }
    for (int i = 0; i < 256; ++i) {
        const double p = (double)((float)hist[i] / nf);
        energy = (float)((double)energy + p * p);
    }
    float n2 = 0.f, avg = 0.f;
    for (int i = 0; i < 256; ++i) {
        n2 += (float)hist[i];
        avg += i * (float)hist[i];
    }
    avg /= n2;
    float d3 = 0.f, d2 = 0.f, d4 = 0.f;
    for (int i = 0; i < 256; ++i) {
        const double e = (float)hist[i];
        const double dd = (double)((float)i - avg);
        d3 = (float)((double)d3 + e * (dd * dd * dd));
        d2 = (float)((double)d2 + e * (dd * dd));
        d4 = (float)((double)d4 + e * ((dd * dd) * (dd * dd)));
    }
    float skew = 0.f, kurt = 0.f;
    if (d3 != 0 && d2 != 0 && n2 != 0) {
        float a = d3 / n2, b = d2 / n2;
        b = sqrtf(b);
        b = (float)((double)b * (double)b * (double)b);
        skew = a / b;
    }
    if (d4 != 0 && d2 != 0 && n2 != 0) {
        float a = d4 / n2, b = d2 / n2;
        b = (float)((double)b * (double)b);
        kurt = a / b;
    }
    int k = 0;
    f[k++] = (float)mean;
    if (with_minmax) {
        f[k++] = (float)mx;
        f[k++] = (float)mn;
    }
    f[k++] = (float)sd;
    f[k++] = -1.0f * ent;
    f[k++] = energy;
    f[k++] = skew;
    f[k++] = kurt;
}

// one thread per nucleus: intensity (8) and gradient (6) statistics from its histograms
__global__ void k_hist_features(const int* __restrict__ hist, int n, float* __restrict__ f8, float* __restrict__ f6) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (f8) hist_stats(hist + (size_t)i * 512, f8 + 8 * (size_t)i, true);
    if (f6) hist_stats(hist + (size_t)i * 512 + 256, f6 + 6 * (size_t)i, false);
}

}  // namespace
}  // namespace nsc

using namespace nsc;

extern "C" int nsc_object_features(const int32_t* labels, int w, int h, const int32_t* bbox5, int n, const uint8_t* img,
                                   int channels, int channel, nsc_object_features_out* out, void* stream) {
    NSC_ARG(labels && bbox5 && out && w > 0 && h > 0 && n >= 0, "object_features");
    NSC_ARG(!img || (channels >= 1 && channel >= 0 && channel < channels), "object_features: channels");
    NSC_ARG(out->area && out->perimeter && out->major_axis && out->minor_axis && out->eccentricity &&
                out->extent_ratio && out->circularity, "object_features: geometric outputs");
    NSC_ARG(img || (!out->intensity && !out->gradient), "object_features: intensity outputs need an image");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    if (n == 0) return NSC_OK;
    hipStream_t s = as_stream(stream);
    const bool hists = img && (out->intensity || out->gradient);
    const size_t np = (size_t)w * h;
    Scratch sc;
    NSC_CHECK(sc.alloc((hists ? sizeof(int) * 512 * (size_t)n : 0) + (out->gradient ? np : 0) + 256, s));
    int* hist = hists ? sc.as<int>() : nullptr;
    uint8_t* grad = out->gradient ? reinterpret_cast<uint8_t*>(sc.as<char>() + (hists ? sizeof(int) * 512 * (size_t)n : 0))
                                  : nullptr;
    const uint8_t* plane = img ? img + channel : nullptr;
    if (grad)
        hipLaunchKernelGGL(k_gradient, dim3(std::min<unsigned>(cdiv(np, 256), 8192u)), dim3(256), 0, s, plane, channels, w,
                           h, grad);
    FeatOut F{out->area, out->perimeter, out->major_axis, out->minor_axis, out->eccentricity, out->extent_ratio,
              out->circularity, hist};
    hipLaunchKernelGGL(k_object_pass, dim3(std::min(n, 8192)), dim3(256), 0, s, labels, w, h, bbox5, n,
                       out->intensity ? plane : nullptr, channels, grad, F);
    if (hists)
        hipLaunchKernelGGL(k_hist_features, dim3(cdiv(n, 64)), dim3(64), 0, s, hist, n, out->intensity, out->gradient);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}
