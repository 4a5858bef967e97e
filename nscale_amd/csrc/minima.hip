// imhmin<float> (MorphologicOperations.cpp:3342-3364) and localMaxima / localMinima<float>
// (MorphologicOperations.cpp:3763-3845).
//
// localMaxima(img, conn):
//   candidate  = !(img - 1 < imreconstruct(img - 1, img, conn))     (float arithmetic as the reference)
//   a candidate is a regional maximum unless its plateau -- the conn-connected set of candidates
//   with the same value, which is exactly the region cv::floodFill(FIXED_RANGE, lo = up = 0,
//   MASK_ONLY) marks from a seed -- holds a pixel with a strictly greater conn-neighbour (pixels
//   outside the image count as -FLT_MAX, the reference's padding).
// The flood fills become: CCL over (candidate, equal value) + OR of the "has a greater neighbour"
// flag per component root.
#include <cfloat>

#include "common.hpp"

namespace nsc {

int ccl_mask(const uint8_t* fg, const float* vals, int32_t* labels, int w, int h, int conn, hipStream_t s);

namespace {

__global__ void k_hmin_prep(const float* __restrict__ img, float hval, float* __restrict__ mask, float* __restrict__ marker,
                            size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float m = -img[i];
        mask[i] = m;
        marker[i] = m - hval;
    }
}

__global__ void k_negate(const float* __restrict__ a, float* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = -a[i];
}

// c = sign * img; marker = c - 1
__global__ void k_lmax_prep(const float* __restrict__ img, float sign, float* __restrict__ c, float* __restrict__ marker,
                            size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = sign < 0 ? -img[i] : img[i];
        c[i] = v;
        marker[i] = v - 1.0f;
    }
}

// cand = recon <= marker (i.e. not marker < recon); greater[i] = some conn-neighbour is > c[i]
template <int CONN>
__global__ void k_lmax_cand(const float* __restrict__ c, const float* __restrict__ marker, const float* __restrict__ rec,
                            uint8_t* __restrict__ cand, uint8_t* __restrict__ greater, int w, int h) {
    const size_t n = (size_t)w * h;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % w), y = (int)(i / w);
        const bool cd = !(marker[i] < rec[i]);
        cand[i] = cd ? 1 : 0;
        bool g = false;
        if (cd) {
            const float v = c[i];
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if ((dx | dy) == 0) continue;
                    if (CONN == 4 && dx != 0 && dy != 0) continue;
                    const int xx = x + dx, yy = y + dy;
                    if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
                    if (v < c[(size_t)yy * w + xx]) g = true;
                }
        }
        greater[i] = g ? 1 : 0;
    }
}

__global__ void k_flag_from(const int32_t* __restrict__ L, const uint8_t* __restrict__ src, uint8_t* flag, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (src[i]) {
            const int r = L[i];
            if (r >= 0) flag[r] = 1;
        }
    }
}

__global__ void k_lmax_out(const int32_t* __restrict__ L, const uint8_t* __restrict__ flag, uint8_t* __restrict__ out,
                           size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int r = L[i];
        out[i] = (r >= 0 && !flag[r]) ? 255 : 0;
    }
}

inline dim3 grid1(size_t n) { return dim3(std::min<unsigned>(cdiv(n, 256), 8192u)); }

}  // namespace

int imhmin_impl(const float* img, float* out, int w, int h, float hval, int conn, float* s1, float* s2, hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(k_hmin_prep, grid1(n), dim3(256), 0, s, img, hval, s1, s2, n);
    NSC_CHECK(recon_impl<float>(s2, s1, out, w, h, conn, s));
    hipLaunchKernelGGL(k_negate, grid1(n), dim3(256), 0, s, out, out, n);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

// scratch: 3 float planes + int32 plane + 3 byte planes
size_t local_max_scratch(int w, int h) { return (size_t)w * h * (3 * sizeof(float) + sizeof(int32_t) + 3); }

int local_max_impl(const float* img, uint8_t* out, int w, int h, int conn, bool minima, void* scratch, hipStream_t s) {
    const size_t n = (size_t)w * h;
    float* c = static_cast<float*>(scratch);
    float* marker = c + n;
    float* rec = marker + n;
    int32_t* L = reinterpret_cast<int32_t*>(rec + n);
    uint8_t* cand = reinterpret_cast<uint8_t*>(L + n);
    uint8_t* greater = cand + n;
    uint8_t* flag = greater + n;
    hipLaunchKernelGGL(k_lmax_prep, grid1(n), dim3(256), 0, s, img, minima ? -1.0f : 1.0f, c, marker, n);
    NSC_CHECK(recon_impl<float>(marker, c, rec, w, h, conn, s));
    if (conn == 8)
        hipLaunchKernelGGL((k_lmax_cand<8>), grid1(n), dim3(256), 0, s, c, marker, rec, cand, greater, w, h);
    else
        hipLaunchKernelGGL((k_lmax_cand<4>), grid1(n), dim3(256), 0, s, c, marker, rec, cand, greater, w, h);
    NSC_CHECK(ccl_mask(cand, c, L, w, h, conn, s));
    NSC_HIP(hipMemsetAsync(flag, 0, n, s));
    hipLaunchKernelGGL(k_flag_from, grid1(n), dim3(256), 0, s, L, greater, flag, n);
    hipLaunchKernelGGL(k_lmax_out, grid1(n), dim3(256), 0, s, L, flag, out, n);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

}  // namespace nsc

using namespace nsc;

extern "C" {

int nsc_imhmin_f32(const float* img, float* out, int w, int h, float hval, int conn, void* stream) {
    NSC_ARG(img && out && w > 0 && h > 0 && valid_conn(conn), "imhmin_f32");
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    Scratch sc;
    NSC_CHECK(sc.alloc(2 * sizeof(float) * (size_t)w * h, s));
    float* a = sc.as<float>();
    return imhmin_impl(img, out, w, h, hval, conn, a, a + (size_t)w * h, s);
}

int nsc_local_maxima_f32(const float* img, uint8_t* out, int w, int h, int conn, void* stream) {
    NSC_ARG(img && out && w > 0 && h > 0 && valid_conn(conn), "local_maxima_f32");
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    Scratch sc;
    NSC_CHECK(sc.alloc(local_max_scratch(w, h), s));
    return local_max_impl(img, out, w, h, conn, false, sc.p, s);
}

int nsc_local_minima_f32(const float* img, uint8_t* out, int w, int h, int conn, void* stream) {
    NSC_ARG(img && out && w > 0 && h > 0 && valid_conn(conn), "local_minima_f32");
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    Scratch sc;
    NSC_CHECK(sc.alloc(local_max_scratch(w, h), s));
    return local_max_impl(img, out, w, h, conn, true, sc.p, s);
}

}  // extern "C"
