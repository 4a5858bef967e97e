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
#include <algorithm>
#include <cfloat>
#include <vector>

#include "common.hpp"

namespace nsc {

int ccl_mask(const uint8_t* fg, const float* vals, int32_t* labels, int w, int h, int conn, hipStream_t s);
int ccl_mask_clear_root_flags(const uint8_t* fg, const float* vals, int32_t* labels, int w, int h, int conn,
                              uint8_t* flags, hipStream_t s);

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
    const auto gc = as_global(c);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % w), y = (int)(i / w);
        // all nine values from clamped (always valid) addresses, issued together; a neighbour
        // outside the image reads the pixel itself, which is never greater
        const int xl = x > 0 ? x - 1 : x, xr = x + 1 < w ? x + 1 : x;
        const size_t ru = (size_t)(y > 0 ? y - 1 : y) * w, r0 = (size_t)y * w, rd = (size_t)(y + 1 < h ? y + 1 : y) * w;
        const float mk = marker[i], rc = rec[i], v = gc[i];
        const float n0 = gc[r0 + xl], n1 = gc[r0 + xr], n2 = gc[ru + x], n3 = gc[rd + x];
        bool g = (v < n0) | (v < n1) | (v < n2) | (v < n3);
        if (CONN == 8) {
            const float n4 = gc[ru + xl], n5 = gc[ru + xr], n6 = gc[rd + xl], n7 = gc[rd + xr];
            g = g | (v < n4) | (v < n5) | (v < n6) | (v < n7);
        }
        const bool cd = !(mk < rc);
        cand[i] = cd ? 1 : 0;
        greater[i] = (cd && g) ? 1 : 0;
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

// batched h-minima: pixels of all images in one flat index space (image found by binary search
// over the offsets)
struct HB {
    const float* img;
    float* mask;
    float* marker;
    float* out;
    long long off, n;
};
__device__ __forceinline__ int hb_find(const HB* __restrict__ b, int nb, long long i) {
    int lo = 0, hi = nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b[mid].off <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__global__ void k_hmin_prep_b(const HB* __restrict__ b, int nb, long long total, float hval) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const HB& e = b[hb_find(b, nb, i)];
        const long long k = i - e.off;
        const float m = -e.img[k];
        e.mask[k] = m;
        e.marker[k] = m - hval;
    }
}
__global__ void k_hmin_post_b(const HB* __restrict__ b, int nb, long long total) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (long long)gridDim.x * blockDim.x) {
        const HB& e = b[hb_find(b, nb, i)];
        const long long k = i - e.off;
        e.out[k] = -e.out[k];
    }
}

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
    NSC_CHECK(ccl_mask_clear_root_flags(cand, c, L, w, h, conn, flag, s));  // flag[] is read at roots only
    hipLaunchKernelGGL(k_flag_from, grid1(n), dim3(256), 0, s, L, greater, flag, n);
    hipLaunchKernelGGL(k_lmax_out, grid1(n), dim3(256), 0, s, L, flag, out, n);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

}  // namespace nsc

using namespace nsc;

extern "C" {

int nsc_imhmin_f32_batched(const nsc_recon_desc* descs, int n, float hval, void* stream) {
    NSC_ARG(descs && n > 0, "imhmin_f32_batched");
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    std::vector<HB> hb(n);
    std::vector<nsc_recon_desc> rd(n);
    long long total = 0;
    for (int i = 0; i < n; ++i) {
        NSC_ARG(descs[i].mask && descs[i].out && descs[i].w > 0 && descs[i].h > 0 && valid_conn(descs[i].conn),
                "imhmin_f32_batched: descriptor");
        hb[i].img = static_cast<const float*>(descs[i].mask);
        hb[i].out = static_cast<float*>(descs[i].out);
        hb[i].off = total;
        hb[i].n = (long long)descs[i].w * descs[i].h;
        total += hb[i].n;
    }
    Scratch planes, tab;
    NSC_CHECK(planes.alloc(2 * sizeof(float) * (size_t)total + 512 * (size_t)n, s));
    char* p = static_cast<char*>(planes.p);
    for (int i = 0; i < n; ++i) {
        const size_t b = ((sizeof(float) * (size_t)hb[i].n + 255) / 256) * 256;
        hb[i].mask = reinterpret_cast<float*>(p); p += b;
        hb[i].marker = reinterpret_cast<float*>(p); p += b;
        rd[i] = nsc_recon_desc{hb[i].marker, hb[i].mask, descs[i].out, descs[i].w, descs[i].h, descs[i].conn};
    }
    NSC_CHECK(tab.alloc(sizeof(HB) * n, s));
    NSC_HIP(hipMemcpyAsync(tab.p, hb.data(), sizeof(HB) * n, hipMemcpyHostToDevice, s));
    const unsigned g = (unsigned)std::min<long long>((total + 255) / 256, 16384);
    hipLaunchKernelGGL(k_hmin_prep_b, dim3(g), dim3(256), 0, s, tab.as<HB>(), n, total, hval);
    NSC_CHECK(recon_batched_impl<float>(rd.data(), n, s));
    hipLaunchKernelGGL(k_hmin_post_b, dim3(g), dim3(256), 0, s, tab.as<HB>(), n, total);
    NSC_HIP(hipGetLastError());
    NSC_HIP(stream_sync(s));  // host tables above are released on return
    return NSC_OK;
}

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
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    Scratch sc;
    NSC_CHECK(sc.alloc(local_max_scratch(w, h), s));
    return local_max_impl(img, out, w, h, conn, false, sc.p, s);
}

int nsc_local_minima_f32(const float* img, uint8_t* out, int w, int h, int conn, void* stream) {
    NSC_ARG(img && out && w > 0 && h > 0 && valid_conn(conn), "local_minima_f32");
    NSC_LABEL_SIZE(w, h);
    NSC_CHECK(ensure_init());
    hipStream_t s = as_stream(stream);
    Scratch sc;
    NSC_CHECK(sc.alloc(local_max_scratch(w, h), s));
    return local_max_impl(img, out, w, h, conn, true, sc.p, s);
}

}  // extern "C"
