// nscale_amd.hpp -- C++ operator surface over the C ABI of libnscale_hip.so, with the names and
// argument meaning of the reference's nscale:: free functions (src/segment/MorphologicOperations.h:
// 32-115) and HistologicalEntities (HistologicalEntities.h:43-66).  Header-only; links against
// libnscale_hip.so.
//
// Differences from the reference's cv::Mat interface, all at the boundary only:
//  - images are views of DEVICE memory (row-major, with a row step like a GpuMat) instead of
//    cv::Mat / GpuMat; a view that is not continuous is copied to a continuous scratch image for
//    the call, as the reference's GPU operators do (MorphologicOperationsGPU.cpp:180-186);
//  - outputs are caller-allocated views instead of returned Mats;
//  - every call takes the hipStream_t it is ordered on (the reference's cv::gpu::Stream&);
//  - misuse raises nscale::hip::Error (std::runtime_error) carrying the library's message, where
//    the reference raises cv::Exception from CV_Assert.
#ifndef NSCALE_AMD_HPP_
#define NSCALE_AMD_HPP_

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "nscale_hip.h"

namespace nscale {
namespace hip {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};

inline void check(int rc, const char* fn) {
    if (rc <= NSC_ERR_ARG) throw Error(rc, std::string(fn) + ": " + nsc_last_error());
}

// A device image: `data` points to the first of height rows of width elements (BGR images: 3
// bytes per pixel), `step` bytes apart (0 = continuous), like a cv::gpu::GpuMat.
template <typename T>
struct Image {
    T* data;
    int width, height;
    size_t step;
    Image(T* d, int w, int h, size_t st = 0) : data(d), width(w), height(h), step(st) {}
    template <typename U, typename = std::enable_if_t<std::is_same<const U, T>::value>>
    Image(const Image<U>& o) : data(o.data), width(o.width), height(o.height), step(o.step) {}
};

namespace detail {
// The continuous form of a view for one call: the view itself when its rows are adjacent, else a
// stream-ordered scratch copy (copied in for inputs; written back by finish() for outputs).
template <typename T>
class Dense {
public:
    Dense(Image<T> v, int channels, hipStream_t s, bool copy_in) : v_(v), s_(s) {
        row_ = sizeof(T) * channels * (size_t)v.width;
        if (v.step == 0 || v.step == row_) {
            p_ = v.data;
            return;
        }
        if (v.step < row_) throw Error(NSC_ERR_ARG, "image step shorter than a row");
        void* q = nullptr;
        if (hipMallocAsync(&q, row_ * v.height, s) != hipSuccess) throw Error(NSC_ERR_NOMEM, "scratch image");
        owned_ = true;
        p_ = static_cast<T*>(q);
        if (copy_in &&
            hipMemcpy2DAsync(q, row_, v.data, v.step, row_, v.height, hipMemcpyDeviceToDevice, s) != hipSuccess)
            throw Error(NSC_ERR_HIP, "pitched copy");
    }
    ~Dense() {
        if (owned_) (void)hipFreeAsync(const_cast<void*>(static_cast<const void*>(p_)), s_);
    }
    Dense(const Dense&) = delete;
    Dense& operator=(const Dense&) = delete;
    T* get() const { return p_; }
    // an output view receives the continuous result
    void finish() {
        if (owned_ && hipMemcpy2DAsync(const_cast<void*>(static_cast<const void*>(v_.data)), v_.step, p_, row_, row_,
                                       v_.height, hipMemcpyDeviceToDevice, s_) != hipSuccess)
            throw Error(NSC_ERR_HIP, "pitched copy");
    }

private:
    Image<T> v_;
    hipStream_t s_;
    size_t row_ = 0;
    T* p_ = nullptr;
    bool owned_ = false;
};
template <typename T>
Dense<const T> in(Image<const T> v, hipStream_t s, int channels = 1) { return Dense<const T>(v, channels, s, true); }
template <typename T>
Dense<T> out(Image<T> v, hipStream_t s, int channels = 1) { return Dense<T>(v, channels, s, false); }
}  // namespace detail

// nscale::imreconstruct<T>(seeds, image, connectivity)  MorphologicOperations.h:34
template <typename T>
void imreconstruct(Image<const T> seeds, Image<const T> image, Image<T> out, int connectivity, hipStream_t s) {
    static_assert(std::is_same<T, uint8_t>::value || std::is_same<T, float>::value ||
                      std::is_same<T, int32_t>::value,
                  "imreconstruct<T>: T is unsigned char, float or int");
    auto m = detail::in(seeds, s), k = detail::in(image, s);
    auto o = detail::out(out, s);
    if (std::is_same<T, uint8_t>::value)
        check(nsc_imreconstruct_u8((const uint8_t*)m.get(), (const uint8_t*)k.get(), (uint8_t*)o.get(), image.width,
                                   image.height, connectivity, s),
              "imreconstruct<uchar>");
    else if (std::is_same<T, float>::value)
        check(nsc_imreconstruct_f32((const float*)m.get(), (const float*)k.get(), (float*)o.get(), image.width,
                                    image.height, connectivity, s),
              "imreconstruct<float>");
    else
        check(nsc_imreconstruct_i32((const int32_t*)m.get(), (const int32_t*)k.get(), (int32_t*)o.get(), image.width,
                                    image.height, connectivity, s),
              "imreconstruct<int>");
    o.finish();
}

// nscale::imreconstructBinary<uchar>(seeds, binaryImage, connectivity)  MorphologicOperations.h:61
inline void imreconstructBinary(Image<const uint8_t> seeds, Image<const uint8_t> image, Image<uint8_t> out,
                                int connectivity, hipStream_t s) {
    auto m = detail::in(seeds, s), k = detail::in(image, s);
    auto o = detail::out(out, s);
    check(nsc_imreconstruct_binary_u8(m.get(), k.get(), o.get(), image.width, image.height, connectivity, s),
          "imreconstructBinary");
    o.finish();
}

// nscale::imfill<uchar>(image, seeds, binary=true, connectivity)  MorphologicOperations.h:65
inline void imfill(Image<const uint8_t> image, Image<const uint8_t> seeds, Image<uint8_t> out, int connectivity,
                   hipStream_t s) {
    auto k = detail::in(image, s), m = detail::in(seeds, s);
    auto o = detail::out(out, s);
    check(nsc_imfill_u8(k.get(), m.get(), o.get(), image.width, image.height, connectivity, s), "imfill");
    o.finish();
}

// nscale::imfillHoles<uchar>(image, binary=true, connectivity)  MorphologicOperations.h:68
inline void imfillHoles(Image<const uint8_t> image, Image<uint8_t> out, int connectivity, hipStream_t s) {
    auto k = detail::in(image, s);
    auto o = detail::out(out, s);
    check(nsc_imfill_holes_u8(k.get(), o.get(), image.width, image.height, connectivity, s), "imfillHoles");
    o.finish();
}

// nscale::bwselect<uchar>(binaryImage, seeds, connectivity)  MorphologicOperations.h:73
inline void bwselect(Image<const uint8_t> bw, Image<const uint8_t> seeds, Image<uint8_t> out, int connectivity,
                     hipStream_t s) {
    auto k = detail::in(bw, s), m = detail::in(seeds, s);
    auto o = detail::out(out, s);
    check(nsc_bwselect_u8(k.get(), m.get(), o.get(), bw.width, bw.height, connectivity, s), "bwselect");
    o.finish();
}

// nscale::bwlabel2(binaryImage, connectivity, relab=false)  MorphologicOperations.h:77 (background -1,
// label = smallest raster index of the component)
inline void bwlabel2(Image<const uint8_t> bw, Image<int32_t> labels, int connectivity, hipStream_t s) {
    auto k = detail::in(bw, s);
    auto o = detail::out(labels, s);
    check(nsc_bwlabel2_u8(k.get(), o.get(), bw.width, bw.height, connectivity, s), "bwlabel2");
    o.finish();
}

// nscale::bwareaopen2(image, labeled=false, flatten=true, minSize, maxSize, connectivity, count)
// MorphologicOperations.h:83; keeps components with minSize <= area < maxSize
inline void bwareaopen2(Image<const uint8_t> bw, Image<uint8_t> out, int minSize, int maxSize, int connectivity,
                        int& count, hipStream_t s) {
    auto k = detail::in(bw, s);
    auto o = detail::out(out, s);
    check(nsc_bwareaopen2_u8(k.get(), o.get(), bw.width, bw.height, minSize, maxSize, connectivity, &count, s),
          "bwareaopen2");
    o.finish();
}

// nscale::imhmin<float>(image, h, connectivity)  MorphologicOperations.h:94
inline void imhmin(Image<const float> image, Image<float> out, float h, int connectivity, hipStream_t s) {
    auto k = detail::in(image, s);
    auto o = detail::out(out, s);
    check(nsc_imhmin_f32(k.get(), o.get(), image.width, image.height, h, connectivity, s), "imhmin<float>");
    o.finish();
}

// nscale::localMaxima<float> / localMinima<float>(image, connectivity)  MorphologicOperations.h:108-111
inline void localMaxima(Image<const float> image, Image<uint8_t> out, int connectivity, hipStream_t s) {
    auto k = detail::in(image, s);
    auto o = detail::out(out, s);
    check(nsc_local_maxima_f32(k.get(), o.get(), image.width, image.height, connectivity, s), "localMaxima");
    o.finish();
}
inline void localMinima(Image<const float> image, Image<uint8_t> out, int connectivity, hipStream_t s) {
    auto k = detail::in(image, s);
    auto o = detail::out(out, s);
    check(nsc_local_minima_f32(k.get(), o.get(), image.width, image.height, connectivity, s), "localMinima");
    o.finish();
}

// nscale::morphOpen<uchar>(image, kernel)  MorphologicOperations.h:114; kernel: HOST k x k 0/1 bytes
inline void morphOpen(Image<const uint8_t> image, Image<uint8_t> out, const uint8_t* kernel, int k, hipStream_t s) {
    auto a = detail::in(image, s);
    auto o = detail::out(out, s);
    check(nsc_morph_open_u8(a.get(), o.get(), image.width, image.height, kernel, k, s), "morphOpen");
    o.finish();
}

// cv::distanceTransform(src, dst, CV_DIST_L2, CV_DIST_MASK_PRECISE) as HistologicalEntities.cpp:359 calls it
inline void distanceTransform(Image<const uint8_t> image, Image<float> dist, hipStream_t s) {
    auto a = detail::in(image, s);
    auto o = detail::out(dist, s);
    check(nsc_dist_transform_u8(a.get(), o.get(), image.width, image.height, s), "distanceTransform");
    o.finish();
}

// nscale::watershed2(origImage (BGR), image, connectivity)  MorphologicOperations.h:103
inline void watershed2(Image<const uint8_t> bgr, Image<const float> image, Image<int32_t> out, int connectivity,
                       hipStream_t s) {
    auto c = detail::in(bgr, s, 3);
    auto a = detail::in(image, s);
    auto o = detail::out(out, s);
    check(nsc_watershed2(c.get(), a.get(), o.get(), image.width, image.height, connectivity, s), "watershed2");
    o.finish();
}

// ConnComponents::boundingBox(w, h, label, bgval, compcount)  ConnComponents.h:43: the reference's
// literal scan on a DEVICE label image; returns a HOST int[5 * compcount] block allocated with
// malloc (the caller frees it with free, as nu-segment.cpp:189 does), or NULL when compcount is 0.
inline int* boundingBox(Image<const int32_t> labels, int bgval, int& compcount, hipStream_t s) {
    auto k = detail::in(labels, s);
    int n = 0;
    check(nsc_cc_bounding_box(k.get(), labels.width, labels.height, bgval, nullptr, 0, &n, s), "boundingBox");
    compcount = n;
    if (n == 0) return nullptr;
    void* d = nullptr;
    if (hipMallocAsync(&d, sizeof(int32_t) * 5 * (size_t)n, s) != hipSuccess) throw Error(NSC_ERR_NOMEM, "boundingBox");
    const int rc = nsc_cc_bounding_box(k.get(), labels.width, labels.height, bgval, (int32_t*)d, n, &n, s);
    int* h = rc == NSC_OK ? static_cast<int*>(std::malloc(sizeof(int) * 5 * (size_t)n)) : nullptr;
    if (h) (void)hipMemcpyAsync(h, d, sizeof(int) * 5 * (size_t)n, hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    (void)hipFreeAsync(d, s);
    check(rc, "boundingBox");
    if (!h) throw Error(NSC_ERR_NOMEM, "boundingBox");
    return h;
}

// nscale::gpu::boundingBox2(labels, compcount, stream)  HistologicalEntities.h:102-103: one box per
// label of a DEVICE image of canonical labels 1..n (background 0) in a DEVICE block of 5 * n ints
// [label | minx | maxx | miny | maxy] allocated with hipMalloc; the caller releases it with
// freeData (the reference's cudaFreeData, HistologicalEntities.h:104).
inline int32_t* boundingBox2(Image<const int32_t> canon_labels, int n, hipStream_t s) {
    auto k = detail::in(canon_labels, s);
    if (n <= 0) return nullptr;
    void* d = nullptr;
    if (hipMalloc(&d, sizeof(int32_t) * 5 * (size_t)n) != hipSuccess) throw Error(NSC_ERR_NOMEM, "boundingBox2");
    const int rc = nsc_bounding_boxes(k.get(), canon_labels.width, canon_labels.height, n, (int32_t*)d, nullptr, s);
    if (rc != NSC_OK) (void)hipFree(d);
    check(rc, "boundingBox2");
    return static_cast<int32_t*>(d);
}
inline void freeData(void* device_ptr) { (void)hipFree(device_ptr); }

// HistologicalEntities::segmentNuclei(img, output, compcount, bbox, ...)  HistologicalEntities.h:52-54,
// plFindNucleusCandidates (S1, :63-64) and plSeparateNuclei (A4, :65-66), with the reference's
// default parameters (HistologicalEntities.h:53).  Return the chain status codes.
class HistologicalEntities {
public:
    static nsc_seg_params defaults() {
        nsc_seg_params p;
        nsc_seg_params_default(&p);
        return p;
    }
    static int segmentNuclei(Image<const uint8_t> bgr, Image<int32_t> output, int& compcount, nsc_workspace* ws,
                             hipStream_t s, const nsc_seg_params* params = nullptr) {
        const nsc_seg_params p = params ? *params : defaults();
        nsc_seg_io io{};
        io.bgr_step = bgr.step;
        io.labels_step = output.step;
        const int rc = nsc_segment_nuclei_io(bgr.data, bgr.width, bgr.height, &p, output.data, &io, ws, nullptr,
                                             nullptr, s);
        check(rc, "segmentNuclei");
        compcount = io.nuclei;
        return rc;
    }
    // With the reference's compcount / bbox outputs: bbox_mode NSC_BBOX_REFERENCE_CPU reproduces
    // the CPU chain's ConnComponents::boundingBox(output, 0) (HistologicalEntities.cpp:590-591),
    // NSC_BBOX_ALL the GPU chain's boundingBox2 (one box per nucleus).  bbox: a HOST int[5 *
    // compcount] block from malloc (NULL when compcount is 0), freed by the caller with free.
    static int segmentNuclei(Image<const uint8_t> bgr, Image<int32_t> output, int& compcount, int*& bbox,
                             int bbox_mode, nsc_workspace* ws, hipStream_t s, const nsc_seg_params* params = nullptr) {
        const nsc_seg_params p = params ? *params : defaults();
        const int cap = bgr.width * bgr.height / 8 + 64;  // more boxes than a tile of nuclei can hold
        void* d = nullptr;
        if (hipMallocAsync(&d, sizeof(int32_t) * 5 * (size_t)cap, s) != hipSuccess)
            throw Error(NSC_ERR_NOMEM, "segmentNuclei bbox");
        nsc_seg_io io{};
        io.bgr_step = bgr.step;
        io.labels_step = output.step;
        io.bbox_mode = bbox_mode;
        io.bbox5 = static_cast<int32_t*>(d);
        io.max_boxes = cap;
        const int rc = nsc_segment_nuclei_io(bgr.data, bgr.width, bgr.height, &p, output.data, &io, ws, nullptr,
                                             nullptr, s);
        bbox = nullptr;
        compcount = rc == NSC_SEG_SUCCESS ? io.compcount : 0;
        if (rc == NSC_SEG_SUCCESS && compcount > 0) {
            bbox = static_cast<int*>(std::malloc(sizeof(int) * 5 * (size_t)compcount));
            if (bbox) (void)hipMemcpyAsync(bbox, d, sizeof(int) * 5 * (size_t)compcount, hipMemcpyDeviceToHost, s);
            (void)hipStreamSynchronize(s);
        }
        (void)hipFreeAsync(d, s);
        check(rc, "segmentNuclei");
        return rc;
    }
    static int plFindNucleusCandidates(Image<const uint8_t> bgr, Image<uint8_t> seg_norbc, nsc_workspace* ws,
                                       hipStream_t s, const nsc_seg_params* params = nullptr) {
        const nsc_seg_params p = params ? *params : defaults();
        const int rc = nsc_pl_find_nucleus_candidates(bgr.data, bgr.step, bgr.width, bgr.height, &p, seg_norbc.data,
                                                      seg_norbc.step, ws, s);
        check(rc, "plFindNucleusCandidates");
        return rc;
    }
    static int plSeparateNuclei(Image<const uint8_t> bgr, Image<const uint8_t> seg_open, Image<uint8_t> seg_nonoverlap,
                                nsc_workspace* ws, hipStream_t s, int minSizePl = 30, int watershedConnectivity = 8) {
        const int rc = nsc_pl_separate_nuclei(bgr.data, bgr.step, seg_open.data, seg_open.step, bgr.width, bgr.height,
                                              minSizePl, watershedConnectivity, seg_nonoverlap.data, seg_nonoverlap.step,
                                              ws, s);
        check(rc, "plSeparateNuclei");
        return rc;
    }
};

// nscale::Normalization (normalization/Normalization.h:30-47): Reinhard colour transfer of the
// foreground of a BGR image to target l-alpha-beta statistics, and those statistics of an image.
class Normalization {
   public:
    // writes the normalized image to `out` (may be `originalI` itself); asynchronous on s
    static void normalization(Image<const uint8_t> originalI, const float targetMean[3], const float targetStd[3],
                              Image<uint8_t> out, hipStream_t s) {
        const size_t ld = originalI.step ? originalI.step : (size_t)3 * originalI.width;
        const size_t old = out.step ? out.step : (size_t)3 * out.width;
        check(nsc_normalization(originalI.data, originalI.width, originalI.height, ld, targetMean, targetStd, out.data,
                                old, s),
              "normalization");
    }
    // synchronizes s
    static void targetParameters(Image<const uint8_t> originalI, float (&targetMean)[3], float (&targetStd)[3],
                                 hipStream_t s) {
        const size_t ld = originalI.step ? originalI.step : (size_t)3 * originalI.width;
        check(nsc_normalization_target_parameters(originalI.data, originalI.width, originalI.height, ld, targetMean,
                                                  targetStd, nullptr, s),
              "targetParameters");
    }
};

}  // namespace hip
}  // namespace nscale

#endif  // NSCALE_AMD_HPP_
