// nscale_amd.hpp -- C++ operator surface over the C ABI of libnscale_hip.so, with the names and
// argument meaning of the reference's nscale:: free functions (src/segment/MorphologicOperations.h:
// 32-115) and HistologicalEntities (HistologicalEntities.h:43-66).  Header-only; links against
// libnscale_hip.so.
//
// Differences from the reference's cv::Mat interface, all at the boundary only:
//  - images are views of DEVICE memory (row-major, contiguous) instead of cv::Mat / GpuMat;
//  - outputs are caller-allocated views instead of returned Mats;
//  - every call takes the hipStream_t it is ordered on (the reference's cv::gpu::Stream&);
//  - misuse raises nscale::hip::Error (std::runtime_error) carrying the library's message, where
//    the reference raises cv::Exception from CV_Assert.
#ifndef NSCALE_AMD_HPP_
#define NSCALE_AMD_HPP_

#include <hip/hip_runtime_api.h>

#include <cstdint>
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

// A device image: `data` points to width*height elements (BGR images: 3 bytes per pixel).
template <typename T>
struct Image {
    T* data;
    int width, height;
    Image(T* d, int w, int h) : data(d), width(w), height(h) {}
    template <typename U, typename = std::enable_if_t<std::is_same<const U, T>::value>>
    Image(const Image<U>& o) : data(o.data), width(o.width), height(o.height) {}
};

// nscale::imreconstruct<T>(seeds, image, connectivity)  MorphologicOperations.h:34
template <typename T>
void imreconstruct(Image<const T> seeds, Image<const T> image, Image<T> out, int connectivity, hipStream_t s) {
    static_assert(std::is_same<T, uint8_t>::value || std::is_same<T, float>::value ||
                      std::is_same<T, int32_t>::value,
                  "imreconstruct<T>: T is unsigned char, float or int");
    if (std::is_same<T, uint8_t>::value)
        check(nsc_imreconstruct_u8((const uint8_t*)seeds.data, (const uint8_t*)image.data, (uint8_t*)out.data,
                                   image.width, image.height, connectivity, s),
              "imreconstruct<uchar>");
    else if (std::is_same<T, float>::value)
        check(nsc_imreconstruct_f32((const float*)seeds.data, (const float*)image.data, (float*)out.data,
                                    image.width, image.height, connectivity, s),
              "imreconstruct<float>");
    else
        check(nsc_imreconstruct_i32((const int32_t*)seeds.data, (const int32_t*)image.data, (int32_t*)out.data,
                                    image.width, image.height, connectivity, s),
              "imreconstruct<int>");
}

// nscale::imreconstructBinary<uchar>(seeds, binaryImage, connectivity)  MorphologicOperations.h:61
inline void imreconstructBinary(Image<const uint8_t> seeds, Image<const uint8_t> image, Image<uint8_t> out,
                                int connectivity, hipStream_t s) {
    check(nsc_imreconstruct_binary_u8(seeds.data, image.data, out.data, image.width, image.height, connectivity, s),
          "imreconstructBinary");
}

// nscale::imfill<uchar>(image, seeds, binary=true, connectivity)  MorphologicOperations.h:65
inline void imfill(Image<const uint8_t> image, Image<const uint8_t> seeds, Image<uint8_t> out, int connectivity,
                   hipStream_t s) {
    check(nsc_imfill_u8(image.data, seeds.data, out.data, image.width, image.height, connectivity, s), "imfill");
}

// nscale::imfillHoles<uchar>(image, binary=true, connectivity)  MorphologicOperations.h:68
inline void imfillHoles(Image<const uint8_t> image, Image<uint8_t> out, int connectivity, hipStream_t s) {
    check(nsc_imfill_holes_u8(image.data, out.data, image.width, image.height, connectivity, s), "imfillHoles");
}

// nscale::bwselect<uchar>(binaryImage, seeds, connectivity)  MorphologicOperations.h:73
inline void bwselect(Image<const uint8_t> bw, Image<const uint8_t> seeds, Image<uint8_t> out, int connectivity,
                     hipStream_t s) {
    check(nsc_bwselect_u8(bw.data, seeds.data, out.data, bw.width, bw.height, connectivity, s), "bwselect");
}

// nscale::bwlabel2(binaryImage, connectivity, relab=false)  MorphologicOperations.h:77 (background -1,
// label = smallest raster index of the component)
inline void bwlabel2(Image<const uint8_t> bw, Image<int32_t> labels, int connectivity, hipStream_t s) {
    check(nsc_bwlabel2_u8(bw.data, labels.data, bw.width, bw.height, connectivity, s), "bwlabel2");
}

// nscale::bwareaopen2(image, labeled=false, flatten=true, minSize, maxSize, connectivity, count)
// MorphologicOperations.h:83; keeps components with minSize <= area < maxSize
inline void bwareaopen2(Image<const uint8_t> bw, Image<uint8_t> out, int minSize, int maxSize, int connectivity,
                        int& count, hipStream_t s) {
    check(nsc_bwareaopen2_u8(bw.data, out.data, bw.width, bw.height, minSize, maxSize, connectivity, &count, s),
          "bwareaopen2");
}

// nscale::imhmin<float>(image, h, connectivity)  MorphologicOperations.h:94
inline void imhmin(Image<const float> image, Image<float> out, float h, int connectivity, hipStream_t s) {
    check(nsc_imhmin_f32(image.data, out.data, image.width, image.height, h, connectivity, s), "imhmin<float>");
}

// nscale::localMaxima<float> / localMinima<float>(image, connectivity)  MorphologicOperations.h:108-111
inline void localMaxima(Image<const float> image, Image<uint8_t> out, int connectivity, hipStream_t s) {
    check(nsc_local_maxima_f32(image.data, out.data, image.width, image.height, connectivity, s), "localMaxima");
}
inline void localMinima(Image<const float> image, Image<uint8_t> out, int connectivity, hipStream_t s) {
    check(nsc_local_minima_f32(image.data, out.data, image.width, image.height, connectivity, s), "localMinima");
}

// nscale::morphOpen<uchar>(image, kernel)  MorphologicOperations.h:114; kernel: HOST k x k 0/1 bytes
inline void morphOpen(Image<const uint8_t> image, Image<uint8_t> out, const uint8_t* kernel, int k, hipStream_t s) {
    check(nsc_morph_open_u8(image.data, out.data, image.width, image.height, kernel, k, s), "morphOpen");
}

// cv::distanceTransform(src, dst, CV_DIST_L2, CV_DIST_MASK_PRECISE) as HistologicalEntities.cpp:359 calls it
inline void distanceTransform(Image<const uint8_t> image, Image<float> dist, hipStream_t s) {
    check(nsc_dist_transform_u8(image.data, dist.data, image.width, image.height, s), "distanceTransform");
}

// nscale::watershed2(origImage (BGR), image, connectivity)  MorphologicOperations.h:103
inline void watershed2(Image<const uint8_t> bgr, Image<const float> image, Image<int32_t> out, int connectivity,
                       hipStream_t s) {
    check(nsc_watershed2(bgr.data, image.data, out.data, image.width, image.height, connectivity, s), "watershed2");
}

// HistologicalEntities::segmentNuclei(img, output, compcount, ...)  HistologicalEntities.h:52-54 with
// the reference's default parameters (HistologicalEntities.h:53).  Returns the chain status code.
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
        const int rc = nsc_segment_nuclei(bgr.data, bgr.width, bgr.height, &p, output.data, &compcount, ws, nullptr,
                                          nullptr, s);
        check(rc, "segmentNuclei");
        return rc;
    }
};

}  // namespace hip
}  // namespace nscale

#endif  // NSCALE_AMD_HPP_
