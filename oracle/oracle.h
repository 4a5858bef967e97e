/*
 * CPU oracle for the nscale segmentNuclei morphology chain.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, single-threaded and in plain C++,
 * the algorithms of nscale's CPU path (the src/segment C++ sources) and of the OpenCV 2.4.9
 * routines that path calls (cv::erode/dilate, cv::distanceTransform(CV_DIST_MASK_PRECISE),
 * cv::floodFill, cv::watershed).  It is the checker for the HIP implementation and the
 * "port" CPU baseline of bench.py; nothing in the product links or calls it.
 *
 * Every function follows the reference line by line in semantics (not in code); the
 * file:line it follows is given at each definition in oracle.cpp.  Images are row-major,
 * contiguous, width w, height h; BGR images are interleaved 3 bytes per pixel.
 */
#ifndef NSCALE_ORACLE_H_
#define NSCALE_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* grayscale reconstruction by dilation (Vincent hybrid: raster, anti-raster, FIFO) */
void orc_imreconstruct_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h, int conn);
void orc_imreconstruct_f32(const float* marker, const float* mask, float* out, int w, int h, int conn);
void orc_imreconstruct_i32(const int32_t* marker, const int32_t* mask, int32_t* out, int w, int h, int conn);
/* binary reconstruction (contour seeds + FIFO), u8 values, nonzero = foreground */
void orc_imreconstruct_binary_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h, int conn);
/* imfill(image, seeds) (seeded fill), binary variant */
void orc_imfill_u8(const uint8_t* img, const uint8_t* seeds, uint8_t* out, int w, int h, int conn);
/* imfillHoles(img, binary=true, conn) */
void orc_imfill_holes_u8(const uint8_t* img, uint8_t* out, int w, int h, int conn);
/* imfillHoles(img, binary=false, conn): grayscale hole filling via imreconstruct */
void orc_imfill_holes_gray_u8(const uint8_t* img, uint8_t* out, int w, int h, int conn);
/* bwselect(bw, seeds, conn) */
void orc_bwselect_u8(const uint8_t* bw, const uint8_t* seeds, uint8_t* out, int w, int h, int conn);
/* ConnComponents::label: union-find on equal nonzero values, root = min raster index, bg = bgval */
void orc_cc_label(const uint8_t* img, int w, int h, int32_t* label, int32_t bgval, int conn);
/* ConnComponents::relabel: roots (label[i]==i) numbered 1..N in raster order, bg -> 0; returns N */
int orc_cc_relabel(int w, int h, int32_t* label, int32_t bgval);
/* ConnComponents::areaThresholdLabeled: keep labels with lower <= area < upper; returns kept count */
int orc_cc_area_threshold_labeled(const int32_t* label, int w, int h, int32_t* n_label, int32_t bgval, int lower, int upper);
/* ConnComponents::boundingBox (literal restatement, including its early 'break' on -1/0 labels).
   Writes up to cap entries per column of an n x 5 SoA block; returns compcount. */
int orc_cc_bounding_box_literal(int w, int h, const int32_t* label, int32_t bgval, int32_t* out5, int cap);
/* bwareaopen2(img, labeled=false, flatten=true, min, max, conn, count) -> u8 255/0 */
int orc_bwareaopen2_u8(const uint8_t* img, uint8_t* out, int w, int h, int minSize, int maxSize, int conn);
/* bwlabel2(img, conn, relab): labels = min raster index, bg -1 */
void orc_bwlabel2(const uint8_t* img, int32_t* labels, int w, int h, int conn);
/* cv::erode / cv::dilate with a k x k structuring element (anchor at centre) and an explicit
   constant for pixels outside the image */
void orc_erode_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside);
void orc_dilate_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside);
/* morphOpen<uchar>(img, kernel) */
void orc_morph_open_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k);
/* cv::distanceTransform(src, dst, CV_DIST_L2, CV_DIST_MASK_PRECISE) */
void orc_dist_transform_u8(const uint8_t* img, float* out, int w, int h);
/* imhmin<float>(img, hval, conn) */
void orc_imhmin_f32(const float* img, float* out, int w, int h, float hval, int conn);
/* localMaxima<float>, localMinima<float> -> u8 255/0 */
void orc_local_maxima_f32(const float* img, uint8_t* out, int w, int h, int conn);
void orc_local_minima_f32(const float* img, uint8_t* out, int w, int h, int conn);
/* cv::watershed on a BGR image and an int32 marker image (in/out), same size */
void orc_cv_watershed(const uint8_t* bgr, int32_t* markers, int w, int h);
/* NeighborOperations::border<int>(padded, background) -> result of the same size; rows 0 and
   columns 0, w-1 are not written (the reference leaves them uninitialised) */
void orc_border_i32(const int32_t* padded, int32_t* out, int w, int h, int32_t background);
/* watershed2(bgr, float image, conn) -> int32 */
void orc_watershed2(const uint8_t* bgr, const float* img, int32_t* out, int w, int h, int conn);
/* getBackground / getRBC */
void orc_get_background(const uint8_t* bgr, uint8_t* out, int w, int h, uint8_t blue, uint8_t green, uint8_t red);
void orc_get_rbc(const uint8_t* bgr, uint8_t* out, int w, int h, double T1, double T2);

typedef struct orc_seg_params {
    uint8_t blue, green, red;
    double T1, T2;
    uint8_t G1;
    int minSize, maxSize;
    uint8_t G2;
    int minSizePl, minSizeSeg, maxSizeSeg;
    int fillHolesConnectivity, reconConnectivity, watershedConnectivity;
} orc_seg_params;

void orc_seg_params_default(orc_seg_params* p);

/* HistologicalEntities::segmentNuclei(img, output, compcount, bbox, ...).
   labels: w*h int32 (min-raster-index labels, bg -1).  stage[i] (may be NULL, or have NULL
   entries) receives the intermediate with stage id i (ids as in HistologicalEntities.cpp;
   u8 planes except 16/17/18 float, 19 BGR u8, 20 int32).  stage_ms (may be NULL) receives
   per-stage wall times in milliseconds, indexed by the same ids.  Returns the status code. */
int orc_segment_nuclei(const uint8_t* bgr, int w, int h, const orc_seg_params* p, int32_t* labels,
                       void** stage, double* stage_ms);
/* segmentNuclei's compcount and bbox outputs (HistologicalEntities.cpp:589-591): the literal
   ConnComponents::boundingBox(w, h, labels, 0, compcount) on the final labels.  bbox5 (may be
   NULL) receives 5 x compcount ints, column stride compcount, when compcount <= cap. */
int orc_segment_nuclei_bbox(const uint8_t* bgr, int w, int h, const orc_seg_params* p, int32_t* labels, int* compcount,
                            int32_t* bbox5, int cap, void** stage, double* stage_ms);


// colour normalization (src/normalization/Normalization.cpp)
void orc_norm_seg_fg(const uint8_t* bgr, uint8_t* out, int w, int h);
void orc_norm_bgr2lab(const uint8_t* bgr, float* lab, int w, int h);
void orc_norm_lab2bgr(const float* lab, uint8_t* bgr, int w, int h);
long long orc_norm_target_parameters(const uint8_t* bgr, int w, int h, float* mean3, float* std3);
void orc_norm_normalization(const uint8_t* bgr, int w, int h, const float* target_mean, const float* target_std,
                            uint8_t* out);

#ifdef __cplusplus
}
#endif
#endif
