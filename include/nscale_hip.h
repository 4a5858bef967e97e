/*
 * nscale_hip.h -- C ABI of libnscale_hip.so, the MI355X (gfx950) implementation of nscale's
 * nuclei-segmentation morphology chain.
 *
 * Conventions
 *  - Image arguments are DEVICE pointers to row-major, contiguous planes of width w and height
 *    h (BGR images: 3 interleaved bytes per pixel, OpenCV order).  Outputs are caller-allocated.
 *  - `stream` is a hipStream_t passed as void* (NULL = the legacy default stream).  Every
 *    operator is stream-ordered; an operator that returns a host scalar (a count, a status)
 *    synchronises that stream before returning.
 *  - Scratch memory comes from HIP's stream-ordered pool (hipMallocAsync) on `stream`, so the
 *    steady state performs no hipMalloc/hipFree.
 *  - Return value: NSC_OK (0) or a negative NSC_ERR_* code; nothing throws across the ABI.
 *    nsc_segment_nuclei returns the reference's chain status codes (>= 0) on success.
 *  - Connectivity arguments are 4 or 8.
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * reference's src/segment/).
 */
#ifndef NSCALE_HIP_H_
#define NSCALE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NSC_OK 0
#define NSC_ERR_ARG (-10)     /* bad argument (null pointer, size, connectivity) */
#define NSC_ERR_HIP (-11)     /* a HIP runtime call failed */
#define NSC_ERR_NOMEM (-12)   /* device allocation failed */
#define NSC_ERR_LIMIT (-13)   /* an internal capacity was exceeded */

/* chain status codes: HistologicalEntities.h:34-40 */
#define NSC_SEG_SUCCESS 0
#define NSC_SEG_BACKGROUND 1
#define NSC_SEG_BACKGROUND_LIKELY 2
#define NSC_SEG_NO_CANDIDATES_LEFT 3
#define NSC_SEG_CONTINUE 4
#define NSC_SEG_INVALID_IMAGE (-1)
#define NSC_SEG_RUNTIME_FAILED (-2)

/* ---- library ---------------------------------------------------------------------------- */
const char* nsc_version(void);
/* last error message of the calling thread (empty string if none) */
const char* nsc_last_error(void);
/* initialise the library on the current device (optional; done lazily otherwise) */
int nsc_init(void);

/* ---- synthetic inputs (HOST pointers; deterministic in (seed, w, h)) --------------------- */
int nsc_synth_mask(uint8_t* out, int w, int h, uint64_t seed);
int nsc_synth_marker(const uint8_t* mask, uint8_t* marker, size_t n);
int nsc_synth_he(uint8_t* bgr, int w, int h, uint64_t seed);
/* Tile (x0, y0, w, h) of a procedural whole-slide H&E image (DEVICE output, BGR): every pixel is
   a function of (x, y, seed) only, so overlapping tiles agree; stain intensity varies by +-15%
   over a 2000-px gradient-noise field.  Stream-ordered. */
int nsc_synth_slide_tile(uint8_t* bgr, int x0, int y0, int w, int h, uint64_t seed, void* stream);

/* ---- morphological reconstruction ------------------------------------------------------- */
/* nscale::imreconstruct<uchar>(seeds, image, conn)  MorphologicOperations.h:34, .cpp:1730-1903 */
int nsc_imreconstruct_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h, int conn,
                         void* stream);
/* nscale::imreconstruct<float>(seeds, image, conn)  MorphologicOperations.h:34, .cpp:1730-1903 */
int nsc_imreconstruct_f32(const float* marker, const float* mask, float* out, int w, int h, int conn,
                          void* stream);
/* nscale::imreconstruct<int>(seeds, image, conn)  MorphologicOperations.h:34, .cpp:1730-1903 */
int nsc_imreconstruct_i32(const int32_t* marker, const int32_t* mask, int32_t* out, int w, int h, int conn,
                          void* stream);
/* nscale::imreconstruct<uchar> without any allocation.  No counterpart in the reference: for
   callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  The pass driver's scratch is a fixed device array of the
   library (about 0.8 MB per GPU, held until the process ends), so images are limited to 65536
   tiles of 62 x 62 pixels (e.g. 15872 x 15872); larger ones return NSC_ERR_LIMIT before anything
   is launched.  Every call on one GPU shares that array: calls that overlap on one GPU (different
   streams or threads) race and must be serialised by the caller.  Same results as
   nsc_imreconstruct_u8. */
int nsc_imreconstruct_u8_noalloc(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h, int conn,
                                 void* stream);
/* Batched reconstruction: n independent (marker, mask) pairs of any sizes and connectivity in
   one launch per pass (the reference calls imreconstruct once per tile, nu-segment.cpp:175).
   Results are identical to n single calls.  marker/mask/out are device pointers of the element
   type of the entry point (uint8 or float). */
typedef struct nsc_recon_desc {
    const void* marker;
    const void* mask;
    void* out;
    int w, h, conn;
} nsc_recon_desc;
int nsc_imreconstruct_u8_batched(const nsc_recon_desc* descs, int n, void* stream);
int nsc_imreconstruct_f32_batched(const nsc_recon_desc* descs, int n, void* stream);
/* A batched reconstruction as a replayable hipGraph (no reference counterpart: for callers that
   replay one batch layout many times or embed it in their own graphs).  Create captures a graph of
   `passes_per_replay` window passes (uint8, or float when is_float) over the images of descs --
   the same buffers every run, the contents may change between runs; the descriptors' buffers must
   stay allocated until destroy.  Run resets the pass state with kernels (no runtime fills are
   recorded), replays the graph until no tile is left marked (one host check per replay) and
   synchronizes the stream; *passes (may be NULL) receives the passes launched.  Results equal
   nsc_imreconstruct_u8_batched / _f32_batched on the same inputs. */
typedef struct nsc_recon_graph nsc_recon_graph;
int nsc_recon_graph_create(const nsc_recon_desc* descs, int n, int is_float, int passes_per_replay,
                           nsc_recon_graph** graph);
int nsc_recon_graph_run(nsc_recon_graph* graph, int* passes, void* stream);
int nsc_recon_graph_destroy(nsc_recon_graph* graph);
/* Autotuner of the reconstruction schedule (no reference counterpart; the reference has fixed
   shapes): times each candidate -- windows per workgroup side 1 or 2 with 4 or 8 passes between
   host checks, and the device-side work queue for the residual wavefront -- on this batch (outputs
   are written each time) and keeps the fastest for the calling thread's later calls of the same
   element type (is_float 0: uint8, 1: float); other threads keep their own schedule.  region /
   passes (may be NULL) receive the choice (passes 0: the work queue), timings (may be NULL, 5
   floats) the measured ms per candidate in that order. */
int nsc_recon_autotune(const nsc_recon_desc* descs, int n, int is_float, void* stream, int* region, int* passes,
                       float* timings);
/* Sets the calling thread's reconstruction schedule explicitly (no reference counterpart): region
   side 1..4 (0 = default), passes between host checks (0 = default), queue 1 = device-side work
   queue after the first pass, 0 = host-launched passes, -1 = the NSC_RECON_DRAIN default. */
int nsc_recon_set_schedule(int is_float, int region, int passes, int queue);
/* nscale::imhmin<float>(image, h, conn) on n images (MorphologicOperations.cpp:3342-3364):
   descs[i].mask = input image, descs[i].out = output, descs[i].marker unused */
int nsc_imhmin_f32_batched(const nsc_recon_desc* descs, int n, float hval, void* stream);
/* nscale::imreconstructBinary<uchar>(seeds, image, conn)  MorphologicOperations.h:61, .cpp:2768-2909 */
int nsc_imreconstruct_binary_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h,
                                int conn, void* stream);
/* nscale::imfill<uchar>(image, seeds, binary=true, conn)  MorphologicOperations.h:65, .cpp:2913-2937 */
int nsc_imfill_u8(const uint8_t* img, const uint8_t* seeds, uint8_t* out, int w, int h, int conn, void* stream);
/* nscale::imfillHoles<uchar>(image, binary=true, conn)  MorphologicOperations.h:68, .cpp:2939-3006 */
int nsc_imfill_holes_u8(const uint8_t* img, uint8_t* out, int w, int h, int conn, void* stream);
/* nscale::bwselect<uchar>(binaryImage, seeds, conn)  MorphologicOperations.h:73, .cpp:3009-3031 */
int nsc_bwselect_u8(const uint8_t* bw, const uint8_t* seeds, uint8_t* out, int w, int h, int conn, void* stream);
/* nscale::imhmin<float>(image, h, conn)  MorphologicOperations.h:94, .cpp:3342-3364 */
int nsc_imhmin_f32(const float* img, float* out, int w, int h, float hval, int conn, void* stream);

/* ---- connected components --------------------------------------------------------------- */
/* ConnComponents::label(img, w, h, label, bgval, conn)  ConnComponents.h:38, .cpp:52-86:
   union of 4/8-neighbours with EQUAL nonzero values; label = smallest raster index of the
   component; background (0) pixels get bgval */
int nsc_cc_label_u8(const uint8_t* img, int32_t* labels, int w, int h, int32_t bgval, int conn, void* stream);
/* nscale::bwlabel2(binaryImage, conn, relab)  MorphologicOperations.h:77, .cpp:3092-3117 (bg -1) */
int nsc_bwlabel2_u8(const uint8_t* img, int32_t* labels, int w, int h, int conn, void* stream);
/* ConnComponents::relabel  ConnComponents.h:41, .cpp:89-118: roots numbered 1..N in raster order
   of their first pixel, bgval -> 0, in place; *count = N */
int nsc_cc_relabel(int32_t* labels, int w, int h, int32_t bgval, int* count, void* stream);
/* nscale::bwareaopen2(img, labeled=false, flatten=true, min, max, conn, count)
   MorphologicOperations.h:89, .cpp:3281-3314; ConnComponents::areaThresholdLabeled .cpp:205-239.
   Keeps components with min <= area < max; out = 255/0; *count = number kept (may be NULL). */
int nsc_bwareaopen2_u8(const uint8_t* img, uint8_t* out, int w, int h, int minSize, int maxSize, int conn,
                       int* count, void* stream);
/* Per-label bounding boxes of canonical labels 1..n (as nsc_cc_relabel produces): out5 is an
   n x 5 SoA block [label | minx | maxx | miny | maxy] (the layout of ConnComponents::boundingBox,
   ConnComponents.cpp:241-284, and of nscale::gpu::boundingBox2, cuda/ccl_uf.cu:993-1107), plus
   the per-label pixel count in area (may be NULL).  Device pointers. */
int nsc_bounding_boxes(const int32_t* canon_labels, int w, int h, int n, int32_t* out5, int32_t* area,
                       void* stream);

/* ConnComponents::boundingBox(w, h, label, bgval, compcount)  ConnComponents.h:43, .cpp:241-284,
   literally: labels are scanned in raster order, bgval skipped, the scan stops at the first -1 or 0
   label; *compcount = number of labels met, out5 (DEVICE, may be NULL) receives the reference's
   block of 5 x compcount ints [label | minx | maxx | miny | maxy] (column stride compcount, labels in
   first-seen order) when compcount <= cap.  Labels met before the stop must lie in 1..w*h-1. */
int nsc_cc_bounding_box(const int32_t* labels, int w, int h, int32_t bgval, int32_t* out5, int cap, int* compcount,
                        void* stream);

/* ---- morphology, distance, minima, watershed -------------------------------------------- */
/* cv::erode / cv::dilate with a k x k structuring element (HOST array, k odd, k <= 31), anchor at the
   centre, and `outside` as the value of pixels beyond the image (morphOpen pads with 255 for the
   erosion and 0 for the dilation, MorphologicOperations.cpp:3848-3882) */
int nsc_erode_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside,
                 void* stream);
int nsc_dilate_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside,
                  void* stream);
/* nscale::morphOpen<uchar>(image, kernel)  MorphologicOperations.h:114, .cpp:3848-3882 */
int nsc_morph_open_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, void* stream);
/* cv::distanceTransform(src, dst, CV_DIST_L2, CV_DIST_MASK_PRECISE) as called at
   HistologicalEntities.cpp:359: exact Euclidean distance of each nonzero pixel to the nearest zero */
int nsc_dist_transform_u8(const uint8_t* img, float* out, int w, int h, void* stream);
/* nscale::localMaxima<float> / localMinima<float>  MorphologicOperations.h:108-111, .cpp:3763-3845 */
int nsc_local_maxima_f32(const float* img, uint8_t* out, int w, int h, int conn, void* stream);
int nsc_local_minima_f32(const float* img, uint8_t* out, int w, int h, int conn, void* stream);
/* cv::watershed(image, markers) on a BGR image, markers in/out (as used by
   MorphologicOperations.cpp:3750) */
int nsc_cv_watershed(const uint8_t* bgr, int32_t* markers, int w, int h, void* stream);
/* process-wide counts of cv::watershed evaluations served by the decomposed parallel flood and by
   the sequential per-component fallback, and the reason bits of the last fallback (1 leaf capacity,
   2 comparison walk, 4 tie group too large / level-0 first background pixel, 8 heap capacity,
   16 tie order matters); no reference counterpart, reporting only */
void nsc_watershed_stats(int* fast, int* fallback, int* last_reason);
/* process-wide sums of the decomposed flood's phase-1 runs per tier of LDS frame size: frames of
   at most 512, 1024, 2048, 4096 and 6400 pixels, then the global-memory bucket; writes min(n, 6)
   counters and returns 6 (no reference counterpart, reporting only) */
int nsc_watershed_tier_runs(long long* runs, int n);
/* nscale::watershed2(origImage, image, conn)  MorphologicOperations.h:103, .cpp:3717-3760 */
int nsc_watershed2(const uint8_t* bgr, const float* img, int32_t* out, int w, int h, int conn, void* stream);

/* ---- the chain -------------------------------------------------------------------------- */
/* parameters of HistologicalEntities::segmentNuclei (HistologicalEntities.h:52-60) */
typedef struct nsc_seg_params {
    uint8_t blue, green, red;
    double T1, T2;
    uint8_t G1;
    int minSize, maxSize;
    uint8_t G2;
    int minSizePl, minSizeSeg, maxSizeSeg;
    int fillHolesConnectivity, reconConnectivity, watershedConnectivity;
} nsc_seg_params;
/* the reference's default arguments (HistologicalEntities.h:53) */
void nsc_seg_params_default(nsc_seg_params* p);

/* Reusable device workspace for tiles up to max_w x max_h (allocation happens here only). */
typedef struct nsc_workspace nsc_workspace;
int nsc_workspace_create(int max_w, int max_h, nsc_workspace** ws);
int nsc_workspace_destroy(nsc_workspace* ws);

/* HistologicalEntities::segmentNuclei(img, output, compcount, bbox, ...)
   HistologicalEntities.h:52-54, .cpp:473-598.
   bgr: device BGR tile; labels: device int32 w*h, the reference's output (label = smallest raster
   index of the nucleus, background -1).  If compcount is non-NULL it receives the number of
   nuclei.  stage (may be NULL) is an array of 32 device pointers; a non-NULL entry i receives the
   intermediate image with the reference's stage id i (IntermediateResultHandler ids,
   UtilsCVImageIO.h:32-51), 24 = filled final mask, 25 = labels.  stage_ms (may be NULL, host,
   32 doubles) accumulates per-stage GPU times in milliseconds under the same ids (the time up to
   the stage's dump point); id 26 is the imreconstruct<uchar>(rc_open, rc) call alone.  Returns
   the chain status code. */
int nsc_segment_nuclei(const uint8_t* bgr, int w, int h, const nsc_seg_params* p, int32_t* labels,
                       int* compcount, nsc_workspace* ws, void** stage, double* stage_ms, void* stream);

/* Bounding-box outputs of the chain (nsc_seg_io.bbox_mode). */
#define NSC_BBOX_NONE 0
/* ConnComponents::boundingBox(w, h, labels, bgval 0, compcount) exactly as segmentNuclei calls it
   (HistologicalEntities.cpp:590-591, ConnComponents.cpp:241-284): the scan skips label 0 and stops
   at the first background (-1) pixel in raster order, so on a tile whose first pixel is background
   compcount is 0 and no box is written.  Boxes come in first-seen order, label = the component's
   label (its smallest raster index), extents over the scanned pixels only. */
#define NSC_BBOX_REFERENCE_CPU 1
/* one box per nucleus in ascending label order, the result of nscale::gpu::boundingBox2
   (cuda/ccl_uf.cu:993-1107) on the returned label image: label column = the nucleus's label in
   the returned image (smallest raster index, or 1..N when canonical), extents over all its pixels;
   area (optional) receives the pixel counts.  compcount = number of nuclei. */
#define NSC_BBOX_ALL 2

/* Pitched images and box outputs of nsc_segment_nuclei_io (the reference's Mat/GpuMat arguments
   carry a row step; its GPU operators make continuous copies internally, MorphologicOperationsGPU.cpp:180-186). */
typedef struct nsc_seg_io {
    /* inputs */
    size_t bgr_step;      /* bytes between rows of bgr; 0 = 3*w */
    size_t labels_step;   /* bytes between rows of labels; 0 = 4*w */
    int canonical;        /* 0: the reference CPU labels (smallest raster index, background -1);
                             1: 1..N in raster order of first pixel, background 0 (the GPU path's relabel) */
    int bbox_mode;        /* NSC_BBOX_NONE / NSC_BBOX_REFERENCE_CPU / NSC_BBOX_ALL */
    int32_t* bbox5;       /* DEVICE, room for 5 * max_boxes ints: receives the reference's n x 5 block
                             [label | minx | maxx | miny | maxy], structure of arrays, column stride n = compcount */
    int32_t* area;        /* DEVICE, optional, max_boxes ints (NSC_BBOX_ALL): pixels per box */
    int max_boxes;        /* capacity; a larger compcount returns NSC_ERR_LIMIT with compcount set */
    /* outputs */
    int compcount;        /* the number of boxes the mode defines (the reference's compcount) */
    int nuclei;           /* number of nuclei (labelled components) */
} nsc_seg_io;
/* HistologicalEntities::segmentNuclei(img, output, compcount, bbox, ...) with pitched images and
   the box outputs selected by io (may be NULL: then as nsc_segment_nuclei without compcount). */
int nsc_segment_nuclei_io(const uint8_t* bgr, int w, int h, const nsc_seg_params* p, int32_t* labels, nsc_seg_io* io,
                          nsc_workspace* ws, void** stage, double* stage_ms, void* stream);

/* HistologicalEntities::plFindNucleusCandidates(img, seg_norbc, ...) (stage S1),
   HistologicalEntities.h:63-64, .cpp:115-304: seg_norbc (DEVICE u8, row step seg_step bytes; 0 =
   w) receives the candidate mask.  Returns NSC_SEG_CONTINUE when the chain goes on, or
   NSC_SEG_BACKGROUND / NSC_SEG_BACKGROUND_LIKELY / NSC_SEG_NO_CANDIDATES_LEFT (seg_norbc then
   unspecified), or a negative error. */
int nsc_pl_find_nucleus_candidates(const uint8_t* bgr, size_t bgr_step, int w, int h, const nsc_seg_params* p,
                                   uint8_t* seg_norbc, size_t seg_step, nsc_workspace* ws, void* stream);
/* HistologicalEntities::plSeparateNuclei(img, seg_open, seg_nonoverlap, minSizePl,
   watershedConnectivity) (stage A4), HistologicalEntities.h:65-66, .cpp:308-449: seg_open and
   seg_nonoverlap are DEVICE u8 planes with row steps (0 = w).  Returns NSC_SEG_CONTINUE or a
   negative error. */
int nsc_pl_separate_nuclei(const uint8_t* bgr, size_t bgr_step, const uint8_t* seg_open, size_t open_step, int w, int h,
                           int minSizePl, int watershedConnectivity, uint8_t* seg_nonoverlap, size_t out_step,
                           nsc_workspace* ws, void* stream);

/* ---- per-nucleus features -------------------------------------------------------------- */
/* Outputs of nsc_object_features: DEVICE arrays of n entries (intensity: n x 8, gradient: n x 6,
   row-major per nucleus); intensity / gradient may be NULL. */
typedef struct nsc_object_features_out {
    int32_t* area;          /* ObjFeatures::area           ObjFeatures.h:38, .cpp:15-45 */
    double* perimeter;      /* ObjFeatures::perimeter      ObjFeatures.h:41, .cpp:144-278 */
    double* major_axis;     /* ObjFeatures::ellipse        ObjFeatures.h:39, .cpp:48-116 */
    double* minor_axis;
    double* eccentricity;
    double* extent_ratio;   /* ObjFeatures::extent_ratio   ObjFeatures.h:40, .cpp:118-141 */
    double* circularity;    /* ObjFeatures::circularity    ObjFeatures.h:42, .cpp:281-293 */
    float* intensity;       /* ObjFeatures::intensityFeatures ObjFeatures.h:49, .cpp:342-375: mean, max,
                               min, std, entropy, energy, skewness, kurtosis of the plane's values */
    float* gradient;        /* ObjFeatures::gradientFeatures  ObjFeatures.h:50, .cpp:377-418: mean, std,
                               entropy, energy, skewness, kurtosis of Operators::gradient's magnitude */
} nsc_object_features_out;
/* Features of the n objects of a DEVICE label image described by a DEVICE box block (5 x n,
   column stride n: [label | minx | maxx | miny | maxy], as nsc_bounding_boxes or the chain's
   NSC_BBOX_ALL output give it); img (DEVICE, may be NULL without intensity / gradient outputs) is
   an interleaved image of `channels` bytes per pixel whose plane `channel` is measured (the
   reference splits BGR and passes one plane, ObjFeatures.cpp:540-560).  The canny features
   (ObjFeatures.h:51) are not provided. */
int nsc_object_features(const int32_t* labels, int w, int h, const int32_t* bbox5, int n, const uint8_t* img,
                        int channels, int channel, nsc_object_features_out* out, void* stream);

/* ---- tile files ------------------------------------------------------------------------ */
/* cv::imread(path) as the reference calls it (HistologicalEntities.cpp:456, IMREAD_COLOR) for PNG
   files: non-interlaced grey (1-16 bits), grey+alpha, RGB, RGBA (8/16 bits) or palette (1-8 bits)
   images become 8-bit BGR (low-depth grey expands to 8 bits, 16-bit samples keep their high byte,
   alpha is dropped).  *bgr receives a HOST buffer of
   3*w*h bytes allocated with malloc (release with nsc_free_host).  Other formats return an error. */
int nsc_png_read(const char* path, uint8_t** bgr, int* w, int* h);
/* cv::imwrite(path, img) for an 8-bit grey (channels 1) or BGR (channels 3) HOST image, row stride
   ld bytes (0 = dense): a PNG that decodes to the same pixels as OpenCV's. */
int nsc_png_write(const char* path, const uint8_t* img, int w, int h, int channels, size_t ld);
void nsc_free_host(void* p);
/* The 8-bit image cv::imwrite stores for an int32 label image (HistologicalEntities.cpp:466):
   saturate_cast<uchar> of every label (background -1 -> 0, labels above 255 -> 255).  DEVICE
   planes with row strides in bytes (0 = dense).  Asynchronous. */
int nsc_labels_to_u8(const int32_t* labels, size_t labels_ld, int w, int h, uint8_t* out, size_t out_ld,
                     void* stream);
/* HistologicalEntities::segmentNuclei(const std::string& in, const std::string& out, int& compcount,
   int*& bbox, ...) (HistologicalEntities.h:46-50, .cpp:453-470): reads the PNG tile `in`, runs the
   chain on the device with params (NULL = defaults) and, on NSC_SEG_SUCCESS, writes the label image
   to the PNG file `out` as imwrite stores it (nsc_labels_to_u8).  compcount and *bbox (HOST,
   malloc'd 5 x compcount block, NULL when compcount is 0; may be passed as NULL) are the
   reference's ConnComponents::boundingBox outputs.  Returns the chain status (NSC_SEG_INVALID_IMAGE
   when the file cannot be read) or a negative error.  Synchronous. */
int nsc_segment_nuclei_file(const char* in, const char* out, const nsc_seg_params* params, int* compcount,
                            int32_t** bbox, void* stream);

/* ---- whole-slide stitching -------------------------------------------------------------- */
/* Not in the reference, which segments the tiles of a slide independently (nu-segment.cpp:274-433):
   the step that removes the duplicates a tiling with overlap creates.  Tile (x0, y0, w, h) of a
   slide_w x slide_h slide cut on a regular grid of stride tile - overlap owns a core (the tile minus
   overlap/2 on every side with a neighbour; the cores partition the slide).  Each nucleus of the
   tile -- box i of the DEVICE bbox5 block (n x 5 [label | minx | maxx | miny | maxy], tile
   coordinates, canonical labels 1..n) -- is kept iff its box centre lies in the core; kept nuclei
   get ids first_id, first_id + 1, ... in label order, written to ids[i] (DEVICE, n ints; 0 =
   dropped).  owned (DEVICE w x h int32, row stride owned_ld bytes, may be NULL) receives the tile's
   labels (DEVICE, canonical, row stride labels_ld) mapped to those ids, 0 elsewhere.  *kept (host):
   number kept; *truncated (host, may be NULL): kept nuclei whose box touches a tile edge that has
   a neighbour (possibly cut: the overlap is smaller than twice their size).  Synchronous. */
int nsc_stitch_tile(const int32_t* labels, size_t labels_ld, const int32_t* bbox5, int n, int x0, int y0, int w, int h,
                    int slide_w, int slide_h, int overlap, int32_t first_id, int32_t* ids, int32_t* owned,
                    size_t owned_ld, int* kept, int* truncated, void* stream);

/* ---- colour normalization -------------------------------------------------------------- */
/* Normalization::targetParameters(originalI, targetMean, targetStd)
   (normalization/Normalization.h:44, .cpp:502-582): l-alpha-beta mean and standard deviation of
   the foreground pixels of a DEVICE BGR image (row stride ld >= 3 w bytes) that the fixed Fisher
   discriminant selects (segFG, .cpp:19-56) and whose L*a*b product is positive (.cpp:549-553).
   mean3 / std3 are host float[3]; count (host, may be NULL) receives the number of pixels used.
   Synchronizes the stream.  No foreground gives NaN statistics, as in the reference. */
int nsc_normalization_target_parameters(const uint8_t* bgr, int w, int h, size_t ld, float* mean3, float* std3,
                                        int64_t* count, void* stream);
/* Normalization::normalization(originalI, targetMean, targetStd) (Normalization.h:43,
   .cpp:436-499): out (DEVICE BGR, row stride out_ld) = the image with its foreground's LAB
   statistics moved to the target's (host float[3] each), background pixels converted to LAB and
   back.  Asynchronous on the stream; out may be the input itself (same stride), any other overlap
   is rejected.  log10f follows the fdlibm formula around a correctly rounded logf and powf(10, x)
   is correctly rounded, so results match a C library except where its own logf / powf is not
   correctly rounded (one level apart on a few channel values in a million). */
int nsc_normalization(const uint8_t* bgr, int w, int h, size_t ld, const float* target_mean3, const float* target_std3,
                      uint8_t* out, size_t out_ld, void* stream);

/* ---- continuous-batching tile server --------------------------------------------------- */
/* Replaces the reference's one-tile-per-MPI-process scale-out of segmentNuclei
   (nu-segment.cpp:274-433, nu-segmentTest.cpp:207) and its task split into stages S1 / A4
   (tasks/seg-tasks/S1Task.cpp, A4Task.cpp): requests of any tile size up to max_w x max_h pixels
   queue FIFO and are admitted when a slot (worker + stream + device memory) is free and the device
   memory the slot must add for them fits the HBM budget.  A slot's memory is a chain workspace for
   the largest tile it has served (grown on demand) plus a staging buffer for host inputs; the
   budget bounds the sum over slots, and an idle slot gives its memory back when a queued request
   could not otherwise be admitted.  Admitted requests run concurrently on their own streams;
   outputs go to the caller's device buffers, which the caller keeps alive until the request is
   waited on or polled complete. */
typedef struct nsc_server nsc_server;
typedef struct nsc_server_config {
    int slots;            /* concurrent requests, 1..64 */
    size_t hbm_budget;    /* device bytes the slots may hold together */
    int max_w, max_h;     /* largest tile: requests may have any shape of at most max_w*max_h pixels */
} nsc_server_config;
/* nsc_request.stage */
#define NSC_STAGE_CHAIN 0             /* HistologicalEntities::segmentNuclei */
#define NSC_STAGE_FIND_CANDIDATES 1   /* plFindNucleusCandidates (S1): bgr -> mask_out = seg_norbc */
#define NSC_STAGE_SEPARATE_NUCLEI 2   /* plSeparateNuclei (A4): bgr, mask_in = seg_open -> mask_out = seg_nonoverlap */
typedef struct nsc_request {
    const uint8_t* bgr;   /* BGR tile: device pointer, or host pointer when host_input != 0 */
    int w, h;
    int host_input;
    int32_t* labels;      /* chain: device int32 w*h, the chain's labels (smallest raster index, bg -1) */
    int canonical;        /* chain: != 0 relabels to 1..N in raster order of first pixel, bg 0 */
    int32_t* bbox5;       /* chain, optional: device n x 5 [label|minx|maxx|miny|maxy] (implies canonical) */
    int32_t* area;        /* chain, optional: device per-label pixel count (with bbox5) */
    int max_labels;       /* capacity of bbox5 / area in labels */
    const nsc_seg_params* params; /* NULL = the reference's defaults */
    int stage;            /* NSC_STAGE_*: the whole chain or one of its two task stages */
    uint8_t* mask_out;    /* stages S1 / A4: device u8 w*h output mask */
    const uint8_t* mask_in; /* stage A4: device u8 w*h seg_open */
} nsc_request;
typedef struct nsc_result {
    int status;           /* chain status code (stages: NSC_SEG_CONTINUE when the chain goes on) */
    int compcount;        /* nuclei found (chain) */
    int error;            /* NSC_OK or NSC_ERR_* */
    int slot;             /* the slot that ran it */
    double queue_ms, run_ms, total_ms;  /* submit -> start -> end, host clock */
    double start_ms, end_ms;            /* start and end, ms since the server was created */
} nsc_result;
/* device bytes a slot holds to serve a w x h tile from device memory (host inputs add 3*w*h) */
size_t nsc_request_bytes(int w, int h);
int nsc_server_create(const nsc_server_config* cfg, nsc_server** server);
int nsc_server_submit(nsc_server* server, const nsc_request* req, int64_t* ticket);
/* block until the request is done; returns its error code (NSC_OK) and fills *res */
int nsc_server_wait(nsc_server* server, int64_t ticket, nsc_result* res);
/* 1 while pending; otherwise like nsc_server_wait */
int nsc_server_poll(nsc_server* server, int64_t ticket, nsc_result* res);
/* queued and running requests, device bytes the slots hold, the budget */
int nsc_server_stats(nsc_server* server, int* queued, int* running, size_t* used_bytes, size_t* budget);
/* drains every submitted request, then frees the slots */
int nsc_server_destroy(nsc_server* server);

#ifdef __cplusplus
}
#endif
#endif /* NSCALE_HIP_H_ */
