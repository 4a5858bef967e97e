/*
 * CPU oracle for the nscale segmentNuclei morphology chain.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Each function restates the reference's algorithm (single-threaded, same visiting order
 * and queue discipline where the reference owns the algorithm) and cites the reference
 * file:line it follows.  Steps owned by OpenCV 2.4.9 (not vendored in the reference) are
 * restated from OpenCV's published algorithms: cv::erode/dilate (min/max over the
 * structuring element), cv::distanceTransform PRECISE (Felzenszwalb-Huttenlocher lower
 * envelope, float arithmetic), cv::floodFill (fixed range, mask only) and cv::watershed
 * (Meyer flooding with 256 FIFO levels).
 *
 * Paths below are relative to the reference's src/segment/.
 * Parity status of each function is listed in tests/test_oracle.py and DESIGN.md.
 */
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <queue>
#include <unordered_map>
#include <vector>

namespace {

// padded copy with a constant border of width b
template <typename T>
std::vector<T> pad(const T* src, int w, int h, int b, T val) {
    int W = w + 2 * b, H = h + 2 * b;
    std::vector<T> out((size_t)W * H, val);
    for (int y = 0; y < h; ++y)
        std::memcpy(&out[(size_t)(y + b) * W + b], src + (size_t)y * w, sizeof(T) * w);
    return out;
}

template <typename T>
void crop(const std::vector<T>& src, int W, int b, T* dst, int w, int h) {
    for (int y = 0; y < h; ++y)
        std::memcpy(dst + (size_t)y * w, &src[(size_t)(y + b) * W + b], sizeof(T) * w);
}

// ---------------------------------------------------------------------------------------
// imreconstruct<T>  (MorphologicOperations.cpp:1730-1903, propagate<T> :430-441)
// ---------------------------------------------------------------------------------------
template <typename T>
inline void propagate(std::vector<T>& O, const std::vector<T>& I, std::queue<int>& xQ,
                      std::queue<int>& yQ, int x, int y, int W, const T& pval) {
    size_t i = (size_t)y * W + x;
    T qval = O[i];
    T ival = I[i];
    if ((qval < pval) && (ival != qval)) {
        O[i] = std::min(pval, ival);
        xQ.push(x);
        yQ.push(y);
    }
}

template <typename T>
void imreconstruct(const T* seeds, const T* image, T* out, int w, int h, int conn) {
    // copyMakeBorder(..., 1, 1, 1, 1, BORDER_CONSTANT, 0)  (:1736-1739)
    std::vector<T> O = pad<T>(seeds, w, h, 1, T(0));
    std::vector<T> I = pad<T>(image, w, h, 1, T(0));
    const int W = w + 2, H = h + 2;
    const int maxx = W - 1, maxy = H - 1;
    std::queue<int> xQ, yQ;
    T pval, preval;

    // raster scan (:1757-1778)
    for (int y = 1; y < maxy; ++y) {
        T* oPtr = &O[(size_t)y * W];
        T* oPtrMinus = &O[(size_t)(y - 1) * W];
        const T* iPtr = &I[(size_t)y * W];
        preval = oPtr[0];
        for (int x = 1; x < maxx; ++x) {
            pval = oPtr[x];
            pval = std::max(pval, std::max(preval, oPtrMinus[x]));
            if (conn == 8) pval = std::max(pval, std::max(oPtrMinus[x + 1], oPtrMinus[x - 1]));
            preval = std::min(pval, iPtr[x]);
            oPtr[x] = preval;
        }
    }
    // anti-raster scan with queue seeding (:1782-1828)
    for (int y = maxy - 1; y > 0; --y) {
        T* oPtr = &O[(size_t)y * W];
        T* oPtrPlus = &O[(size_t)(y + 1) * W];
        const T* iPtr = &I[(size_t)y * W];
        const T* iPtrPlus = &I[(size_t)(y + 1) * W];
        preval = oPtr[maxx];
        for (int x = maxx - 1; x > 0; --x) {
            pval = oPtr[x];
            pval = std::max(pval, std::max(preval, oPtrPlus[x]));
            if (conn == 8) pval = std::max(pval, std::max(oPtrPlus[x + 1], oPtrPlus[x - 1]));
            preval = std::min(pval, iPtr[x]);
            oPtr[x] = preval;
            pval = oPtr[x];
            if ((oPtr[x + 1] < std::min(pval, iPtr[x + 1])) || (oPtrPlus[x] < std::min(pval, iPtrPlus[x]))) {
                xQ.push(x); yQ.push(y);
                continue;
            }
            if (conn == 8) {
                if ((oPtrPlus[x + 1] < std::min(pval, iPtrPlus[x + 1])) ||
                    (oPtrPlus[x - 1] < std::min(pval, iPtrPlus[x - 1]))) {
                    xQ.push(x); yQ.push(y);
                    continue;
                }
            }
        }
    }
    // FIFO propagation (:1837-1893)
    while (!xQ.empty()) {
        int x = xQ.front(), y = yQ.front();
        xQ.pop(); yQ.pop();
        T pv = O[(size_t)y * W + x];
        if (y > 0) propagate<T>(O, I, xQ, yQ, x, y - 1, W, pv);
        if (y < maxy) propagate<T>(O, I, xQ, yQ, x, y + 1, W, pv);
        if (x > 0) propagate<T>(O, I, xQ, yQ, x - 1, y, W, pv);
        if (x < maxx) propagate<T>(O, I, xQ, yQ, x + 1, y, W, pv);
        if (conn == 8) {
            if (y > 0) {
                if (x > 0) propagate<T>(O, I, xQ, yQ, x - 1, y - 1, W, pv);
                if (x < maxx) propagate<T>(O, I, xQ, yQ, x + 1, y - 1, W, pv);
            }
            if (y < maxy) {
                if (x > 0) propagate<T>(O, I, xQ, yQ, x - 1, y + 1, W, pv);
                if (x < maxx) propagate<T>(O, I, xQ, yQ, x + 1, y + 1, W, pv);
            }
        }
    }
    crop<T>(O, W, 1, out, w, h);  // (:1901)
}

// ---------------------------------------------------------------------------------------
// imreconstructBinary<T>  (MorphologicOperations.cpp:2768-2909, propagateBinary :2746-2753)
// ---------------------------------------------------------------------------------------
template <typename T>
void imreconstruct_binary(const T* seeds, const T* image, T* out, int w, int h, int conn) {
    std::vector<T> O = pad<T>(seeds, w, h, 1, T(0));
    std::vector<T> I = pad<T>(image, w, h, 1, T(0));
    const int W = w + 2, H = h + 2;
    const int maxx = W - 1, maxy = H - 1;
    std::queue<int> xQ, yQ;
    // contour pixels (:2794-2833)
    for (int y = 1; y < maxy; ++y) {
        const T* oPtr = &O[(size_t)y * W];
        const T* oPtrPlus = &O[(size_t)(y + 1) * W];
        const T* oPtrMinus = &O[(size_t)(y - 1) * W];
        const T* iPtr = &I[(size_t)y * W];
        for (int x = 1; x < maxx; ++x) {
            T pval = oPtr[x], ival = iPtr[x];
            if (pval != 0 && ival != 0) {
                if ((oPtrMinus[x] == 0) || (oPtrPlus[x] == 0) || (oPtr[x + 1] == 0) || (oPtr[x - 1] == 0)) {
                    xQ.push(x); yQ.push(y);
                    continue;
                }
                if (conn == 8) {
                    if ((oPtrMinus[x - 1] == 0) || (oPtrMinus[x + 1] == 0) || (oPtrPlus[x - 1] == 0) ||
                        (oPtrPlus[x + 1] == 0)) {
                        xQ.push(x); yQ.push(y);
                        continue;
                    }
                }
            }
        }
    }
    const T outval = std::numeric_limits<T>::max();
    auto prop = [&](int x, int y) {
        size_t i = (size_t)y * W + x;
        if ((O[i] == 0) && (I[i] != 0)) {
            O[i] = outval;
            xQ.push(x); yQ.push(y);
        }
    };
    // (:2842-2904)
    while (!xQ.empty()) {
        int x = xQ.front(), y = yQ.front();
        xQ.pop(); yQ.pop();
        if (y > 0) prop(x, y - 1);
        if (y < maxy) prop(x, y + 1);
        if (x > 0) prop(x - 1, y);
        if (x < maxx) prop(x + 1, y);
        if (conn == 8) {
            if (y > 0) {
                if (x > 0) prop(x - 1, y - 1);
                if (x < maxx) prop(x + 1, y - 1);
            }
            if (y < maxy) {
                if (x > 0) prop(x - 1, y + 1);
                if (x < maxx) prop(x + 1, y + 1);
            }
        }
    }
    crop<T>(O, W, 1, out, w, h);
}

// PixelOperations::invert<uchar> (PixelOperations.cpp:21-43): max - x
inline void invert_u8(std::vector<uint8_t>& v) {
    for (auto& x : v) x = (uint8_t)(255 - x);
}

// ---------------------------------------------------------------------------------------
// ConnComponents (ConnComponents.cpp).  find/merge/flatten are recursive in the reference;
// restated iteratively with identical results (root = smallest index of the set).
// ---------------------------------------------------------------------------------------
inline int cc_find(int32_t* label, int x) {  // :24-32
    int r = x;
    while (label[r] != r) r = label[r];
    while (label[x] != r) { int n = label[x]; label[x] = r; x = n; }
    return r;
}
inline void cc_merge(int32_t* label, int x, int y) {  // :34-43
    x = cc_find(label, x);
    y = cc_find(label, y);
    if (label[x] < label[y]) label[y] = x;
    else if (label[x] > label[y]) label[x] = y;
}
inline int cc_flatten(int32_t* label, int x, int32_t bgval) {  // :137-147
    // follow the chain until a root or a bgval entry, then compress
    int r = x;
    while (label[r] != bgval && label[r] != r) r = label[r];
    int v = (label[r] == bgval) ? bgval : r;
    while (x != r && label[x] != bgval && label[x] != x) { int n = label[x]; label[x] = v; x = n; }
    return v;
}

}  // namespace

extern "C" {

void orc_imreconstruct_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h, int conn) {
    imreconstruct<uint8_t>(marker, mask, out, w, h, conn);
}
void orc_imreconstruct_f32(const float* marker, const float* mask, float* out, int w, int h, int conn) {
    imreconstruct<float>(marker, mask, out, w, h, conn);
}
void orc_imreconstruct_i32(const int32_t* marker, const int32_t* mask, int32_t* out, int w, int h, int conn) {
    imreconstruct<int32_t>(marker, mask, out, w, h, conn);
}
void orc_imreconstruct_binary_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h, int conn) {
    imreconstruct_binary<uint8_t>(marker, mask, out, w, h, conn);
}

// imfill<T>(image, seeds, binary=true, conn)  (MorphologicOperations.cpp:2913-2937)
void orc_imfill_u8(const uint8_t* img, const uint8_t* seeds, uint8_t* out, int w, int h, int conn) {
    size_t n = (size_t)w * h;
    std::vector<uint8_t> mask(img, img + n);
    invert_u8(mask);
    std::vector<uint8_t> marker(n, 0), rec(n);
    for (size_t i = 0; i < n; ++i) if (seeds[i]) marker[i] = mask[i];
    imreconstruct_binary<uint8_t>(marker.data(), mask.data(), rec.data(), w, h, conn);
    for (size_t i = 0; i < n; ++i) out[i] = img[i] | rec[i];
}

// imfillHoles<T>(image, binary, conn)  (MorphologicOperations.cpp:2939-3006); T=uchar: mn=0, mx=255
static void imfill_holes_u8(const uint8_t* img, uint8_t* out, int w, int h, int conn, bool binary) {
    const int W = w + 2, H = h + 2;
    std::vector<uint8_t> mask = pad<uint8_t>(img, w, h, 1, 0);
    std::vector<uint8_t> marker((size_t)W * H, 255);
    for (int y = 1; y <= h; ++y) std::memset(&marker[(size_t)y * W + 1], 0, w);
    invert_u8(mask);
    std::vector<uint8_t> rec((size_t)W * H);
    if (binary) imreconstruct_binary<uint8_t>(marker.data(), mask.data(), rec.data(), W, H, conn);
    else imreconstruct<uint8_t>(marker.data(), mask.data(), rec.data(), W, H, conn);
    invert_u8(rec);
    crop<uint8_t>(rec, W, 1, out, w, h);
}
void orc_imfill_holes_u8(const uint8_t* img, uint8_t* out, int w, int h, int conn) {
    imfill_holes_u8(img, out, w, h, conn, true);
}
void orc_imfill_holes_gray_u8(const uint8_t* img, uint8_t* out, int w, int h, int conn) {
    imfill_holes_u8(img, out, w, h, conn, false);
}

// bwselect<T>  (MorphologicOperations.cpp:3009-3031)
void orc_bwselect_u8(const uint8_t* bw, const uint8_t* seeds, uint8_t* out, int w, int h, int conn) {
    size_t n = (size_t)w * h;
    std::vector<uint8_t> marker(n, 0), rec(n);
    for (size_t i = 0; i < n; ++i) if (seeds[i]) marker[i] = bw[i];
    imreconstruct_binary<uint8_t>(marker.data(), bw, rec.data(), w, h, conn);
    for (size_t i = 0; i < n; ++i) out[i] = rec[i] & bw[i];
}

// ConnComponents::label  (ConnComponents.cpp:52-86)
void orc_cc_label(const uint8_t* img, int w, int h, int32_t* label, int32_t bgval, int conn) {
    int length = w * h;
    for (int i = 0; i < length; ++i) label[i] = i;
    int i = -1;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            ++i;
            uint8_t p = img[i];
            if (p == 0) {
                label[i] = bgval;
            } else {
                int imw = i - w;
                if (x > 0 && p == img[i - 1]) cc_merge(label, i, i - 1);
                if (y > 0 && p == img[imw]) cc_merge(label, i, imw);
                if (conn == 8) {
                    if (y > 0 && x > 0 && p == img[imw - 1]) cc_merge(label, i, imw - 1);
                    if (y > 0 && x < w - 1 && p == img[imw + 1]) cc_merge(label, i, imw + 1);
                }
            }
        }
    }
    for (int k = 0; k < length; ++k) label[k] = cc_flatten(label, k, bgval);
}

// ConnComponents::relabel  (ConnComponents.cpp:89-118)
int orc_cc_relabel(int w, int h, int32_t* label, int32_t bgval) {
    int length = w * h;
    std::unordered_map<int, int> labelmap;
    int j = 1;
    labelmap[bgval] = 0;
    for (int i = 0; i < length; ++i) if (label[i] == i) labelmap[i] = j++;
    for (int i = 0; i < length; ++i) label[i] = labelmap[label[i]];
    return j - 1;
}

// ConnComponents::areaThresholdLabeled  (ConnComponents.cpp:205-239)
int orc_cc_area_threshold_labeled(const int32_t* label, int w, int h, int32_t* n_label, int32_t bgval,
                                  int lower, int upper) {
    int length = w * h;
    std::unordered_map<int, int> areas;
    for (int i = 0; i < length; ++i) {
        if (label[i] != bgval) areas[label[i]] += 1;
        n_label[i] = bgval;
    }
    std::unordered_map<int, int> keep;
    int j = 0;
    for (auto& kv : areas) {
        if (kv.second >= lower && kv.second < upper) { keep[kv.first] = 1; ++j; }
        else keep[kv.first] = 0;
    }
    for (int i = 0; i < length; ++i)
        if (label[i] != bgval && keep[label[i]] == 1) n_label[i] = label[i];
    return j;
}

// ConnComponents::boundingBox  (ConnComponents.cpp:241-284), literal: it stops at the first
// label equal to -1 or 0 that is not bgval (:252-253).
int orc_cc_bounding_box_literal(int w, int h, const int32_t* label, int32_t bgval, int32_t* out5, int cap) {
    int length = w * h;
    std::unordered_map<int, int> l2i;
    std::vector<int> tb, mnx, mxx, mny, mxy;
    for (int i = 0; i < length; ++i) {
        if (label[i] == bgval) continue;
        if (label[i] == -1) break;
        if (label[i] == 0) break;
        int x = i % w, y = i / w;
        auto it = l2i.find(label[i]);
        if (it == l2i.end()) {
            l2i[label[i]] = (int)tb.size();
            tb.push_back(label[i]); mnx.push_back(x); mxx.push_back(x); mny.push_back(y); mxy.push_back(y);
        } else {
            int k = it->second;
            mnx[k] = std::min(mnx[k], x); mxx[k] = std::max(mxx[k], x);
            mny[k] = std::min(mny[k], y); mxy[k] = std::max(mxy[k], y);
        }
    }
    int n = (int)tb.size();
    if (out5) {
        int m = std::min(n, cap);
        for (int k = 0; k < m; ++k) {
            out5[k] = tb[k]; out5[cap + k] = mnx[k]; out5[2 * cap + k] = mxx[k];
            out5[3 * cap + k] = mny[k]; out5[4 * cap + k] = mxy[k];
        }
    }
    return n;
}

// bwareaopen2(image, labeled=false, flatten=true, min, max, conn, count)
// (MorphologicOperations.cpp:3281-3314)
int orc_bwareaopen2_u8(const uint8_t* img, uint8_t* out, int w, int h, int minSize, int maxSize, int conn) {
    size_t n = (size_t)w * h;
    std::vector<int32_t> temp(n), o(n);
    orc_cc_label(img, w, h, temp.data(), -1, conn);
    int count = orc_cc_area_threshold_labeled(temp.data(), w, h, o.data(), -1, minSize, maxSize);
    for (size_t i = 0; i < n; ++i) out[i] = (o[i] > -1) ? 255 : 0;
    return count;
}

// bwlabel2(img, conn, relab)  (MorphologicOperations.cpp:3092-3117; relabel is commented out)
void orc_bwlabel2(const uint8_t* img, int32_t* labels, int w, int h, int conn) {
    orc_cc_label(img, w, h, labels, -1, conn);
}

// cv::erode / cv::dilate: dst(x,y) = min/max over kernel(i,j) != 0 of src(x+j-a, y+i-a), anchor a
// at the centre; the reference pads explicitly (morphOpen) so the outside value is given here.
static void morph_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k,
                     int outside, bool is_erode) {
    // Same result as a per-tap min/max over the kernel's nonzero offsets; evaluated per kernel
    // row over runs of ones, row-at-a-time, so the loops vectorise.
    int a = k / 2;
    std::vector<uint8_t> P = pad<uint8_t>(img, w, h, a, (uint8_t)outside);
    int W = w + 2 * a;
    std::vector<uint8_t> acc(w);
    for (int y = 0; y < h; ++y) {
        std::memset(acc.data(), is_erode ? 255 : 0, w);
        for (int i = 0; i < k; ++i) {
            const uint8_t* src = &P[(size_t)(y + i) * W];
            for (int j = 0; j < k; ++j) {
                if (!kernel[i * k + j]) continue;
                const uint8_t* s = src + j;
                if (is_erode) for (int x = 0; x < w; ++x) acc[x] = std::min(acc[x], s[x]);
                else for (int x = 0; x < w; ++x) acc[x] = std::max(acc[x], s[x]);
            }
        }
        std::memcpy(out + (size_t)y * w, acc.data(), w);
    }
}
void orc_erode_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside) {
    morph_u8(img, out, w, h, kernel, k, outside, true);
}
void orc_dilate_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k, int outside) {
    morph_u8(img, out, w, h, kernel, k, outside, false);
}
// morphOpen<T>  (MorphologicOperations.cpp:3848-3882): erode with border 255, then dilate with
// border 0 (the padded cv::erode/cv::dilate never read past the explicit padding).
void orc_morph_open_u8(const uint8_t* img, uint8_t* out, int w, int h, const uint8_t* kernel, int k) {
    std::vector<uint8_t> e((size_t)w * h);
    morph_u8(img, e.data(), w, h, kernel, k, 255, true);
    morph_u8(e.data(), out, w, h, kernel, k, 0, false);
}

// cv::distanceTransform(src, dst, CV_DIST_L2, CV_DIST_MASK_PRECISE) (OpenCV 2.4.9
// imgproc/distransform.cpp, trueDistTrans): column pass = exact 1-D distance to the nearest
// zero (squared, 1e15 when the column has none); row pass = lower envelope of parabolas in
// float arithmetic, d = sqrt(float((q-p)^2) + f[p]).
void orc_dist_transform_u8(const uint8_t* src, float* dst, int n /*cols*/, int m /*rows*/) {
    const float inf = 1e15f;
    std::vector<float> sqr_tab((size_t)std::max(2 * m, 2 * n) + 2);
    for (int i = 0; i < m; ++i) sqr_tab[i] = (float)((double)i * i);
    for (int i = m; i < 2 * m; ++i) sqr_tab[i] = inf;
    // column pass evaluated for all columns at once, bottom-up then top-down (same values as the
    // per-column loops of DTColumnInvoker)
    std::vector<int> d((size_t)m * n);
    std::vector<int> dist(n, m - 1);
    for (int j = m - 1; j >= 0; --j)
        for (int c = 0; c < n; ++c) {
            dist[c] = (src[(size_t)j * n + c] == 0) ? 0 : dist[c] + 1;
            d[(size_t)j * n + c] = dist[c];
        }
    std::fill(dist.begin(), dist.end(), m - 1);
    for (int j = 0; j < m; ++j)
        for (int c = 0; c < n; ++c) {
            dist[c] = std::min(dist[c] + 1, d[(size_t)j * n + c]);
            dst[(size_t)j * n + c] = sqr_tab[std::min(dist[c], 2 * m - 1)];
        }
    std::vector<float> sq(n), inv(n);
    sq[0] = inv[0] = 0.f;
    for (int i = 1; i < n; ++i) { inv[i] = (float)(0.5 / i); sq[i] = (float)((double)i * i); }
    std::vector<float> f(n), z(n + 1);
    std::vector<int> v(n);
    for (int r = 0; r < m; ++r) {
        float* dr = dst + (size_t)r * n;
        int p, q, k;
        v[0] = 0; z[0] = -inf; z[1] = inf; f[0] = dr[0];
        for (q = 1, k = 0; q < n; q++) {
            float fq = dr[q];
            f[q] = fq;
            for (;; k--) {
                p = v[k];
                volatile float num = fq + sq[q];
                num = num - dr[p];
                num = num - sq[p];
                float s = num * inv[q - p];
                if (s > z[k]) {
                    k++;
                    v[k] = q;
                    z[k] = s;
                    z[k + 1] = inf;
                    break;
                }
            }
        }
        for (q = 0, k = 0; q < n; q++) {
            while (z[k + 1] < q) k++;
            p = v[k];
            dr[q] = std::sqrt(sq[std::abs(q - p)] + f[p]);
        }
    }
}

// imhmin<T>  (MorphologicOperations.cpp:3342-3364), T = float: invert = negate
void orc_imhmin_f32(const float* img, float* out, int w, int h, float hval, int conn) {
    size_t n = (size_t)w * h;
    std::vector<float> mask(n), marker(n), rec(n);
    for (size_t i = 0; i < n; ++i) mask[i] = -img[i];
    for (size_t i = 0; i < n; ++i) marker[i] = mask[i] - hval;
    imreconstruct<float>(marker.data(), mask.data(), rec.data(), w, h, conn);
    for (size_t i = 0; i < n; ++i) out[i] = -rec[i];
}

// localMaxima<float>  (MorphologicOperations.cpp:3763-3836).  cv::floodFill(inputBlock, output,
// seed, ..., lo=up=0, FLOODFILL_FIXED_RANGE | FLOODFILL_MASK_ONLY | conn) marks in the mask
// (output, padded by one) every pixel conn-connected to the seed through pixels whose value
// equals the seed value and whose mask entry is still 0.
void orc_local_maxima_f32(const float* img, uint8_t* out, int w, int h, int conn) {
    size_t n = (size_t)w * h;
    std::vector<float> marker(n), rec(n);
    for (size_t i = 0; i < n; ++i) marker[i] = img[i] - 1.0f;
    imreconstruct<float>(marker.data(), img, rec.data(), w, h, conn);
    std::vector<uint8_t> cand(n);
    for (size_t i = 0; i < n; ++i) cand[i] = (marker[i] < rec[i]) ? 255 : 0;
    const float mn = -std::numeric_limits<float>::max();
    std::vector<uint8_t> O = pad<uint8_t>(cand.data(), w, h, 1, 255);
    std::vector<float> I = pad<float>(img, w, h, 1, mn);
    const int W = w + 2, maxy = h + 1, maxx = w + 1;
    std::vector<int> stack;
    auto flood = [&](int sx, int sy) {  // padded coordinates
        float sv = I[(size_t)sy * W + sx];
        stack.clear();
        stack.push_back(sy * W + sx);
        O[(size_t)sy * W + sx] = 1;
        while (!stack.empty()) {
            int i = stack.back();
            stack.pop_back();
            int y = i / W, x = i % W;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    if (dx == 0 && dy == 0) continue;
                    if (conn == 4 && dx != 0 && dy != 0) continue;
                    int yy = y + dy, xx = x + dx;
                    if (yy < 1 || yy > h || xx < 1 || xx > w) continue;
                    size_t j = (size_t)yy * W + xx;
                    if (O[j] == 0 && I[j] >= sv && I[j] <= sv) { O[j] = 1; stack.push_back((int)j); }
                }
        }
    };
    for (int y = 1; y < maxy; ++y) {
        for (int x = 1; x < maxx; ++x) {
            size_t i = (size_t)y * W + x;
            if (O[i] > 0) continue;
            float val = I[i];
            if ((val < I[i - W]) || (val < I[i + W]) || (val < I[i - 1]) || (val < I[i + 1])) {
                flood(x, y);
                continue;
            }
            if (conn == 8) {
                if ((val < I[i - W - 1]) || (val < I[i - W + 1]) || (val < I[i + W - 1]) || (val < I[i + W + 1])) {
                    flood(x, y);
                    continue;
                }
            }
        }
    }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) out[(size_t)y * w + x] = (O[(size_t)(y + 1) * W + x + 1] == 0) ? 255 : 0;
}

// localMinima<float>  (MorphologicOperations.cpp:3838-3845)
void orc_local_minima_f32(const float* img, uint8_t* out, int w, int h, int conn) {
    size_t n = (size_t)w * h;
    std::vector<float> c(n);
    for (size_t i = 0; i < n; ++i) c[i] = -img[i];
    orc_local_maxima_f32(c.data(), out, w, h, conn);
}

// cv::watershed (OpenCV 2.4.9 imgproc/segmentation.cpp, cvWatershed): 1-px frame of -1; interior
// markers < 0 reset to 0; unknown pixels next to a positive marker queued at the smallest
// max-channel |difference| to such a neighbour; then pop from the lowest non-empty of 256 FIFO
// levels; label = the common positive label of the 4 neighbours, -1 if they disagree; a labelled
// pixel queues its unknown neighbours at their max-channel difference to it.
void orc_cv_watershed(const uint8_t* img, int32_t* mask, int w, int h) {
    const int IN_QUEUE = -2, WSHED = -1, NQ = 256;
    const int mstep = w, istep = 3 * w;
    std::vector<std::deque<std::pair<int, int>>> q(NQ);
    auto c_diff = [&](const uint8_t* p1, const uint8_t* p2) {
        int db = std::abs((int)p1[0] - (int)p2[0]);
        int dg = std::abs((int)p1[1] - (int)p2[1]);
        int dr = std::abs((int)p1[2] - (int)p2[2]);
        return std::max(std::max(db, dg), dr);
    };
    for (int j = 0; j < w; ++j) mask[j] = mask[j + mstep * (h - 1)] = WSHED;
    for (int i = 1; i < h - 1; ++i) {
        int32_t* mrow = mask + (size_t)i * mstep;
        const uint8_t* irow = img + (size_t)i * istep;
        mrow[0] = mrow[w - 1] = WSHED;
        for (int j = 1; j < w - 1; ++j) {
            int32_t* m = mrow + j;
            if (m[0] < 0) m[0] = 0;
            if (m[0] == 0 && (m[-1] > 0 || m[1] > 0 || m[-mstep] > 0 || m[mstep] > 0)) {
                const uint8_t* ptr = irow + j * 3;
                int idx = 256, t;
                if (m[-1] > 0) idx = c_diff(ptr, ptr - 3);
                if (m[1] > 0) { t = c_diff(ptr, ptr + 3); idx = std::min(idx, t); }
                if (m[-mstep] > 0) { t = c_diff(ptr, ptr - istep); idx = std::min(idx, t); }
                if (m[mstep] > 0) { t = c_diff(ptr, ptr + istep); idx = std::min(idx, t); }
                q[idx].push_back({i * mstep + j, i * istep + j * 3});
                m[0] = IN_QUEUE;
            }
        }
    }
    int active = 0;
    for (active = 0; active < NQ; ++active) if (!q[active].empty()) break;
    if (active == NQ) return;
    for (;;) {
        if (q[active].empty()) {
            int i;
            for (i = active + 1; i < NQ; ++i) if (!q[i].empty()) break;
            if (i == NQ) break;
            active = i;
        }
        auto node = q[active].front();
        q[active].pop_front();
        int mofs = node.first, iofs = node.second;
        int32_t* m = mask + mofs;
        const uint8_t* ptr = img + iofs;
        int lab = 0, t;
        t = m[-1];
        if (t > 0) lab = t;
        t = m[1];
        if (t > 0) { if (lab == 0) lab = t; else if (t != lab) lab = WSHED; }
        t = m[-mstep];
        if (t > 0) { if (lab == 0) lab = t; else if (t != lab) lab = WSHED; }
        t = m[mstep];
        if (t > 0) { if (lab == 0) lab = t; else if (t != lab) lab = WSHED; }
        m[0] = lab;
        if (lab == WSHED) continue;
        if (m[-1] == 0) {
            t = c_diff(ptr, ptr - 3);
            q[t].push_back({mofs - 1, iofs - 3}); active = std::min(active, t); m[-1] = IN_QUEUE;
        }
        if (m[1] == 0) {
            t = c_diff(ptr, ptr + 3);
            q[t].push_back({mofs + 1, iofs + 3}); active = std::min(active, t); m[1] = IN_QUEUE;
        }
        if (m[-mstep] == 0) {
            t = c_diff(ptr, ptr - istep);
            q[t].push_back({mofs - mstep, iofs - istep}); active = std::min(active, t); m[-mstep] = IN_QUEUE;
        }
        if (m[mstep] == 0) {
            t = c_diff(ptr, ptr + istep);
            q[t].push_back({mofs + mstep, iofs + istep}); active = std::min(active, t); m[mstep] = IN_QUEUE;
        }
    }
}

// NeighborOperations::border<T>  (NeighborOperations.cpp:17-44)
void orc_border_i32(const int32_t* img, int32_t* res, int w, int h, int32_t background) {
    for (int y = 1; y < h; ++y) {
        const int32_t* ptr = img + (size_t)y * w;
        const int32_t* ptrm1 = img + (size_t)(y - 1) * w;
        int32_t* r = res + (size_t)y * w;
        for (int x = 1; x < w - 1; ++x) {
            if (ptrm1[x] == background && ptr[x] != background &&
                ((ptrm1[x - 1] != background && ptr[x - 1] == background) ||
                 (ptrm1[x + 1] != background && ptr[x + 1] == background)))
                r[x] = background;
            else
                r[x] = ptr[x];
        }
    }
}

// watershed2  (MorphologicOperations.cpp:3717-3760)
void orc_watershed2(const uint8_t* bgr, const float* image, int32_t* out, int w, int h, int conn) {
    size_t n = (size_t)w * h;
    std::vector<uint8_t> minima(n);
    orc_local_minima_f32(image, minima.data(), w, h, conn);
    std::vector<int32_t> labels(n);
    orc_bwlabel2(minima.data(), labels.data(), w, h, conn);
    const int W = w + 2, H = h + 2;
    std::vector<int32_t> temp = pad<int32_t>(labels.data(), w, h, 1, 0);
    std::vector<uint8_t> input((size_t)W * H * 3, 0);
    for (int y = 0; y < h; ++y) std::memcpy(&input[((size_t)(y + 1) * W + 1) * 3], bgr + (size_t)y * w * 3, (size_t)w * 3);
    orc_cv_watershed(input.data(), temp.data(), W, H);
    std::vector<int32_t> res((size_t)W * H, 0);
    orc_border_i32(temp.data(), res.data(), W, H, -1);
    crop<int32_t>(res, W, 1, out, w, h);
}

// HistologicalEntities::getBackground  (HistologicalEntities.cpp:93-109)
void orc_get_background(const uint8_t* bgr, uint8_t* out, int w, int h, uint8_t blue, uint8_t green, uint8_t red) {
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; ++i)
        out[i] = (bgr[3 * i] > blue && bgr[3 * i + 1] > green && bgr[3 * i + 2] > red) ? 255 : 0;
}

// HistologicalEntities::getRBC  (HistologicalEntities.cpp:33-82): float planes
// bd = B*1 + FLT_EPSILON, gd = G*1 + FLT_EPSILON, rd = R; imR2G = rd/gd; imR2B = (rd/bd) > 1.
void orc_get_rbc(const uint8_t* bgr, uint8_t* out, int w, int h, double T1, double T2) {
    size_t n = (size_t)w * h;
    std::vector<uint8_t> bw1(n), bw2(n), r2b(n), sel(n);
    const float t1 = (float)T1, t2 = (float)T2;
    size_t nz = 0;
    for (size_t i = 0; i < n; ++i) {
        volatile float bd = (float)bgr[3 * i] * 1.0f + FLT_EPSILON;
        volatile float gd = (float)bgr[3 * i + 1] * 1.0f + FLT_EPSILON;
        float rd = (float)bgr[3 * i + 2];
        volatile float r2g = rd / gd;
        volatile float r2bv = rd / bd;
        r2b[i] = (r2bv > 1.0f) ? 255 : 0;
        bw1[i] = (r2g > t1) ? 255 : 0;
        bw2[i] = (r2g > t2) ? 255 : 0;
        nz += bw1[i] != 0;
    }
    if (nz > 0) {
        orc_bwselect_u8(bw2.data(), bw1.data(), sel.data(), w, h, 8);
        for (size_t i = 0; i < n; ++i) out[i] = sel[i] & r2b[i];
    } else {
        std::memset(out, 0, n);
    }
}

void orc_seg_params_default(orc_seg_params* p) {
    // HistologicalEntities.h:53
    p->blue = 220; p->green = 220; p->red = 220;
    p->T1 = 5.0; p->T2 = 4.0;
    p->G1 = 80;
    p->minSize = 11; p->maxSize = 1000;
    p->G2 = 45;
    p->minSizePl = 30; p->minSizeSeg = 21; p->maxSizeSeg = 1000;
    p->fillHolesConnectivity = 4; p->reconConnectivity = 8; p->watershedConnectivity = 8;
}

static const uint8_t kDisk19[361] = {  // HistologicalEntities.cpp:169-188 (MATLAB strel('disk',10))
    0,0,0,0,1,1,1,1,1,1,1,1,1,1,1,0,0,0,0,
    0,0,0,1,1,1,1,1,1,1,1,1,1,1,1,1,0,0,0,
    0,0,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,0,0,
    0,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,0,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,
    0,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,0,
    0,0,1,1,1,1,1,1,1,1,1,1,1,1,1,1,1,0,0,
    0,0,0,1,1,1,1,1,1,1,1,1,1,1,1,1,0,0,0,
    0,0,0,0,1,1,1,1,1,1,1,1,1,1,1,0,0,0,0};
// getStructuringElement(MORPH_ELLIPSE, Size(3,3)) is the 3x3 cross
static const uint8_t kCross3[9] = {0, 1, 0, 1, 1, 1, 0, 1, 0};

// HistologicalEntities::segmentNuclei (HistologicalEntities.cpp:473-598) with
// plFindNucleusCandidates (:115-304) and plSeparateNuclei (:308-449) inlined.
int orc_segment_nuclei(const uint8_t* img, int w, int h, const orc_seg_params* p, int32_t* labels,
                       void** stage, double* stage_ms) {
    using clk = std::chrono::steady_clock;
    auto t_last = clk::now();
    auto lap = [&](int id) {
        auto t = clk::now();
        if (stage_ms) stage_ms[id] += std::chrono::duration<double, std::milli>(t - t_last).count();
        t_last = t;
    };
    const size_t n = (size_t)w * h;
    auto save = [&](int id, const void* src, size_t bytes) {
        if (stage && stage[id]) std::memcpy(stage[id], src, bytes);
    };
    if (!img || w <= 0 || h <= 0) return -1;  // INVALID_IMAGE
    save(0, img, n * 3);

    // ---- S1: plFindNucleusCandidates ----
    std::vector<uint8_t> background(n);
    orc_get_background(img, background.data(), w, h, p->blue, p->green, p->red);
    size_t bgArea = 0;
    for (size_t i = 0; i < n; ++i) bgArea += background[i] != 0;
    float ratio = (float)bgArea / (float)(w * h);  // :125
    lap(1);
    if (ratio >= 0.99) return 1;       // BACKGROUND
    else if (ratio >= 0.9) return 2;   // BACKGROUND_LIKELY
    save(1, background.data(), n);

    std::vector<uint8_t> rbc(n);
    orc_get_rbc(img, rbc.data(), w, h, p->T1, p->T2);
    save(2, rbc.data(), n);
    lap(2);

    std::vector<uint8_t> rc(n), rc_open(n), rc_recon(n), diffIm(n);
    for (size_t i = 0; i < n; ++i) rc[i] = (uint8_t)(255 - img[3 * i + 2]);  // invert(bgr[2]) :158
    orc_morph_open_u8(rc.data(), rc_open.data(), w, h, kDisk19, 19);         // :189
    save(3, rc_open.data(), n);
    lap(3);
    orc_imreconstruct_u8(rc_open.data(), rc.data(), rc_recon.data(), w, h, p->reconConnectivity);  // :207
    save(4, rc_recon.data(), n);
    for (size_t i = 0; i < n; ++i) diffIm[i] = (uint8_t)std::max(0, (int)rc[i] - (int)rc_recon[i]);  // :211
    save(5, diffIm.data(), n);
    lap(4);

    std::vector<uint8_t> diffIm2(n), bw1(n), bw1_t(n), bw2(n), seg_norbc(n);
    for (size_t i = 0; i < n; ++i) diffIm2[i] = diffIm[i] > p->G1 ? 255 : 0;  // :227
    save(6, diffIm2.data(), n);
    orc_imfill_holes_u8(diffIm2.data(), bw1.data(), w, h, p->fillHolesConnectivity);  // :232
    save(7, bw1.data(), n);
    lap(7);
    int compcount2 = orc_bwareaopen2_u8(bw1.data(), bw1_t.data(), w, h, p->minSize, p->maxSize, 8);  // :263
    save(8, bw1_t.data(), n);
    lap(8);
    if (compcount2 == 0) return 3;  // NO_CANDIDATES_LEFT
    for (size_t i = 0; i < n; ++i) bw2[i] = diffIm[i] > p->G2 ? 255 : 0;  // :279
    save(9, bw2.data(), n);
    orc_bwselect_u8(bw2.data(), bw1_t.data(), seg_norbc.data(), w, h, 8);  // :292
    save(10, seg_norbc.data(), n);
    for (size_t i = 0; i < n; ++i) seg_norbc[i] = seg_norbc[i] & (rbc[i] == 0 ? 255 : 0);  // :295
    save(11, seg_norbc.data(), n);
    lap(10);

    // ---- segmentNuclei body ----
    std::vector<uint8_t> seg_nohole(n), seg_open(n);
    orc_imfill_holes_u8(seg_norbc.data(), seg_nohole.data(), w, h, 4);  // :494
    save(12, seg_nohole.data(), n);
    lap(12);
    orc_morph_open_u8(seg_nohole.data(), seg_open.data(), w, h, kCross3, 3);  // :498-499
    save(13, seg_open.data(), n);
    lap(13);

    // ---- A4: plSeparateNuclei ----
    std::vector<uint8_t> seg_big_t(n), seg_big(n);
    orc_bwareaopen2_u8(seg_open.data(), seg_big_t.data(), w, h, p->minSizePl, std::numeric_limits<int>::max(), 8);  // :317
    save(14, seg_big_t.data(), n);
    lap(14);
    orc_dilate_u8(seg_big_t.data(), seg_big.data(), w, h, kCross3, 3, 0);  // :327-330
    save(15, seg_big.data(), n);
    lap(15);
    std::vector<float> dist(n), distance(n), distance2(n);
    orc_dist_transform_u8(seg_big.data(), dist.data(), w, h);  // :359
    save(16, dist.data(), n * 4);
    for (size_t i = 0; i < n; ++i) dist[i] = -dist[i];  // :366
    for (size_t i = 0; i < n; ++i) distance[i] = seg_big[i] ? dist[i] : -std::numeric_limits<float>::max();  // :373-374
    save(17, distance.data(), n * 4);
    lap(16);
    orc_imhmin_f32(distance.data(), distance2.data(), w, h, 1.0f, 8);  // :384
    save(18, distance2.data(), n * 4);
    lap(18);
    std::vector<uint8_t> nuclei(n * 3, 0);
    for (size_t i = 0; i < n; ++i)
        if (seg_big[i]) { nuclei[3 * i] = img[3 * i]; nuclei[3 * i + 1] = img[3 * i + 1]; nuclei[3 * i + 2] = img[3 * i + 2]; }  // :407
    save(19, nuclei.data(), n * 3);
    std::vector<int32_t> watermask(n);
    orc_watershed2(nuclei.data(), distance2.data(), watermask.data(), w, h, p->watershedConnectivity);  // :414
    save(20, watermask.data(), n * 4);
    lap(20);
    std::vector<uint8_t> seg_nonoverlap(n), seg(n), fin(n);
    for (size_t i = 0; i < n; ++i) seg_nonoverlap[i] = (watermask[i] >= 0) ? seg_big[i] : 0;  // :423-424
    save(21, seg_nonoverlap.data(), n);

    compcount2 = orc_bwareaopen2_u8(seg_nonoverlap.data(), seg.data(), w, h, p->minSizeSeg, p->maxSizeSeg, 4);  // :548
    lap(23);
    if (compcount2 == 0) return 3;
    save(23, seg.data(), n);
    orc_imfill_holes_u8(seg.data(), fin.data(), w, h, p->fillHolesConnectivity);  // :573
    save(24, fin.data(), n);
    lap(24);
    orc_bwlabel2(fin.data(), labels, w, h, 8);  // :585
    save(25, labels, n * 4);
    lap(25);
    return 0;  // SUCCESS
}

// The chain's compcount and bbox outputs: ConnComponents::boundingBox(w, h, labels, 0, compcount)
// on the final labels, as HistologicalEntities.cpp:589-591 calls it (bgval 0 while the labels use
// -1 for background: the scan stops at the first -1 pixel, ConnComponents.cpp:250-253).  bbox5
// receives the reference's block, 5 x compcount ints with column stride compcount, when
// compcount <= cap.
int orc_segment_nuclei_bbox(const uint8_t* bgr, int w, int h, const orc_seg_params* p, int32_t* labels, int* compcount,
                            int32_t* bbox5, int cap, void** stage, double* stage_ms) {
    const int st = orc_segment_nuclei(bgr, w, h, p, labels, stage, stage_ms);
    if (compcount) *compcount = 0;
    if (st != 0) return st;
    const int n = orc_cc_bounding_box_literal(w, h, labels, 0, nullptr, 0);
    if (compcount) *compcount = n;
    if (bbox5 && n > 0 && n <= cap) {
        std::vector<int32_t> tmp(5 * (size_t)n);
        orc_cc_bounding_box_literal(w, h, labels, 0, tmp.data(), n);
        std::memcpy(bbox5, tmp.data(), sizeof(int32_t) * 5 * n);
    }
    return st;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Per-nucleus features (src/features/ObjFeatures.cpp, Operators.cpp) on a label image and the
// reference's box block (5 x n, column stride n: [label | minx | maxx | miny | maxy]).
extern "C" {

// ObjFeatures::area (ObjFeatures.cpp:15-45): pixels of the label inside its box
void orc_feat_area(const int32_t* bbox5, int n, const int32_t* labels, int w, int h, int32_t* area) {
    (void)h;
    for (int i = 0; i < n; ++i) {
        const int label = bbox5[i], minx = bbox5[n + i], maxx = bbox5[2 * n + i], miny = bbox5[3 * n + i],
                  maxy = bbox5[4 * n + i];
        int a = 0;
        for (int y = miny; y <= maxy; ++y)
            for (int x = minx; x <= maxx; ++x) a += labels[(size_t)y * w + x] == label;
        area[i] = a;
    }
}

// ObjFeatures::ellipse (ObjFeatures.cpp:48-116): moments about the box centre (float offsets,
// double sums), major / minor axis and eccentricity
void orc_feat_ellipse(const int32_t* bbox5, const int32_t* area, int n, const int32_t* labels, int w, int h,
                      double* major, double* minor, double* ecc) {
    (void)h;
    const double frac = 1.0 / 12.0, root = std::sqrt(8.0);
    for (int i = 0; i < n; ++i) {
        const int label = bbox5[i];
        const int minX = bbox5[n + i], maxX = bbox5[2 * n + i], minY = bbox5[3 * n + i], maxY = bbox5[4 * n + i];
        const float midX = (float)(minX + maxX) / 2.0f, midY = (float)(minY + maxY) / 2.0f;
        double sx = 0, sy = 0, sxy = 0, ssqx = 0, ssqy = 0;
        for (int y = minY; y <= maxY; ++y) {
            const float cy = (float)y - midY;
            for (int x = minX; x <= maxX; ++x) {
                const float cx = (float)x - midX;
                if (labels[(size_t)y * w + x] == label) {
                    sx += cx;
                    sy += cy;
                    sxy += cx * cy;
                    ssqx += cx * cx;
                    ssqy += cy * cy;
                }
            }
        }
        const double a = area[i];
        const double xbar = sx / a, ybar = sy / a;
        const double mxx = ((ssqx + a * xbar * xbar - 2.0 * xbar * sx) / a) + frac;
        const double myy = ((ssqy + a * ybar * ybar - 2.0 * ybar * sy) / a) + frac;
        const double mxy = (sxy - ybar * sx - xbar * sy + xbar * ybar * a) / a;
        const double delta = std::sqrt((mxx - myy) * (mxx - myy) + 4.0 * mxy * mxy);
        major[i] = root * std::sqrt(mxx + myy + delta);
        minor[i] = root * std::sqrt(mxx + myy - delta);
        ecc[i] = (2.0 * std::sqrt(std::fabs(major[i] * major[i] - minor[i] * minor[i]) / 4.0)) / major[i];
    }
}

// ObjFeatures::extent_ratio (ObjFeatures.cpp:118-141) and ::circularity (:281-293)
void orc_feat_extent_circ(const int32_t* bbox5, const int32_t* area, const double* perimeter, int n, double* extent,
                          double* circ) {
    for (int i = 0; i < n; ++i) {
        const int width = bbox5[2 * n + i] - bbox5[n + i] + 1, height = bbox5[4 * n + i] - bbox5[3 * n + i] + 1;
        extent[i] = (double)area[i] / (double)(width * height);
        circ[i] = (4.0 * 3.14159265359 * (double)area[i]) / (perimeter[i] * perimeter[i]);
    }
}

// ObjFeatures::perimeter (ObjFeatures.cpp:144-278): marching squares over the box (cell code from
// lower-left, lower-right, upper-right, upper-left as bits 3..0, looked up in a table), plus the
// box's edge and corner half-cells, summed in the reference's order
void orc_feat_perimeter(const int32_t* bbox5, int n, const int32_t* labels, int w, int h, double* perimeter) {
    (void)h;
    double lut[16] = {0};
    const double hs = 0.70710678118, s2 = 1.41421356237;
    lut[1] = lut[2] = lut[4] = lut[8] = hs;
    lut[3] = lut[6] = lut[9] = lut[12] = 1.0;
    lut[7] = lut[11] = lut[13] = lut[14] = hs;
    lut[5] = lut[10] = s2;
    for (int i = 0; i < n; ++i) {
        const int label = bbox5[i];
        const int xmin = bbox5[n + i], xmax = bbox5[2 * n + i], ymin = bbox5[3 * n + i], ymax = bbox5[4 * n + i];
        auto L = [&](int x, int y) { return labels[(size_t)y * w + x] == label ? 1 : 0; };
        double p = 0.0;
        for (int y = ymin; y < ymax; ++y)
            for (int x = xmin; x < xmax; ++x) {
                int m = L(x, y);
                m = (m << 1) | L(x + 1, y);
                m = (m << 1) | L(x + 1, y + 1);
                m = (m << 1) | L(x, y + 1);
                p += lut[m];
            }
        for (int x = xmin; x < xmax; ++x) {
            int m = L(x, ymax);
            m = (m << 1) | L(x + 1, ymax);
            m <<= 2;
            p += lut[m];
            m = L(x + 1, ymin);
            m = (m << 1) | L(x, ymin);
            p += lut[m];
        }
        for (int y = ymin; y < ymax; ++y) {
            int m = L(xmin, y);
            m = (m << 1) | L(xmin, y + 1);
            m <<= 1;
            p += lut[m];
            m = L(xmax, y);
            m <<= 2;
            m = (m << 1) | L(xmax, y + 1);
            p += lut[m];
        }
        p += lut[L(xmin, ymin) << 1];
        p += lut[L(xmax, ymin)];
        p += lut[L(xmax, ymax) << 3];
        p += lut[L(xmin, ymax) << 2];
        perimeter[i] = p;
    }
}

// The histogram statistics of Operators.cpp:281-505 as ObjFeatures::intensityFeatures
// (ObjFeatures.cpp:342-375: mean, max, min, std, entropy, energy, skewness, kurtosis) and
// ::gradientFeatures (:377-418: the same without max and min) use them
static void hist_stats(const int* hist, float* f, bool with_minmax) {
    double hm = 0.0;  // calcMeanFromHistogram
    int ne = 0;
    for (int i = 0; i < 256; ++i) {
        hm += hist[i] * i;
        ne += hist[i];
    }
    const double mean = ne == 0 ? 0.0 : hm / ne;
    unsigned mx = 0, mn = 256;  // calcMaxFromHistogram / calcMinFromHistogram
    for (int i = 255; i >= 0; --i)
        if (hist[i] > 0) { mx = i; break; }
    for (int i = 0; i < 256; ++i)
        if (hist[i] > 0) { mn = i; break; }
    double sum = 0.0, nd = 0;  // calcStdFromHistogram
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
    const double sd = std::sqrt(sq / nd);
    const float nf = (float)ne;  // calcEntropyFromHistogram / calcEnergyFromHistogram (float)
    float ent = 0.0f, energy = 0.0f;
    for (int i = 0; i < 256; ++i) {
        const float p = (float)hist[i] / nf;
        if (hist[i] != 0) ent += p * std::log2(p);
    }
    for (int i = 0; i < 256; ++i) energy += std::pow((float)hist[i] / nf, 2);
    float n2 = 0.f, avg = 0.f;  // calcSkewnessFromHistogram / calcKurtosisFromHistogram (float)
    for (int i = 0; i < 256; ++i) {
        n2 += (float)hist[i];
        avg += i * (float)hist[i];
    }
    avg /= n2;
    float d3 = 0.f, d2 = 0.f, d4 = 0.f;
    for (int i = 0; i < 256; ++i) {
        const float e = (float)hist[i];
        d3 += e * std::pow(i - avg, 3);
        d2 += e * std::pow(i - avg, 2);
        d4 += e * std::pow(i - avg, 4);
    }
    float skew = 0.f, kurt = 0.f;
    if (d3 != 0 && d2 != 0 && n2 != 0) {
        float a = d3 / n2, b = d2 / n2;
        b = std::sqrt(b);
        b = std::pow(b, 3);
        skew = a / b;
    }
    if (d4 != 0 && d2 != 0 && n2 != 0) {
        float a = d4 / n2, b = d2 / n2;
        b = std::pow(b, 2);
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

// Operators::buildHistogram256CPU (Operators.cpp:515-531) of each nucleus, then its statistics:
// out8[8 i + k] = mean, max, min, std, entropy, energy, skewness, kurtosis
void orc_feat_intensity(const int32_t* bbox5, int n, const int32_t* labels, const uint8_t* gray, int w, int h,
                        float* out8) {
    (void)h;
    for (int i = 0; i < n; ++i) {
        int hist[256] = {0};
        const int label = bbox5[i];
        for (int y = bbox5[3 * n + i]; y <= bbox5[4 * n + i]; ++y)
            for (int x = bbox5[n + i]; x <= bbox5[2 * n + i]; ++x)
                if (labels[(size_t)y * w + x] == label) hist[gray[(size_t)y * w + x]]++;
        hist_stats(hist, out8 + 8 * (size_t)i, true);
    }
}

// Operators::gradient (Operators.cpp:568-640): central differences [-0.5 0 0.5] along x and y
// with replicated borders (cvFilter2D), the first and last column (row) doubled into one-sided
// differences, magnitude in float, rounded half up and clamped to uint8
void orc_gradient_u8(const uint8_t* gray, uint8_t* out, int w, int h) {
    auto G = [&](int xx, int yy) {
        xx = std::min(std::max(xx, 0), w - 1);
        yy = std::min(std::max(yy, 0), h - 1);
        return (float)gray[(size_t)yy * w + xx];
    };
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float dx = 0.5f * G(x + 1, y) - 0.5f * G(x - 1, y);
            if (x == 0 || x == w - 1) dx *= 2;
            float dy = 0.5f * G(x, y + 1) - 0.5f * G(x, y - 1);
            if (y == 0 || y == h - 1) dy *= 2;
            float mag = 0.f;
            mag += dx * dx;
            mag += dy * dy;
            mag = std::sqrt(mag);
            int v;
            if (mag < 0.f) v = 0;
            else if (mag > 255.f) v = 255;
            else {
                v = (int)std::floor(mag);
                if (mag - std::floor(mag) >= 0.5f) v += 1;
            }
            out[(size_t)y * w + x] = (uint8_t)v;
        }
}

// ObjFeatures::gradientFeatures (ObjFeatures.cpp:377-418) on the gradient image:
// out6[6 i + k] = mean, std, entropy, energy, skewness, kurtosis
void orc_feat_gradient(const int32_t* bbox5, int n, const int32_t* labels, const uint8_t* gray, int w, int h,
                       float* out6) {
    std::vector<uint8_t> g((size_t)w * h);
    orc_gradient_u8(gray, g.data(), w, h);
    for (int i = 0; i < n; ++i) {
        int hist[256] = {0};
        const int label = bbox5[i];
        for (int y = bbox5[3 * n + i]; y <= bbox5[4 * n + i]; ++y)
            for (int x = bbox5[n + i]; x <= bbox5[2 * n + i]; ++x)
                if (labels[(size_t)y * w + x] == label) hist[g[(size_t)y * w + x]]++;
        hist_stats(hist, out6 + 6 * (size_t)i, false);
    }
}


}  // extern "C"

// ---- colour normalization (src/normalization/Normalization.cpp) ----------------------------
// Paths in this section are relative to the reference's src/normalization/.  The OpenCV 2.4
// matrix expressions are restated with the arithmetic OpenCV 2.4.9 runs for them:
// addWeighted on float images works in double (arithm.cpp addWeighted32f), convertTo
// float -> float scales in float, Mat -= / += scalar and Mat.mul are float operations.

namespace {

// segFG (Normalization.cpp:19-56) for one pixel: the expression
// M00*rd + M01*gd + M02*bd + M03 evaluates as addWeighted(addWeighted(rd, M00, gd, M01, 0),
// 1, bd, M02, M03), each addWeighted rounding its double result to float.
bool norm_fg(uint8_t b, uint8_t g, uint8_t r) {
    static const float M[8] = {-0.154f, 0.035f, 0.549f, -45.718f, -0.057f, -0.817f, 1.170f, -49.887f};
    float d[2];
    for (int k = 0; k < 2; ++k) {
        const float* m = M + 4 * k;
        const float t = (float)((double)r * (double)m[0] + (double)g * (double)m[1] + 0.0);
        d[k] = (float)((double)t * 1.0 + (double)b * (double)m[2] + (double)m[3]);
    }
    return d[0] < d[1];
}

struct NormConsts {
    float k3, k6, k2;  // 1/sqrtf(3), 1/sqrtf(6), 1/sqrtf(2) (Normalization.cpp:112-114, 172-174)
    float s255;        // the float scale convertTo applies for "/= 255" (Normalization.cpp:331)
    NormConsts() {
        k3 = 1 / sqrtf(3.0f);
        k6 = 1 / sqrtf(6.0f);
        k2 = 1 / sqrtf(2.0f);
        s255 = (float)(1. / 255);
    }
};
const NormConsts& nc() {
    static const NormConsts c;
    return c;
}

// bgr2Lab (Normalization.cpp:64-146) of one pixel after PixelClass's float conversion and
// scaling (Normalization.cpp:328-333)
void norm_lab(uint8_t b8, uint8_t g8, uint8_t r8, float lab[3]) {
    const NormConsts& c = nc();
    const float b = (float)b8 * c.s255, g = (float)g8 * c.s255, r = (float)r8 * c.s255;
    const float L = (float)(0.3811 * r + 0.5783 * g + 0.0402 * b);
    const float M = (float)(0.1967 * r + 0.7244 * g + 0.0782 * b);
    const float S = (float)(0.0241 * r + 0.1288 * g + 0.8444 * b);
    const float l = log10f(L), m = log10f(M), s = log10f(S);
    // Matrix2 = diag(k3, k6, k2) * [1 1 1; 1 1 -2; 1 -1 0] (exact products)
    lab[0] = l * c.k3 + m * c.k3 + s * c.k3;
    lab[1] = l * c.k6 + m * c.k6 + s * (-2 * c.k6);
    lab[2] = l * c.k2 + m * (-c.k2) + s * 0.0f;
}

// lab2BGR (Normalization.cpp:162-306) of one pixel
void norm_bgr(const float lab[3], uint8_t out[3]) {
    const NormConsts& c = nc();
    // Matrix1 = [1 1 1; 1 1 -1; 1 -2 0] * diag(k3, k6, k2) (exact products)
    const float ll = lab[0] * c.k3 + lab[1] * c.k6 + lab[2] * c.k2;
    const float lm = lab[0] * c.k3 + lab[1] * c.k6 + lab[2] * (-c.k2);
    const float ls = lab[0] * c.k3 + lab[1] * (-2 * c.k6) + lab[2] * (0 * c.k2);
    float lms[3] = {powf(10.0f, ll), powf(10.0f, lm), powf(10.0f, ls)};
    for (float& v : lms)
        if (!(v <= DBL_MAX && v >= -DBL_MAX)) v = 0.0f;  // NaN and +-inf
    const float L = lms[0], M = lms[1], S = lms[2];
    const float bgrf[3] = {L * 0.0585f + M * -0.2611f + S * 1.2057f,
                           L * -1.2197f + M * 2.3831f + S * -0.1626f,
                           L * 4.4687f + M * -3.5887f + S * 0.1196f};
    for (int k = 0; k < 3; ++k) {
        float v = (float)((double)bgrf[k] * 255.0);
        if (v < 0.0) v = 0.0f;
        if (v > 255.0) v = 255.0f;
        // rndint (Normalization.cpp:149-158): floor, plus one from a fraction of 0.5 up.  A NaN
        // converts to INT_MIN on x86-64, whose low byte is 0.
        if (v != v) {
            out[k] = 0;
            continue;
        }
        const float f = floorf(v);
        int q = (int)f;
        if (v - f >= 0.5f) q += 1;
        out[k] = (uint8_t)q;
    }
}

// The foreground statistics of TransferI / targetParameters (Normalization.cpp:365-394,
// 540-569): pixels that are foreground and whose float product L*a*b is positive; per channel a
// double sum of the values and of their float squares, in raster order.
void norm_stats(const uint8_t* bgr, int w, int h, double mean[3], double stdv[3], long long* count) {
    double sum[3] = {0, 0, 0}, sq[3] = {0, 0, 0};
    long long n = 0;
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint8_t* p = bgr + 3 * i;
        if (!norm_fg(p[0], p[1], p[2])) continue;
        float lab[3];
        norm_lab(p[0], p[1], p[2], lab);
        if (!((lab[0] * lab[1]) * lab[2] > 0)) continue;
        for (int k = 0; k < 3; ++k) {
            sum[k] += lab[k];
            sq[k] += lab[k] * lab[k];
        }
        ++n;
    }
    for (int k = 0; k < 3; ++k) {
        // int count in the reference: images below 2^31 pixels
        mean[k] = sum[k] / (int)n;
        const double var = sq[k] / (int)n - mean[k] * mean[k];
        stdv[k] = sqrt(var);
    }
    if (count) *count = n;
}

}  // namespace

extern "C" {

// segFG (Normalization.cpp:19-56): 255 where the foreground discriminant wins, else 0
void orc_norm_seg_fg(const uint8_t* bgr, uint8_t* out, int w, int h) {
    for (size_t i = 0; i < (size_t)w * h; ++i)
        out[i] = norm_fg(bgr[3 * i], bgr[3 * i + 1], bgr[3 * i + 2]) ? 255 : 0;
}

// bgr2Lab of every pixel (Normalization.cpp:64-146, 328-333): lab[3 i + k]
void orc_norm_bgr2lab(const uint8_t* bgr, float* lab, int w, int h) {
    for (size_t i = 0; i < (size_t)w * h; ++i) norm_lab(bgr[3 * i], bgr[3 * i + 1], bgr[3 * i + 2], lab + 3 * i);
}

// lab2BGR of every pixel (Normalization.cpp:162-306)
void orc_norm_lab2bgr(const float* lab, uint8_t* bgr, int w, int h) {
    for (size_t i = 0; i < (size_t)w * h; ++i) norm_bgr(lab + 3 * i, bgr + 3 * i);
}

// Normalization::targetParameters (Normalization.cpp:502-582); returns the pixel count used
long long orc_norm_target_parameters(const uint8_t* bgr, int w, int h, float* mean3, float* std3) {
    double m[3], s[3];
    long long n;
    norm_stats(bgr, w, h, m, s, &n);
    for (int k = 0; k < 3; ++k) {
        mean3[k] = (float)m[k];
        std3[k] = (float)s[k];
    }
    return n;
}

// Normalization::normalization (Normalization.cpp:436-499): foreground LAB values move to the
// target statistics (TransferI, Normalization.cpp:396-404), background ones are kept, and every
// pixel converts back (lab2BGR).  Background-labelled pixels of the transferred image are exactly
// those copied from o_bg_lab (Normalization.cpp:487), so each pixel depends on itself and on the
// six statistics only.
void orc_norm_normalization(const uint8_t* bgr, int w, int h, const float* target_mean, const float* target_std,
                            uint8_t* out) {
    double m[3], s[3];
    norm_stats(bgr, w, h, m, s, nullptr);
    float meanI[3], ratio[3];
    for (int k = 0; k < 3; ++k) {
        meanI[k] = (float)m[k];
        ratio[k] = target_std[k] / (float)s[k];
    }
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint8_t* p = bgr + 3 * i;
        float lab[3];
        norm_lab(p[0], p[1], p[2], lab);
        if (norm_fg(p[0], p[1], p[2]))
            for (int k = 0; k < 3; ++k) lab[k] = (lab[k] - meanI[k]) * ratio[k] + target_mean[k];
        norm_bgr(lab, out + 3 * i);
    }
}

}  // extern "C"
