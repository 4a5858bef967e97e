/*
 * C entry points over the reference's own ConnComponents class (ConnComponents.cpp/.h), which
 * is OpenCV-free.  Built by oracle/Makefile from the reference sources in place, into
 * oracle/_ref/ (not committed).  TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's
 * restatement of ConnComponents against the reference code itself.
 */
#include <cstdlib>
#include <cstring>

#include "ConnComponents.h"

extern "C" {
void ref_cc_label(unsigned char* img, int w, int h, int* label, int bgval, int conn) {
    nscale::ConnComponents cc;
    cc.label(img, w, h, label, bgval, conn);
}
int ref_cc_relabel(int w, int h, int* label, int bgval) {
    nscale::ConnComponents cc;
    return cc.relabel(w, h, label, bgval);
}
int ref_cc_area_threshold_labeled(const int* label, int w, int h, int* n_label, int bgval, int lower, int upper) {
    nscale::ConnComponents cc;
    return cc.areaThresholdLabeled(label, w, h, n_label, bgval, lower, upper);
}
int ref_cc_area_threshold(unsigned char* img, int w, int h, int* label, int bgval, int lower, int upper, int conn) {
    nscale::ConnComponents cc;
    return cc.areaThreshold(img, w, h, label, bgval, lower, upper, conn);
}
int ref_cc_bounding_box(int w, int h, const int* label, int bgval, int* out5, int cap) {
    nscale::ConnComponents cc;
    int n = 0;
    int* b = cc.boundingBox(w, h, label, bgval, n);
    int m = n < cap ? n : cap;
    for (int c = 0; c < 5; ++c) std::memcpy(out5 + (size_t)c * cap, b + (size_t)c * n, sizeof(int) * m);
    std::free(b);
    return n;
}
}
