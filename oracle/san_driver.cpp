// Sanitizer driver for the CPU oracle (test infrastructure): runs every oracle entry point on seeded
// synthetic inputs and on the edge shapes the reference's operators meet (1 x 1, single rows and
// columns, all-background and all-foreground planes), so that an AddressSanitizer /
// UndefinedBehaviorSanitizer build of oracle.cpp (oracle/Makefile, target `sanitize`) reports any
// out-of-bounds access, use after free, leak or undefined arithmetic.  Exit status 0 = clean run;
// the sanitizers abort the process on the first report.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

struct Rng {  // splitmix64
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    int range(int lo, int hi) { return lo + (int)(uni() * (hi - lo + 1)); }  // [lo, hi]
};

// BGR tile: pale background RGB (230,220,230), purple nuclei ellipses RGB (90,60,140) +-20 (one per
// 1600 px^2), red discs RGB (200,80,90), then +-6 noise on every channel
std::vector<uint8_t> he_tile(int w, int h, uint64_t seed) {
    Rng r{seed};
    std::vector<uint8_t> img((size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        img[3 * i + 0] = 230;
        img[3 * i + 1] = 220;
        img[3 * i + 2] = 230;
    }
    const int nuclei = std::max(1, w * h / 1600);
    for (int k = 0; k < nuclei; ++k) {
        const double cx = r.uni() * w, cy = r.uni() * h, a = 4 + r.uni() * 10, b = 4 + r.uni() * 10, t = r.uni() * M_PI;
        const int c0 = 140 + r.range(-20, 20), c1 = 60 + r.range(-20, 20), c2 = 90 + r.range(-20, 20);
        for (int y = std::max(0, (int)(cy - 15)); y < std::min(h, (int)(cy + 15)); ++y)
            for (int x = std::max(0, (int)(cx - 15)); x < std::min(w, (int)(cx + 15)); ++x) {
                const double dx = x - cx, dy = y - cy;
                const double u = dx * std::cos(t) + dy * std::sin(t), v = -dx * std::sin(t) + dy * std::cos(t);
                if (u * u / (a * a) + v * v / (b * b) <= 1.0) {
                    uint8_t* p = &img[3 * ((size_t)y * w + x)];
                    p[0] = (uint8_t)c0;
                    p[1] = (uint8_t)c1;
                    p[2] = (uint8_t)c2;
                }
            }
    }
    const int rbc = std::max(1, w * h / 5000);
    for (int k = 0; k < rbc; ++k) {
        const int cx = r.range(0, w - 1), cy = r.range(0, h - 1), rad = r.range(3, 6);
        for (int y = std::max(0, cy - rad); y <= std::min(h - 1, cy + rad); ++y)
            for (int x = std::max(0, cx - rad); x <= std::min(w - 1, cx + rad); ++x)
                if ((x - cx) * (x - cx) + (y - cy) * (y - cy) <= rad * rad) {
                    uint8_t* p = &img[3 * ((size_t)y * w + x)];
                    p[0] = 90;
                    p[1] = 80;
                    p[2] = 200;
                }
    }
    for (auto& c : img) c = (uint8_t)std::min(255, std::max(0, (int)c + r.range(-6, 6)));
    return img;
}

std::vector<uint8_t> noise_u8(int w, int h, uint64_t seed, int lo, int hi) {
    Rng r{seed};
    std::vector<uint8_t> v((size_t)w * h);
    for (auto& x : v) x = (uint8_t)r.range(lo, hi);
    return v;
}

std::vector<uint8_t> blobs(int w, int h, uint64_t seed) {  // binary 0/255 blobs
    Rng r{seed};
    std::vector<uint8_t> v((size_t)w * h, 0);
    const int n = std::max(1, w * h / 200);
    for (int k = 0; k < n; ++k) {
        const int cx = r.range(0, w - 1), cy = r.range(0, h - 1), rad = r.range(0, 5);
        for (int y = std::max(0, cy - rad); y <= std::min(h - 1, cy + rad); ++y)
            for (int x = std::max(0, cx - rad); x <= std::min(w - 1, cx + rad); ++x)
                if ((x - cx) * (x - cx) + (y - cy) * (y - cy) <= rad * rad) v[(size_t)y * w + x] = 255;
    }
    return v;
}

// every operator on one shape
void operators(int w, int h, uint64_t seed, const std::vector<uint8_t>& bw) {
    const size_t n = (size_t)w * h;
    const std::vector<uint8_t> mask = noise_u8(w, h, seed, 0, 255);
    std::vector<uint8_t> marker(n), o8(n), o8b(n);
    for (size_t i = 0; i < n; ++i) marker[i] = mask[i] > 10 ? (uint8_t)(mask[i] - 10) : 0;
    std::vector<float> fm(n), fk(n), fo(n);
    std::vector<int32_t> im(n), ik(n), io(n), lab(n), lab2(n);
    for (size_t i = 0; i < n; ++i) {
        fk[i] = (float)mask[i] * 0.5f;
        fm[i] = fk[i] - 1.0f;
        ik[i] = (int32_t)mask[i] * 1000 - 7;
        im[i] = ik[i] - 3000;
    }
    for (int conn : {4, 8}) {
        orc_imreconstruct_u8(marker.data(), mask.data(), o8.data(), w, h, conn);
        orc_imreconstruct_f32(fm.data(), fk.data(), fo.data(), w, h, conn);
        orc_imreconstruct_i32(im.data(), ik.data(), io.data(), w, h, conn);
        orc_imreconstruct_binary_u8(bw.data(), bw.data(), o8.data(), w, h, conn);
        orc_imfill_u8(bw.data(), blobs(w, h, seed + 1).data(), o8.data(), w, h, conn);
        orc_imfill_holes_u8(bw.data(), o8.data(), w, h, conn);
        orc_imfill_holes_gray_u8(mask.data(), o8.data(), w, h, conn);
        orc_bwselect_u8(bw.data(), blobs(w, h, seed + 2).data(), o8.data(), w, h, conn);
        orc_cc_label(bw.data(), w, h, lab.data(), -1, conn);
        lab2 = lab;
        const int nl = orc_cc_relabel(w, h, lab2.data(), -1);
        std::vector<int32_t> box(5 * (size_t)(nl + 1));
        (void)orc_cc_bounding_box_literal(w, h, lab2.data(), -1, box.data(), nl + 1);
        std::vector<int32_t> thr = lab;
        (void)orc_cc_area_threshold_labeled(lab.data(), w, h, thr.data(), -1, 2, 40);
        (void)orc_bwareaopen2_u8(bw.data(), o8.data(), w, h, 3, 60, conn);
        orc_bwlabel2(bw.data(), lab.data(), w, h, conn);
        orc_imhmin_f32(fk.data(), fo.data(), w, h, 1.0f, conn);
        orc_local_maxima_f32(fk.data(), o8.data(), w, h, conn);
        orc_local_minima_f32(fk.data(), o8.data(), w, h, conn);
    }
    const uint8_t cross[9] = {0, 1, 0, 1, 1, 1, 0, 1, 0};
    std::vector<uint8_t> disk(19 * 19);
    for (int y = 0; y < 19; ++y)
        for (int x = 0; x < 19; ++x) disk[y * 19 + x] = (x - 9) * (x - 9) + (y - 9) * (y - 9) <= 81 ? 1 : 0;
    orc_erode_u8(mask.data(), o8.data(), w, h, cross, 3, 255);
    orc_dilate_u8(mask.data(), o8.data(), w, h, cross, 3, 0);
    orc_morph_open_u8(mask.data(), o8.data(), w, h, disk.data(), 19);
    orc_dist_transform_u8(bw.data(), fo.data(), w, h);
    // watershed inputs: BGR of the mask, markers from labelled blobs
    std::vector<uint8_t> bgr(3 * n);
    for (size_t i = 0; i < n; ++i) bgr[3 * i] = bgr[3 * i + 1] = bgr[3 * i + 2] = mask[i];
    orc_cc_label(bw.data(), w, h, lab.data(), 0, 8);
    std::vector<int32_t> mk(n);
    for (size_t i = 0; i < n; ++i) mk[i] = lab[i] > 0 ? (int32_t)(lab[i] % 97 + 1) : 0;
    orc_cv_watershed(bgr.data(), mk.data(), w, h);
    std::vector<int32_t> border(n, 0);
    orc_border_i32(mk.data(), border.data(), w, h, 0);
    orc_watershed2(bgr.data(), fk.data(), io.data(), w, h, 8);
    orc_get_background(bgr.data(), o8.data(), w, h, 220, 220, 220);
    orc_get_rbc(bgr.data(), o8.data(), w, h, 5.0, 4.0);
    // colour normalization
    std::vector<float> lab3(3 * n);
    orc_norm_seg_fg(bgr.data(), o8.data(), w, h);
    orc_norm_bgr2lab(bgr.data(), lab3.data(), w, h);
    orc_norm_lab2bgr(lab3.data(), bgr.data(), w, h);
    float mean3[3], std3[3];
    (void)orc_norm_target_parameters(bgr.data(), w, h, mean3, std3);
    std::vector<uint8_t> nout(3 * n);
    orc_norm_normalization(bgr.data(), w, h, mean3, std3, nout.data());
}

int chain(int w, int h, uint64_t seed) {
    const std::vector<uint8_t> img = he_tile(w, h, seed);
    orc_seg_params p;
    orc_seg_params_default(&p);
    std::vector<int32_t> labels((size_t)w * h);
    int compcount = 0;
    std::vector<int32_t> bbox(5 * 65536);
    const int st = orc_segment_nuclei_bbox(img.data(), w, h, &p, labels.data(), &compcount, bbox.data(), 65536, nullptr,
                                           nullptr);
    long roots = 0;  // nuclei: labels are the smallest raster index of their nucleus
    for (size_t i = 0; i < labels.size(); ++i) roots += labels[i] == (int32_t)i;
    std::printf("chain %dx%d seed %llu: status %d, %ld nuclei (literal compcount %d)\n", w, h, (unsigned long long)seed, st,
                roots, compcount);
    return st;
}

}  // namespace

int main() {
    const int shapes[][2] = {{1, 1}, {1, 37}, {53, 1}, {2, 2}, {17, 9}, {64, 64}, {130, 67}, {256, 256}};
    uint64_t seed = 2026;
    for (const auto& s : shapes) {
        const int w = s[0], h = s[1];
        const size_t n = (size_t)w * h;
        operators(w, h, seed++, blobs(w, h, seed));
        operators(w, h, seed++, std::vector<uint8_t>(n, 0));    // all background
        operators(w, h, seed++, std::vector<uint8_t>(n, 255));  // all foreground
    }
    (void)chain(512, 512, 2026);
    (void)chain(384, 257, 7);
    (void)chain(64, 64, 11);
    std::printf("sanitizer run clean\n");
    return 0;
}
