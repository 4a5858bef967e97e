// PNG files for the tile pipeline's inputs and outputs, with zlib: what the reference does with
// cv::imread (IMREAD_COLOR: any grey / grey+alpha / RGB / RGBA / palette image of 1-16 bits to
// 8-bit BGR: low-depth grey expands to 8 bits, 16-bit samples keep their high byte, alpha is dropped) and cv::imwrite (8-bit grey or
// BGR image to PNG; an int32 label image is first saturated to 8 bits, as imwrite converts images
// of depths PNG cannot store).  Files written here decode to the same pixels as OpenCV's; their
// bytes differ (compression settings).
//
// Also the file-level segmentNuclei (HistologicalEntities.cpp:453-470): read a tile, run the
// chain on the device, write the label image as imwrite stores it.
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/nscale_hip.h"
#include "common.hpp"

namespace {

const uint8_t kSig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }
void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& f, const char* type, const uint8_t* data, size_t n) {
    put32(f, (uint32_t)n);
    const size_t at = f.size();
    f.insert(f.end(), type, type + 4);
    if (n) f.insert(f.end(), data, data + n);
    put32(f, (uint32_t)crc32(0L, f.data() + at, (uInt)(n + 4)));
}

struct File {
    FILE* f;
    explicit File(FILE* p) : f(p) {}
    ~File() {
        if (f) fclose(f);
    }
};

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    return (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
}

}  // namespace

using nsc::set_error;

static int png_write_impl(const char* path, const uint8_t* img, int w, int h, int channels, size_t ld) {
    NSC_ARG(path && img && w > 0 && h > 0 && (channels == 1 || channels == 3), "png_write");
    if (!ld) ld = (size_t)w * channels;
    NSC_ARG(ld >= (size_t)w * channels, "png_write: row stride");
    // scanlines with filter type 0; BGR pixels are stored as RGB
    const size_t row = (size_t)w * channels + 1;
    std::vector<uint8_t> raw(row * h);
    for (int y = 0; y < h; ++y) {
        uint8_t* d = raw.data() + row * y;
        const uint8_t* s = img + ld * y;
        d[0] = 0;
        if (channels == 1) {
            memcpy(d + 1, s, w);
        } else {
            for (int x = 0; x < w; ++x) {
                d[1 + 3 * x] = s[3 * x + 2];
                d[2 + 3 * x] = s[3 * x + 1];
                d[3 + 3 * x] = s[3 * x];
            }
        }
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 3) != Z_OK) {
        set_error("png_write: deflate failed");
        return NSC_ERR_LIMIT;
    }
    std::vector<uint8_t> f(kSig, kSig + 8);
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.push_back(8);                          // bit depth
    ihdr.push_back(channels == 1 ? 0 : 2);      // grey / RGB
    ihdr.push_back(0);                          // deflate
    ihdr.push_back(0);                          // adaptive filtering
    ihdr.push_back(0);                          // no interlace
    chunk(f, "IHDR", ihdr.data(), ihdr.size());
    chunk(f, "IDAT", z.data(), zn);
    chunk(f, "IEND", nullptr, 0);
    File out(fopen(path, "wb"));
    if (!out.f || fwrite(f.data(), 1, f.size(), out.f) != f.size()) {
        set_error("png_write: cannot write %s", path);
        return NSC_ERR_ARG;
    }
    return NSC_OK;
}

static int png_read_impl(const char* path, uint8_t** bgr, int* w_out, int* h_out) {
    NSC_ARG(path && bgr && w_out && h_out, "png_read");
    *bgr = nullptr;
    File in(fopen(path, "rb"));
    if (!in.f) {
        set_error("png_read: cannot open %s", path);
        return NSC_ERR_ARG;
    }
    std::vector<uint8_t> f;
    {
        uint8_t buf[1 << 16];
        size_t n;
        while ((n = fread(buf, 1, sizeof(buf), in.f)) > 0) f.insert(f.end(), buf, buf + n);
    }
    if (f.size() < 8 || memcmp(f.data(), kSig, 8) != 0) {
        set_error("png_read: %s is not a PNG file", path);
        return NSC_ERR_ARG;
    }
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte;
    for (size_t p = 8; p + 12 <= f.size();) {
        const uint32_t n = be32(&f[p]);
        if (p + 12 + (size_t)n > f.size()) break;
        const uint8_t* t = &f[p + 4];
        const uint8_t* d = &f[p + 8];
        if (!memcmp(t, "IHDR", 4) && n >= 13) {
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (!memcmp(t, "PLTE", 4)) {
            plte.assign(d, d + n);
        } else if (!memcmp(t, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + n);
        } else if (!memcmp(t, "IEND", 4)) {
            break;
        }
        p += 12 + (size_t)n;
    }
    const int samples = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    const bool packed = depth < 8;  // 1/2/4-bit grey or palette samples, several per byte
    const bool depth_ok = depth == 8 || (depth == 16 && ctype != 3) ||
                          ((depth == 1 || depth == 2 || depth == 4) && (ctype == 0 || ctype == 3));
    if (!w || !h || !samples || interlace || !depth_ok || (ctype == 3 && plte.size() < 3) ||
        (uint64_t)w * h > (1ull << 31)) {
        set_error("png_read: %s: unsupported PNG (%ux%u, depth %d, colour type %d, interlace %d)", path, w, h, depth,
                  ctype, interlace);
        return NSC_ERR_ARG;
    }
    const int bpp = packed ? 1 : samples * depth / 8;  // filter distance in bytes
    const size_t row = packed ? ((size_t)w * depth + 7) / 8 : (size_t)w * bpp;
    std::vector<uint8_t> raw((row + 1) * h);
    uLongf rn = (uLongf)raw.size();
    if (uncompress(raw.data(), &rn, idat.data(), (uLong)idat.size()) != Z_OK || rn != raw.size()) {
        set_error("png_read: %s: corrupt image data", path);
        return NSC_ERR_ARG;
    }
    // undo the scanline filters in place
    std::vector<uint8_t> zero(row, 0);
    for (uint32_t y = 0; y < h; ++y) {
        uint8_t* c = raw.data() + (row + 1) * y;
        const int ft = c[0];
        uint8_t* s = c + 1;
        const uint8_t* up = y ? raw.data() + (row + 1) * (y - 1) + 1 : zero.data();
        for (size_t i = 0; i < row; ++i) {
            const int a = i >= (size_t)bpp ? s[i - bpp] : 0, b = up[i], cc = i >= (size_t)bpp ? up[i - bpp] : 0;
            int v = s[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, cc); break;
                default:
                    set_error("png_read: %s: bad filter type %d", path, ft);
                    return NSC_ERR_ARG;
            }
            s[i] = (uint8_t)v;
        }
    }
    uint8_t* out = static_cast<uint8_t*>(malloc((size_t)w * h * 3));
    if (!out) {
        set_error("png_read: out of host memory");
        return NSC_ERR_NOMEM;
    }
    const int step = depth / 8;  // 16-bit samples: the high byte comes first
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* s = raw.data() + (row + 1) * y + 1;
        uint8_t* d = out + (size_t)w * 3 * y;
        for (uint32_t x = 0; x < w; ++x) {
            const uint8_t* px = s + (size_t)x * bpp;
            uint8_t r, g, b;
            if (packed) {
                // MSB first; grey expands to 8 bits (png_set_expand_gray_1_2_4_to_8: v * 255 / (2^d - 1))
                const size_t bit = (size_t)x * depth;
                const int v = (s[bit >> 3] >> (8 - depth - (int)(bit & 7))) & ((1 << depth) - 1);
                if (ctype == 0) {
                    r = g = b = (uint8_t)(v * 255 / ((1 << depth) - 1));
                } else {
                    const size_t k = (size_t)v * 3;
                    r = k + 2 < plte.size() ? plte[k] : 0;
                    g = k + 2 < plte.size() ? plte[k + 1] : 0;
                    b = k + 2 < plte.size() ? plte[k + 2] : 0;
                }
            } else if (ctype == 3) {
                const size_t k = (size_t)px[0] * 3;
                r = k + 2 < plte.size() ? plte[k] : 0;
                g = k + 2 < plte.size() ? plte[k + 1] : 0;
                b = k + 2 < plte.size() ? plte[k + 2] : 0;
            } else if (ctype == 0 || ctype == 4) {
                r = g = b = px[0];
            } else {
                r = px[0];
                g = px[step];
                b = px[2 * step];
            }
            d[3 * x] = b;
            d[3 * x + 1] = g;
            d[3 * x + 2] = r;
        }
    }
    *bgr = out;
    *w_out = (int)w;
    *h_out = (int)h;
    return NSC_OK;
}

extern "C" void nsc_free_host(void* p) { free(p); }

namespace nsc {
namespace {
// cv::saturate_cast<uchar>(int): the conversion imwrite applies to an int32 label image for PNG
__global__ void k_labels_to_u8(const int32_t* __restrict__ lab, size_t lab_ld, int w, int h, uint8_t* __restrict__ out,
                               size_t out_ld) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= w) return;
    for (int y = blockIdx.y; y < h; y += gridDim.y) {
        const int32_t v = *(const int32_t*)((const uint8_t*)lab + lab_ld * y + sizeof(int32_t) * x);
        out[out_ld * y + x] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};
}  // namespace
}  // namespace nsc

extern "C" int nsc_labels_to_u8(const int32_t* labels, size_t lab_ld, int w, int h, uint8_t* out, size_t out_ld,
                                void* stream) {
    NSC_CHECK(nsc::ensure_init());
    NSC_ARG(labels && out && w > 0 && h > 0, "labels_to_u8");
    if (!lab_ld) lab_ld = sizeof(int32_t) * (size_t)w;
    if (!out_ld) out_ld = (size_t)w;
    NSC_ARG(lab_ld >= sizeof(int32_t) * (size_t)w && out_ld >= (size_t)w, "labels_to_u8: row strides");
    hipLaunchKernelGGL(nsc::k_labels_to_u8, dim3(nsc::cdiv(w, 256), std::min(h, 4096)), dim3(256), 0,
                       nsc::as_stream(stream), labels,
                       lab_ld, w, h, out, out_ld);
    NSC_HIP(hipGetLastError());
    return NSC_OK;
}

static int segment_file_impl(const char* in, const char* out, const nsc_seg_params* params, int* compcount,
                             int32_t** bbox, void* stream) {
    NSC_ARG(in && out && compcount, "segment_nuclei_file");
    NSC_CHECK(nsc::ensure_init());
    if (bbox) *bbox = nullptr;
    *compcount = 0;
    uint8_t* host = nullptr;
    int w = 0, h = 0;
    if (png_read_impl(in, &host, &w, &h) != NSC_OK) return NSC_SEG_INVALID_IMAGE;  // imread failed
    std::unique_ptr<uint8_t, void (*)(void*)> hold(host, free);
    const hipStream_t s = nsc::as_stream(stream);
    const size_t px = (size_t)w * h;
    nsc::DevBuf dbgr, dlab, dbox, dmask;
    NSC_HIP(hipMalloc(&dbgr.p, 3 * px));
    NSC_HIP(hipMalloc(&dlab.p, sizeof(int32_t) * px));
    NSC_HIP(hipMalloc(&dmask.p, px));
    NSC_HIP(hipMemcpyAsync(dbgr.p, host, 3 * px, hipMemcpyHostToDevice, s));
    nsc_workspace* ws = nullptr;
    NSC_CHECK(nsc_workspace_create(w, h, &ws));
    std::unique_ptr<nsc_workspace, int (*)(nsc_workspace*)> wsh(ws, nsc_workspace_destroy);
    nsc_seg_params p;
    if (params) p = *params; else nsc_seg_params_default(&p);
    // the reference's ConnComponents::boundingBox output (compcount, 5 x compcount block)
    int cap = 1 << 16;
    nsc_seg_io io;
    memset(&io, 0, sizeof(io));
    int rc;
    for (;;) {
        if (bbox) NSC_HIP(hipMalloc(&dbox.p, sizeof(int32_t) * 5 * (size_t)cap));
        io.bbox_mode = bbox ? NSC_BBOX_REFERENCE_CPU : NSC_BBOX_NONE;
        io.bbox5 = static_cast<int32_t*>(dbox.p);
        io.max_boxes = cap;
        rc = nsc_segment_nuclei_io(static_cast<const uint8_t*>(dbgr.p), w, h, &p, static_cast<int32_t*>(dlab.p), &io,
                                   ws, nullptr, nullptr, s);
        if (rc != NSC_ERR_LIMIT || !bbox || io.compcount <= cap) break;
        cap = io.compcount;  // more boxes than the first guess: size the block and run again
        (void)hipFree(dbox.p);
        dbox.p = nullptr;
    }
    if (rc < 0 && rc != NSC_SEG_INVALID_IMAGE && rc != NSC_SEG_RUNTIME_FAILED) return rc;
    if (rc != NSC_SEG_SUCCESS) return rc;
    *compcount = io.compcount;
    if (bbox && io.compcount > 0) {
        *bbox = static_cast<int32_t*>(malloc(sizeof(int32_t) * 5 * (size_t)io.compcount));
        if (!*bbox) {
            set_error("segment_nuclei_file: out of host memory");
            return NSC_ERR_NOMEM;
        }
        NSC_HIP(hipMemcpyAsync(*bbox, dbox.p, sizeof(int32_t) * 5 * (size_t)io.compcount, hipMemcpyDeviceToHost, s));
    }
    NSC_CHECK(nsc_labels_to_u8(static_cast<const int32_t*>(dlab.p), 0, w, h, static_cast<uint8_t*>(dmask.p), 0, s));
    std::vector<uint8_t> mask(px);
    NSC_HIP(hipMemcpyAsync(mask.data(), dmask.p, px, hipMemcpyDeviceToHost, s));
    NSC_HIP(hipStreamSynchronize(s));
    NSC_CHECK(png_write_impl(out, mask.data(), w, h, 1, (size_t)w));
    return NSC_SEG_SUCCESS;
}

// host buffers of image size are std::vectors: an allocation failure becomes an error code, never
// an exception across the C ABI
extern "C" int nsc_png_write(const char* path, const uint8_t* img, int w, int h, int channels, size_t ld) {
    try {
        return png_write_impl(path, img, w, h, channels, ld);
    } catch (const std::bad_alloc&) {
        set_error("png_write: out of host memory");
        return NSC_ERR_NOMEM;
    }
}

extern "C" int nsc_png_read(const char* path, uint8_t** bgr, int* w, int* h) {
    try {
        return png_read_impl(path, bgr, w, h);
    } catch (const std::bad_alloc&) {
        set_error("png_read: out of host memory");
        return NSC_ERR_NOMEM;
    }
}

extern "C" int nsc_segment_nuclei_file(const char* in, const char* out, const nsc_seg_params* params, int* compcount,
                                       int32_t** bbox, void* stream) {
    try {
        return segment_file_impl(in, out, params, compcount, bbox, stream);
    } catch (const std::bad_alloc&) {
        set_error("segment_nuclei_file: out of host memory");
        return NSC_ERR_NOMEM;
    }
}
