// Library-level C entry points: errors, init, and the thin operator wrappers whose
// implementations live in the kernel translation units.
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "common.hpp"

namespace nsc {

static thread_local char g_err[512] = {0};

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
void clear_error() { g_err[0] = 0; }

namespace {
struct SyncState {
    bool blocking = false;
    hipEvent_t ev = nullptr;
    ~SyncState() {
        if (ev) (void)hipEventDestroy(ev);
    }
};
SyncState& sync_state() {
    static thread_local SyncState st;
    return st;
}
}  // namespace

void set_thread_blocking_sync(bool on) { sync_state().blocking = on; }

namespace {
// per-thread pinned word and event for the pass loops' asynchronous count reads
struct CountReader {
    int* host = nullptr;
    hipEvent_t ev = nullptr;
    int dev = -1;
    bool bridge = true;
    bool failed = false;
    ~CountReader() {
        if (ev) (void)hipEventDestroy(ev);
        if (host) (void)hipHostFree(host);
    }
};
CountReader& count_reader() {
    static thread_local CountReader r;
    return r;
}
}  // namespace

bool& thread_serving() {
    static thread_local bool serving = false;
    return serving;
}

bool set_thread_pass_bridge(bool on) {
    const bool prev = count_reader().bridge;
    count_reader().bridge = on;
    return prev;
}

int* count_read_async(const int* d, hipStream_t s) {
    CountReader& r = count_reader();
    if (!r.bridge || r.failed) return nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    if (!r.host && hipHostMalloc(reinterpret_cast<void**>(&r.host), 64, hipHostMallocDefault) != hipSuccess) {
        r.host = nullptr;
        r.failed = true;
        (void)hipGetLastError();
        return nullptr;
    }
    if (!r.ev || r.dev != dev) {
        if (r.ev) (void)hipEventDestroy(r.ev);
        r.ev = nullptr;
        if (hipEventCreateWithFlags(&r.ev, hipEventDisableTiming) != hipSuccess) {
            r.ev = nullptr;
            r.failed = true;
            (void)hipGetLastError();
            return nullptr;
        }
        r.dev = dev;
    }
    if (hipMemcpyAsync(r.host, d, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipEventRecord(r.ev, s) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return r.host;
}

hipError_t count_wait() { return hipEventSynchronize(count_reader().ev); }

Arena*& thread_arena() {
    static thread_local Arena* a = nullptr;
    return a;
}

hipError_t stream_sync(hipStream_t s) {
    SyncState& st = sync_state();
    if (!st.blocking) return hipStreamSynchronize(s);
    if (!st.ev) {
        const hipError_t e = hipEventCreateWithFlags(&st.ev, hipEventBlockingSync | hipEventDisableTiming);
        if (e != hipSuccess) return e;
    }
    hipError_t e = hipEventRecord(st.ev, s);
    if (e != hipSuccess) return e;
    e = hipEventSynchronize(st.ev);
    if (e != hipSuccess) return e;
    // the work is done; this returns at once and completes any pageable host copy bookkeeping
    return hipStreamSynchronize(s);
}

SideStream& side_stream() {
    static thread_local SideStream ss;
    return ss;
}

int ensure_init() {
    static std::once_flag once;
    static int status = NSC_OK;
    std::call_once(once, [] {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) {
            set_error("no HIP device: %s", hipGetErrorString(e));
            status = NSC_ERR_HIP;
            return;
        }
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t thr = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
    });
    return status;
}

}  // namespace nsc

using namespace nsc;

extern "C" {

const char* nsc_version(void) { return "nscale_amd 0.1 (gfx950)"; }
const char* nsc_last_error(void) { return g_err; }
int nsc_init(void) { return ensure_init(); }

int nsc_imreconstruct_u8(const uint8_t* marker, const uint8_t* mask, uint8_t* out, int w, int h, int conn,
                         void* stream) {
    NSC_ARG(marker && mask && out && w > 0 && h > 0 && valid_conn(conn), "imreconstruct_u8");
    NSC_CHECK(ensure_init());
    return recon_impl<uint8_t>(marker, mask, out, w, h, conn, as_stream(stream));
}

int nsc_imreconstruct_f32(const float* marker, const float* mask, float* out, int w, int h, int conn,
                          void* stream) {
    NSC_ARG(marker && mask && out && w > 0 && h > 0 && valid_conn(conn), "imreconstruct_f32");
    NSC_CHECK(ensure_init());
    return recon_impl<float>(marker, mask, out, w, h, conn, as_stream(stream));
}

int nsc_imreconstruct_u8_batched(const nsc_recon_desc* descs, int n, void* stream) {
    NSC_ARG(descs && n > 0, "imreconstruct_u8_batched");
    NSC_CHECK(ensure_init());
    return recon_batched_impl<uint8_t>(descs, n, as_stream(stream));
}

int nsc_imreconstruct_f32_batched(const nsc_recon_desc* descs, int n, void* stream) {
    NSC_ARG(descs && n > 0, "imreconstruct_f32_batched");
    NSC_CHECK(ensure_init());
    return recon_batched_impl<float>(descs, n, as_stream(stream));
}

int nsc_imreconstruct_i32(const int32_t* marker, const int32_t* mask, int32_t* out, int w, int h, int conn,
                          void* stream) {
    NSC_ARG(marker && mask && out && w > 0 && h > 0 && valid_conn(conn), "imreconstruct_i32");
    NSC_CHECK(ensure_init());
    return recon_impl<int32_t>(marker, mask, out, w, h, conn, as_stream(stream));
}

}  // extern "C"
