// Shared host/device helpers for libnscale_hip (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "nscale_hip.h"

namespace nsc {

// ---- error reporting --------------------------------------------------------------------
void set_error(const char* fmt, ...);
void clear_error();

#define NSC_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            ::nsc::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),      \
                             __FILE__, __LINE__);                                        \
            return NSC_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)

#define NSC_CHECK(call)                 \
    do {                                \
        int r_ = (call);                \
        if (r_ != NSC_OK) return r_;    \
    } while (0)

#define NSC_ARG(cond, msg)                                   \
    do {                                                     \
        if (!(cond)) {                                       \
            ::nsc::set_error("invalid argument: %s", msg);   \
            return NSC_ERR_ARG;                              \
        }                                                    \
    } while (0)

// int32 labels and root indices address pixels, so entry points that produce or consume them
// reject images of INT_MAX pixels or more
#define NSC_LABEL_SIZE(w, h) \
    NSC_ARG((size_t)(w) * (size_t)(h) < (size_t)0x7FFFFFFF, "image too large for int32 labels (w*h must be below 2^31-1)")

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

int ensure_init();

// ---- stream-ordered scratch memory -------------------------------------------------------
// Scratch comes from the calling thread's arena when one is active (a chain call activates its
// workspace's arena: a device buffer allocated once, handed out as a stack -- scopes nest, so
// frees arrive in reverse order, and a buffer freed out of order is reclaimed when everything
// above it is gone), else from the device's default memory pool (hipMallocAsync, release
// threshold raised at init).  Either way the memory is stream-ordered: reuse by a later
// operator on the same stream is safe without synchronisation.
struct Arena {
    uint8_t* base = nullptr;
    size_t cap = 0, top = 0;
    struct Entry { size_t off, end; bool live; };
    Entry stack[64];
    int depth = 0;
    void* take(size_t bytes) {
        const size_t off = (top + 255) & ~(size_t)255;
        if (!base || depth == 64 || off + bytes > cap) return nullptr;
        stack[depth++] = Entry{off, off + bytes, true};
        top = off + bytes;
        return base + off;
    }
    bool give(void* p) {
        if (!base || p < (void*)base || p >= (void*)(base + cap)) return false;
        const size_t off = (size_t)((uint8_t*)p - base);
        for (int k = depth - 1; k >= 0; --k)
            if (stack[k].off == off && stack[k].live) { stack[k].live = false; break; }
        while (depth > 0 && !stack[depth - 1].live) --depth;
        top = depth ? stack[depth - 1].end : 0;
        return true;
    }
};
Arena*& thread_arena();

// activates an arena for the calling thread for the lifetime of the guard
struct ArenaScope {
    Arena* prev;
    explicit ArenaScope(Arena* a) : prev(thread_arena()) { thread_arena() = a; }
    ~ArenaScope() { thread_arena() = prev; }
};

struct Scratch {
    void* p = nullptr;
    hipStream_t s = nullptr;
    bool from_arena = false;
    Scratch() = default;
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    ~Scratch() { release(); }
    int alloc(size_t bytes, hipStream_t stream) {
        release();
        s = stream;
        if (bytes == 0) bytes = 256;
        if (Arena* a = thread_arena()) {
            if ((p = a->take(bytes)) != nullptr) {
                from_arena = true;
                return NSC_OK;
            }
        }
        hipError_t e = hipMallocAsync(&p, bytes, stream);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMallocAsync(%zu) failed: %s", bytes, hipGetErrorString(e));
            return NSC_ERR_NOMEM;
        }
        return NSC_OK;
    }
    void release() {
        if (p) {
            if (from_arena) {
                Arena* a = thread_arena();
                if (!a || !a->give(p)) { /* arena gone: nothing to return */ }
            } else {
                (void)hipFreeAsync(p, s);
            }
        }
        p = nullptr;
        from_arena = false;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// Carves 256-byte-aligned sub-buffers out of one allocation; construct with base = nullptr to
// measure the total size first.
struct Carver {
    uint8_t* base;
    size_t off = 0;
    explicit Carver(void* b) : base(static_cast<uint8_t*>(b)) {}
    template <typename T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
    size_t size() const { return off + 256; }
};

inline int valid_conn(int c) { return c == 4 || c == 8; }

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ---- fork/join onto a side stream ---------------------------------------------------------
// Independent launches (e.g. long-latency work on a few runs beside many short ones) can overlap
// on per-thread side streams (default priority); join() makes the caller's stream wait for them.
struct SideStream {
    static constexpr int K = 3;
    hipStream_t side[K] = {nullptr, nullptr, nullptr};
    hipEvent_t forked = nullptr, joined[K] = {nullptr, nullptr, nullptr};
    int fork(hipStream_t s) {
        if (!forked) {
            NSC_HIP(hipEventCreateWithFlags(&forked, hipEventDisableTiming));
            for (int k = 0; k < K; ++k) {
                NSC_HIP(hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking));
                NSC_HIP(hipEventCreateWithFlags(&joined[k], hipEventDisableTiming));
            }
        }
        NSC_HIP(hipEventRecord(forked, s));
        for (int k = 0; k < K; ++k) NSC_HIP(hipStreamWaitEvent(side[k], forked, 0));
        return NSC_OK;
    }
    int join(hipStream_t s) {
        for (int k = 0; k < K; ++k) {
            NSC_HIP(hipEventRecord(joined[k], side[k]));
            NSC_HIP(hipStreamWaitEvent(s, joined[k], 0));
        }
        return NSC_OK;
    }
};
// one per host thread: concurrent callers never share a side stream
SideStream& side_stream();

// Wait for the stream.  By default the thread spins (lowest latency); a thread that has asked for
// blocking waits (the server's workers, many of which wait at once) sleeps on an event instead.
hipError_t stream_sync(hipStream_t s);
void set_thread_blocking_sync(bool on);
// Pass loops (reconstruction, geodesic distance) read each check's list length through a pinned
// word and queue a short bridge of further passes before waiting, so the GPU keeps working during
// the host round trip (passes after the last busy one find empty lists and exit).  On by default;
// the server's slot threads turn it off (other slots fill those gaps, and the bridge's empty
// passes would take their GPU time), and so does the no-alloc reconstruction (the bridge's pinned
// word is allocated on a thread's first use).  set_thread_pass_bridge returns the previous setting.
// count_read_async queues the copy of *d and returns the host word (valid after count_wait), or
// nullptr when the bridge is off or unavailable.
bool set_thread_pass_bridge(bool on);
// Set on the server's slot threads, whose chains share the GPU with other slots: host round trips
// cost those threads little there, and kernels sized to their work leave more of the GPU to the
// other slots (the watershed reads its phase-1 tier counts and launches each tier's grid exactly).
bool& thread_serving();
int* count_read_async(const int* d, hipStream_t s);
hipError_t count_wait();

// ---- device helpers ----------------------------------------------------------------------
// max/min with the selection rule of std::max / std::min (first argument on ties)
template <typename R>
__device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }
template <typename R>
__device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
// The same pointer in the global address space: loads and stores through it compile to global_*
// instructions instead of flat ones (flat accesses also count against the LDS wait counter).
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* as_global(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Counters shared by a whole launch: atomics on one address from every wave serialise (one
// address lives in one L2 channel), so counts are summed per workgroup first and added with one
// atomic per workgroup.  Every thread of the block calls these, with the same control flow.
__device__ __forceinline__ int block_sum(int v) {
    __shared__ int part[16];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    const int wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();  // part[] may still be read by a previous call
    if ((threadIdx.x & 63) == 0) part[wv] = v;
    __syncthreads();
    int s = 0;
    for (int k = 0; k < nw; ++k) s += part[k];
    return s;
}
__device__ __forceinline__ void block_count_add(int v, int* counter) {
    const int s = block_sum(v);
    if (threadIdx.x == 0 && s) atomicAdd(counter, s);
}
// appends v for the threads whose `take` is set: one atomic per workgroup, list order arbitrary
__device__ __forceinline__ void block_append(bool take, int v, int* __restrict__ list, int* __restrict__ count) {
    __shared__ int wbase[16];
    __shared__ int bbase;
    const unsigned long long m = __ballot(take);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    __syncthreads();  // wbase / bbase may still be read by a previous call
    if (lane == 0) wbase[wv] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int k = 0; k < nw; ++k) { const int c = wbase[k]; wbase[k] = tot; tot += c; }
        bbase = tot ? atomicAdd(count, tot) : 0;
    }
    __syncthreads();
    if (take) list[bbase + wbase[wv] + __popcll(m & ((1ull << lane) - 1ull))] = v;
}
// grid for a grid-stride pass over n items with one counter update per workgroup
inline dim3 grid_counted(size_t n) { return dim3((unsigned)std::min<size_t>(cdiv(n, 256), 1024)); }
// x where keep is all ones, 0 (all bits clear) where keep is 0.  Written as a bit operation on
// purpose: a select between a loaded value and a constant may be compiled into a branch around
// the load, which serialises the loads of an unrolled window.
template <typename R>
__device__ __forceinline__ R keep_bits(R x, uint32_t keep) {
    static_assert(sizeof(R) == 4, "four-byte values");
    return __builtin_bit_cast(R, __builtin_bit_cast(uint32_t, x) & keep);
}


// value from lane-1 (lane 0 receives `old`) and from lane+1 (lane 63 receives `old`):
// full-wave DPP shifts (wave_shr:1 / wave_shl:1, GFX9 DPP controls 0x138 / 0x130)
__device__ __forceinline__ int dpp_from_left(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ int dpp_from_right(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, 0x130, 0xF, 0xF, false);
}
template <typename R>
__device__ __forceinline__ R from_left(R old, R v) {
    return __builtin_bit_cast(R, dpp_from_left(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v)));
}
template <typename R>
__device__ __forceinline__ R from_right(R old, R v) {
    return __builtin_bit_cast(R, dpp_from_right(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v)));
}

// ---- internal entry points shared between translation units ------------------------------
// connected-component labelling with min-index roots; pred: 0 = nonzero, 1 = equal u8 values,
// 2 = equal float values restricted to fg != 0 (fg given separately)
struct CCL;
int ccl_u8(const uint8_t* img, int32_t* labels, int w, int h, int conn, bool equal_values, int32_t bgval,
           hipStream_t s);
int ccl_mask(const uint8_t* fg, const float* vals, int32_t* labels, int w, int h, int conn, hipStream_t s);

template <typename T>
int recon_impl(const T* marker, const T* mask, T* out, int w, int h, int conn, hipStream_t s);
template <typename T>
int recon_batched_impl(const nsc_recon_desc* d, int n, hipStream_t s);

}  // namespace nsc
