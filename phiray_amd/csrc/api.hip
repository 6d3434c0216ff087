// The device and its arrays behind the C ABI: phiray's `Gpu` (src/gpu/mod.rs:84-331)
// and `GpuArray<T>` (src/gpu/array.rs:7-188), one HIP stream per device.
#include "phr_runtime.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>

namespace {
thread_local std::string g_error;

// A plain streaming copy, 16 bytes a thread: the reference point the HBM
// byte counters are scaled by.
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
    uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * 256;
    for (; i < n16; i += stride) dst[i] = src[i];
}
}

void phr_set_error(const char* what, hipError_t e) {
    g_error = std::string(what) + ": " + hipGetErrorString(e);
}
void phr_set_error_msg(const std::string& msg) { g_error = msg; }

int phr_array_alloc(phr_gpu* g, uint64_t elem, uint64_t len, phr_array** out) {
    PHR_CHECK(g && out && elem > 0, PHR_ERR_ARG, "array: bad argument");
    phr_array* a = new phr_array();
    a->gpu = g;
    a->elem = elem;
    a->len = len;
    a->cap = std::max<uint64_t>(len, 1);
    uint64_t bytes = std::max<uint64_t>(a->cap * elem, 16);
    hipError_t e = hipMalloc(&a->ptr, bytes);
    if (e == hipSuccess) e = hipMemsetAsync(a->ptr, 0, bytes, g->stream);
    if (e != hipSuccess) {
        if (a->ptr) (void)hipFree(a->ptr);
        delete a;
        phr_set_error("hipMalloc", e);
        return PHR_ERR_HIP;
    }
    *out = a;
    return PHR_OK;
}

void phr_array_release(phr_array* a) {
    if (!a) return;
    if (a->ptr) (void)hipFree(a->ptr);
    delete a;
}

extern "C" {

const char* phr_last_error(void) { return g_error.c_str(); }

uint32_t phr_abi_version(void) { return 1u; }

// Gpu::new (src/gpu/mod.rs:108): None when there is no device.
int phr_gpu_open(int device, phr_gpu** out) {
    PHR_CHECK(out, PHR_ERR_ARG, "phr_gpu_open: null out");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) {
        phr_set_error_msg("phr_gpu_open: no such HIP device");
        return PHR_ERR_UNSUPPORTED;
    }
    PHR_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    PHR_HIP(hipGetDeviceProperties(&prop, device));
    phr_gpu* g = new phr_gpu();
    g->device = device;
    g->cu_count = prop.multiProcessorCount;
    g->max_alloc = prop.totalGlobalMem;
    std::snprintf(g->name, sizeof g->name, "%s", prop.gcnArchName);
    hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete g;
        phr_set_error("hipStreamCreateWithFlags", e);
        return PHR_ERR_HIP;
    }
    *out = g;
    return PHR_OK;
}

// Gpu::shared (src/gpu/mod.rs:113-118): the process's device 0, opened on
// first use and kept until the process ends; every caller gets the same one.
int phr_gpu_shared(phr_gpu** out) {
    PHR_CHECK(out, PHR_ERR_ARG, "phr_gpu_shared: null out");
    static std::mutex m;
    static phr_gpu* dev = nullptr;
    std::lock_guard<std::mutex> lock(m);
    if (!dev) {
        int rc = phr_gpu_open(0, &dev);
        if (rc != PHR_OK) {
            *out = nullptr;
            return rc;
        }
        dev->shared = true;
    }
    *out = dev;
    return PHR_OK;
}

void phr_gpu_close(phr_gpu* g) {
    if (!g || g->shared) return;
    if (g->stream) {
        (void)hipStreamSynchronize(g->stream);
        (void)hipStreamDestroy(g->stream);
    }
    if (g->sums) (void)hipFree(g->sums);
    delete g;
}

// Gpu::wait (src/gpu/mod.rs:272): drain the device's queue.
int phr_gpu_wait(phr_gpu* g) {
    PHR_CHECK(g, PHR_ERR_ARG, "phr_gpu_wait: null");
    PHR_HIP(hipStreamSynchronize(g->stream));
    return PHR_OK;
}

// Gpu::wait_for (src/gpu/mod.rs:292-303): drained within `limit_ns`? The
// queue is polled, never waited on unboundedly.
int phr_gpu_wait_for(phr_gpu* g, uint64_t limit_ns, int* done) {
    PHR_CHECK(g && done, PHR_ERR_ARG, "phr_gpu_wait_for: null");
    auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        hipError_t e = hipStreamQuery(g->stream);
        if (e == hipSuccess) { *done = 1; return PHR_OK; }
        if (e != hipErrorNotReady) { phr_set_error("hipStreamQuery", e); return PHR_ERR_HIP; }
        auto dt = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
        if ((uint64_t)dt >= limit_ns) { *done = 0; return PHR_OK; }
        std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
}

// Gpu::max_binding (src/gpu/mod.rs:253): the largest one buffer may be.
uint64_t phr_gpu_max_binding(phr_gpu* g) { return g ? g->max_alloc : 0; }
int phr_gpu_cu_count(phr_gpu* g) { return g ? g->cu_count : 0; }
const char* phr_gpu_arch(phr_gpu* g) { return g ? g->name : ""; }
void* phr_gpu_stream(phr_gpu* g) { return g ? (void*)g->stream : nullptr; }

// ---- GpuArray --------------------------------------------------------------
int phr_array_new(phr_gpu* g, uint32_t elem, uint64_t cap, phr_array** out) {
    int rc = phr_array_alloc(g, elem, cap, out);
    if (rc == PHR_OK) (*out)->len = 0;
    return rc;
}

int phr_array_zeroed(phr_gpu* g, uint32_t elem, uint64_t len, phr_array** out) {
    return phr_array_alloc(g, elem, len, out);
}

int phr_array_from(phr_gpu* g, uint32_t elem, const void* items, uint64_t len, phr_array** out) {
    PHR_CHECK(items || len == 0, PHR_ERR_ARG, "phr_array_from: null items");
    int rc = phr_array_alloc(g, elem, len, out);
    if (rc != PHR_OK) return rc;
    if (len) {
        hipError_t e = hipMemcpyAsync((*out)->ptr, items, len * elem, hipMemcpyHostToDevice, g->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
        if (e != hipSuccess) {
            phr_array_release(*out);
            *out = nullptr;
            phr_set_error("hipMemcpyAsync", e);
            return PHR_ERR_HIP;
        }
    }
    return PHR_OK;
}

// GpuArray::write (src/gpu/array.rs:47-56): replace the contents; `grown` is 1
// when the buffer was reallocated (to the next power of two), so bindings to
// the old one are stale.
int phr_array_write(phr_array* a, const void* items, uint64_t len, int* grown) {
    PHR_CHECK(a && (items || len == 0), PHR_ERR_ARG, "phr_array_write: null");
    int g = 0;
    if (len > a->cap) {
        uint64_t cap = 1;
        while (cap < len) cap <<= 1;
        void* p = nullptr;
        PHR_HIP(hipStreamSynchronize(a->gpu->stream));
        PHR_HIP(hipMalloc(&p, cap * a->elem));
        (void)hipFree(a->ptr);
        a->ptr = p;
        a->cap = cap;
        g = 1;
    }
    a->len = len;
    if (len) PHR_HIP(hipMemcpyAsync(a->ptr, items, len * a->elem, hipMemcpyHostToDevice, a->gpu->stream));
    PHR_HIP(hipStreamSynchronize(a->gpu->stream));
    if (grown) *grown = g;
    return PHR_OK;
}

// GpuArray::write_at (src/gpu/array.rs:96-107): overwrite len items from
// item `at`, keeping the rest; *written = 0, and nothing written, when they
// reach past the array's length.
int phr_array_write_at(phr_array* a, uint64_t at, const void* items, uint64_t len, int* written) {
    PHR_CHECK(a && written && (items || len == 0), PHR_ERR_ARG, "phr_array_write_at: null");
    *written = 0;
    if (at + len > a->len || at + len < at) return PHR_OK;
    if (len) PHR_HIP(hipMemcpyAsync((char*)a->ptr + at * a->elem, items, len * a->elem, hipMemcpyHostToDevice, a->gpu->stream));
    PHR_HIP(hipStreamSynchronize(a->gpu->stream));
    *written = 1;
    return PHR_OK;
}

// GpuArray::write_front (src/gpu/array.rs:85-94): write_at from item 0.
int phr_array_write_front(phr_array* a, const void* items, uint64_t len, int* written) {
    return phr_array_write_at(a, 0, items, len, written);
}

// GpuArray::split (src/gpu/array.rs:58-83): items over `chunks` arrays of at
// most limit / elem items each, `cut` items a chunk (at least 1); an unused
// chunk holds one placeholder item. PHR_ERR_SIZE (the reference's None) when
// they do not fit. out receives `chunks` handles.
int phr_array_split(phr_gpu* g, uint32_t elem, const void* items, uint64_t len, uint64_t limit, uint32_t chunks,
                    phr_array** out, uint32_t* cut) {
    PHR_CHECK(g && elem && out && cut && (items || len == 0) && chunks >= 1, PHR_ERR_ARG, "phr_array_split: bad argument");
    const uint64_t cap = std::max<uint64_t>(limit / elem, 1);
    PHR_CHECK(len <= (uint64_t)chunks * cap, PHR_ERR_SIZE, "phr_array_split: the items do not fit the chunks");
    const uint64_t c = std::min<uint64_t>(std::max<uint64_t>(len, 1), cap);
    for (uint32_t k = 0; k < chunks; ++k) out[k] = nullptr;
    for (uint32_t k = 0; k < chunks; ++k) {
        const uint64_t lo = k * c, hi = std::min<uint64_t>((k + 1) * c, len);
        int rc = lo < hi ? phr_array_from(g, elem, (const char*)items + lo * elem, hi - lo, &out[k])
                         : phr_array_zeroed(g, elem, 1, &out[k]);
        if (rc != PHR_OK) {
            for (uint32_t j = 0; j < k; ++j) phr_array_free(out[j]);
            for (uint32_t j = 0; j < chunks; ++j) out[j] = nullptr;
            return rc;
        }
    }
    *cut = (uint32_t)c;
    return PHR_OK;
}

// GpuArray::read_range (src/gpu/array.rs:136-187), blocking.
int phr_array_read_range(phr_array* a, uint64_t from, uint64_t count, void* out) {
    PHR_CHECK(a && (out || count == 0), PHR_ERR_ARG, "phr_array_read_range: null");
    PHR_CHECK(from + count <= a->len, PHR_ERR_SIZE, "phr_array_read_range: past the end");
    if (count)
        PHR_HIP(hipMemcpyAsync(out, (const char*)a->ptr + from * a->elem, count * a->elem, hipMemcpyDeviceToHost, a->gpu->stream));
    PHR_HIP(hipStreamSynchronize(a->gpu->stream));
    return PHR_OK;
}

// GpuArray::read_first (src/gpu/array.rs:138-142): the first min(count, len)
// items, their number in *got.
int phr_array_read_first(phr_array* a, uint64_t count, void* out, uint64_t* got) {
    PHR_CHECK(a && got, PHR_ERR_ARG, "phr_array_read_first: null");
    *got = std::min<uint64_t>(count, a->len);
    return phr_array_read_range(a, 0, *got, out);
}

// ---- Clock (src/gpu/clock.rs:17-87) ------------------------------------------
// A pair of HIP events on the device's stream around whatever is enqueued
// between start and stop: the GPU's own time, and the wall time to submit.
struct phr_clock {
    phr_gpu* gpu = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    std::chrono::steady_clock::time_point t0;
};

int phr_clock_new(phr_gpu* g, phr_clock** out) {
    PHR_CHECK(g && out, PHR_ERR_ARG, "phr_clock_new: null");
    phr_clock* c = new phr_clock();
    c->gpu = g;
    if (hipEventCreate(&c->a) != hipSuccess || hipEventCreate(&c->b) != hipSuccess) {
        if (c->a) (void)hipEventDestroy(c->a);
        delete c;
        phr_set_error_msg("phr_clock_new: hipEventCreate failed");
        return PHR_ERR_HIP;
    }
    *out = c;
    return PHR_OK;
}

int phr_clock_start(phr_clock* c) {
    PHR_CHECK(c, PHR_ERR_ARG, "phr_clock_start: null");
    c->t0 = std::chrono::steady_clock::now();
    PHR_HIP(hipEventRecord(c->a, c->gpu->stream));
    return PHR_OK;
}

// Clock::run's read (src/gpu/clock.rs:54-72): stop, wait for the device, and
// report the stamped span and the wall time since start.
int phr_clock_stop(phr_clock* c, double* gpu_s, double* wall_s) {
    PHR_CHECK(c, PHR_ERR_ARG, "phr_clock_stop: null");
    PHR_HIP(hipEventRecord(c->b, c->gpu->stream));
    PHR_HIP(hipEventSynchronize(c->b));
    float ms = 0.f;
    PHR_HIP(hipEventElapsedTime(&ms, c->a, c->b));
    if (gpu_s) *gpu_s = ms * 1e-3;
    if (wall_s) *wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - c->t0).count();
    return PHR_OK;
}

void phr_clock_free(phr_clock* c) {
    if (!c) return;
    (void)hipEventDestroy(c->a);
    (void)hipEventDestroy(c->b);
    delete c;
}

uint64_t phr_array_len(const phr_array* a) { return a ? a->len : 0; }
uint64_t phr_array_cap(const phr_array* a) { return a ? a->cap : 0; }
uint32_t phr_array_elem(const phr_array* a) { return a ? (uint32_t)a->elem : 0; }
void* phr_array_device_ptr(phr_array* a) { return a ? a->ptr : nullptr; }

// Copy `src` into `dst` (same size, a multiple of 16 bytes) with a plain
// streaming kernel, enqueued on the device's stream.
int phr_array_copy_kernel(phr_array* dst, phr_array* src) {
    PHR_CHECK(dst && src && dst->gpu == src->gpu, PHR_ERR_ARG, "phr_array_copy_kernel: arrays of one device");
    uint64_t bytes = src->len * src->elem;
    PHR_CHECK(bytes == dst->len * dst->elem && bytes % 16 == 0, PHR_ERR_SIZE, "phr_array_copy_kernel: equal sizes, 16-byte multiple");
    uint64_t n16 = bytes / 16;
    uint32_t blocks = (uint32_t)std::min<uint64_t>((n16 + 255) / 256, 8192);
    if (n16) hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, dst->gpu->stream, (const uint4*)src->ptr, (uint4*)dst->ptr, n16);
    PHR_HIP(hipGetLastError());
    return PHR_OK;
}
void phr_array_free(phr_array* a) { phr_array_release(a); }

}  // extern "C"
