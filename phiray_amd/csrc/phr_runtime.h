// Host-side objects behind the C ABI: one device + stream (`Gpu`), device
// arrays (`GpuArray`), and the error slot every entry point reports into.
#pragma once
#include "phr_common.h"
#include "phiray_hip.h"

#include <string>

struct phr_gpu {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 0;
    uint64_t max_alloc = 0;
    char name[256] = {0};
    bool shared = false;  // the process's device (Gpu::shared): close leaves it open
    double* sums = nullptr;  // per-block partial sums of the reductions read back to the host (made on first use)
};

// A device buffer grown by doubling (src/gpu/array.rs:7-56): `len` items in
// use, `cap` allocated.
struct phr_array {
    phr_gpu* gpu = nullptr;
    void* ptr = nullptr;
    uint64_t elem = 0, len = 0, cap = 0;
};

void phr_set_error(const char* what, hipError_t e);
void phr_set_error_msg(const std::string& msg);

#define PHR_HIP(x)                                  \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) {                     \
            phr_set_error(#x, e_);                  \
            return PHR_ERR_HIP;                     \
        }                                           \
    } while (0)

#define PHR_CHECK(cond, code, msg)                  \
    do {                                            \
        if (!(cond)) {                              \
            phr_set_error_msg(msg);                 \
            return (code);                          \
        }                                           \
    } while (0)

// Internal helpers shared by the translation units.
int phr_array_alloc(phr_gpu* g, uint64_t elem, uint64_t len, phr_array** out);
void phr_array_release(phr_array* a);
