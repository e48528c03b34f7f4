// noalloc.cpp -- the gate path with no device allocation
// (fheacc_bootstrap_batch_dev_noalloc, include/fheacc.h).
//
// Its scratch is a file-scope device array, one copy per GPU, held until the
// process ends.  On the host the array's name is a stand-in, not device
// memory: host code uses only the address hipGetSymbolAddress returns for the
// current device, looked up at the entry point's first call on each device.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "fheacc_internal.hpp"

using namespace fheacc;

static __device__ __align__(256) unsigned char noalloc_scratch[FHEACC_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of noalloc_scratch on each device (null: not looked up yet)

// This device's copy of the array; `device` is current.
int scratch_of(int device, void **out) {
    if (device < 0 || device >= kMaxDevices) return fail("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
        if (e != hipSuccess || !ptr)
            return fail(std::string("hipGetSymbolAddress(noalloc_scratch): ") + hipGetErrorString(e));
        g_ptr[device] = ptr;
    }
    *out = g_ptr[device];
    return 0;
}
}  // namespace

// Every scratch word the kernels read is written earlier in the same call
// (the key-switch digits and b words by the blind rotation, the partial sums
// by the gather key switch), so nothing is cleared between calls.  No event is
// recorded (per-kernel profiling is skipped): the call issues kernel launches
// on the context stream and nothing else.
int fheacc_bootstrap_batch_dev_noalloc(fheacc_ctx *c, fheacc_keys *k, const int32_t *gates, const uint64_t *c1,
                                       const uint64_t *c2, uint64_t *out, int64_t B, const fheacc_capture *cap) {
    if (!c || !k || !gates || !c1 || !c2 || !out) return fail("null argument");
    if (k->ctx != c) return fail("keys belong to another context");
    if (!k->bsk || !k->ksk) return fail("bootstrap and key-switch keys must be resident");
    if (B <= 0) return B == 0 ? 0 : fail("negative batch");
    if (B > 0x7fffffff) return fail("batch too large");
    const size_t need = gate_scratch_bytes(c, B);
    if (need > (size_t)FHEACC_NOALLOC_SCRATCH_BYTES)
        return fail("a batch of " + std::to_string(B) + " needs " + std::to_string(need) +
                    " bytes of scratch; the allocation-free gate path holds " +
                    std::to_string((size_t)FHEACC_NOALLOC_SCRATCH_BYTES));
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    void *scratch = nullptr;
    if (int rc = scratch_of(c->device, &scratch)) return rc;
    GateLaunch g;
    g.gates = gates;
    g.c1 = c1;
    g.c2 = c2;
    g.out = out;
    g.B = B;
    g.cap = cap;
    g.scratch = scratch;
    g.stream = c->stream;
    g.profile = false;
    return launch_gate_path(c, k, g);
}
