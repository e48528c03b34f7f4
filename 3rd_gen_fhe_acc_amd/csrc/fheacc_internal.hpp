// fheacc_internal.hpp -- context, key set and kernel argument structures.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fheacc.h"

namespace fheacc {

// Derived parameter block: what the reference keeps in its globals
// (FHE_ACC_runtime/include/params.h:13-134), one instance per set.
struct HostParams {
    int32_t n;
    uint64_t q;
    int32_t N, logN;
    uint64_t Q, iN, ROU;
    uint32_t BGbits, digitG;
    uint64_t Br;
    uint32_t digitR;
    uint64_t Bks;
    uint32_t digitKS;
    uint32_t Bks_R_bits, digitKS_R;
    uint64_t E, E_rlwe;
    double stddev;
    uint64_t gate_constant[6];
    std::vector<uint64_t> rou, irou;  // reference tables (params.cpp:149-175)
    std::vector<uint64_t> cdt;        // Gaussian magnitude table (2^63 scale)
    bool narrow;                      // Q < 2^30: 32-bit words on device
    bool harvey;                      // narrow with Harvey's forward butterflies (28- to 30-bit Q)
};

int derive_params(const fheacc_params *in, HostParams *out, std::string *err);

// Per-ciphertext status flags of the gate path (the context's status word
// ORs them over a batch; the serving loop keeps one per ciphertext).
enum : int { FHEACC_STATUS_SAME = 1, FHEACC_STATUS_GATE = 2 };

// Host mirror of the device sampler (identical word streams).
uint64_t rng_word(uint64_t seed, uint64_t stream, uint64_t ctr);
int64_t gauss_sample(const HostParams &p, uint64_t word);

enum : uint64_t {
    TAG_LWE_SK = 1, TAG_RLWE_SK = 2, TAG_BSK = 3, TAG_KSK = 4, TAG_SUBS = 5,
    TAG_LWE_ENC = 6, TAG_RLWE_ENC = 7, TAG_RGSW_ENC = 8
};
__host__ __device__ inline uint64_t stream_id(uint64_t tag, uint64_t idx) { return (tag << 56) | idx; }

struct DeviceTables {
    void *fw = nullptr, *fwp = nullptr, *iw = nullptr, *iwp = nullptr;  // W-typed [N]
    uint64_t iN = 0, iNp = 0;
    uint64_t iNw = 0, iNwp = 0;  // N^-1 irou[1] mod Q and its Shoup companion
    uint64_t *cdt = nullptr;  // [cdt_len]
    int cdt_len = 0;
};

// Tuning knobs picked per parameter set by the autotuner (fheacc_autotune).
struct Tuning {
    int br_threads = 256;     // threads per blind-rotation workgroup
    int br_pb = 0;            // digit polys per NTT batch (0: as many as fit)
    int ks_threads = 0;       // staged key switch: ciphertexts per workgroup (0: ks_staged_threads)
    int ks_gather_max = 32;   // batches up to this size use the gather key switch
    // Batches of at most small_max ciphertexts (one workgroup per CU) leave
    // most of the chip idle; they run at br_threads_small threads per gate,
    // which shortens each gate (0: br_threads).  Set by the autotuner.
    int br_threads_small = 0;
    int64_t small_max = 0;
    // Blind-rotation schedule of dense batches (fheacc_ctx_set_schedule):
    // br_persist > 0 splits the batch into tiles of that many workgroups per
    // CU (at most the occupancy), one launch each; with br_sync_window > 0
    // the workgroups of a group keep within that many steps of each other,
    // polling a counter at most br_spin_limit times.
    int br_persist = 0;
    int br_sync_window = 0;
    int br_spin_limit = 256;
    // fheacc_ctx_set_tuning was called: the CMUX chains take br_threads too.
    // Otherwise they run 1024 threads from N = 1024 up (one workgroup fills
    // a CU's LDS there, and 1024 threads give it 16 waves).
    bool shape_set = false;
};

// Error text of the calling thread (fheacc_last_error) and the C ABI's
// failure return.
int fail(const std::string &m, int rc = FHEACC_ERR);

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            return ::fheacc::fail(std::string(#expr) + ": " + hipGetErrorString(e_),       \
                                  e_ == hipErrorOutOfMemory ? FHEACC_ENOMEM : FHEACC_ERR);  \
        }                                                                                   \
    } while (0)

// One gate-path launch (blind rotation + key switch + mod switch) on an
// explicit stream and scratch buffer.  Dense batches leave the gathered
// fields null; the serving loop fills them.
struct GateLaunch {
    const int32_t *gates = nullptr;
    const uint64_t *c1 = nullptr, *c2 = nullptr;
    uint64_t *out = nullptr;
    int64_t B = 0;
    const fheacc_capture *cap = nullptr;
    // gathered form: operand / output rows and per-ciphertext key sets
    const int64_t *idx1 = nullptr, *idx2 = nullptr, *oidx = nullptr;
    const int64_t *gidx = nullptr;  // gate of ciphertext c at gates[gidx[c]] (null: gates[c])
    int64_t row_words = 0;          // input and output row stride in words (0: n + 1)
    const int32_t *key_of = nullptr;
    const void *const *bsk_tab = nullptr;  // device array of device key pointers
    int *ct_status = nullptr;              // device [B]: 1 where the operands are identical
    // key-switch segments: ciphertexts [seg_off[s], seg_off[s+1]) use seg_ksk[s]
    int nseg = 0;
    const int64_t *seg_off = nullptr;  // host
    const void *const *seg_ksk = nullptr;  // host array of device pointers
    void *scratch = nullptr;  // gate_scratch_bytes(ctx, B) bytes of device memory
    hipStream_t stream = nullptr;
    bool profile = false;
};
size_t gate_scratch_bytes(const fheacc_ctx *c, int64_t B);
// lock-step groups of a paced tile: workgroup w is in
// group w mod 8, one group per XCD when workgroups are dealt round robin
constexpr int kPacingGroups = 8;
// Launches only: no locking, no allocation, no synchronisation.
int launch_gate_path(fheacc_ctx *c, const fheacc_keys *k, const GateLaunch &g);

// Lifetimes: a context is freed when its creator and every key set, graph and
// server made from it have let go of it; a key set when its owner and every
// graph and server that uses it have.
void ctx_retain(fheacc_ctx *c);
void ctx_release(fheacc_ctx *c);
void keys_retain(fheacc_keys *k);
void keys_release(fheacc_keys *k);

}  // namespace fheacc

struct fheacc_ctx {
    fheacc::HostParams p;
    fheacc::DeviceTables t;
    fheacc::Tuning tune;
    int device = 0;
    int cus = 0;                    // compute units of the device
    size_t lds_budget = 64 * 1024;  // bytes of LDS one workgroup may use
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // per-context scratch, grown on demand
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    int *status = nullptr;  // device status word (identical-operand flag)
    // blind-rotation order, FHEACC_EVAL_*: read at every launch and capture,
    // including by a serving loop's thread
    std::atomic<int> eval_mode{FHEACC_EVAL_SOFTWARE};
    std::mutex mu;
    // profiling: (class, start, stop) event records while enabled
    bool profile = false;
    struct EvRec {
        int cls;
        hipEvent_t a, b;
    };
    std::vector<EvRec> events;
    std::atomic<int> refs{1};
};

struct fheacc_keys {
    fheacc_ctx *ctx = nullptr;
    void *bsk = nullptr;   // device layout, see kernels_impl.hpp
    void *ksk = nullptr;
    void *subs = nullptr;
    size_t bsk_bytes = 0, ksk_bytes = 0, subs_bytes = 0;
    uint64_t bsk_digest = 0, ksk_digest = 0;
    std::atomic<int> refs{1};
};
