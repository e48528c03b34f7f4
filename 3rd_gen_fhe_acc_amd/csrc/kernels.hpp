// kernels.hpp -- kernel argument blocks and host launchers (kernels_impl.hpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "modarith.hpp"
#include "ntt_lds.hpp"

namespace fheacc {

struct ShoupGen;

// Product forms of the blind rotation (the launcher picks one from pb):
// the whole 2d-poly digit batch in place (pb >= 2d), two in-place batches of
// d polys, the digits of a then of b (pb == d), or an accumulator plus a
// separate pb-poly digit buffer (any other pb).
enum ProductForm { kSeparate = 0, kInPlace = 1, kSplit = 2 };
__host__ __device__ __forceinline__ int product_form(int pb, int d) {
    return pb >= 2 * d ? kInPlace : (pb == d ? kSplit : kSeparate);
}

template <typename W>
struct BRArgs {
    Arith<W> A;
    Twiddles<W> tw;
    const uint64_t *c1, *c2;
    const int32_t *gates;
    const W *bsk;
    size_t rgsw_elems;  // W elements per device RGSW
    int n;
    uint32_t q;
    int digitR;
    uint32_t Br;
    int d, bits, pb;  // pb: digit polys transformed per LDS batch
    uint64_t gate_constant[6];
    uint64_t *cap_acc;
    uint8_t *ksd;
    uint64_t *ksb;
    int dks;
    uint32_t Bks;
    int *status;
    // Gathered batches (the serving loop): ciphertext c reads its operands
    // from rows idx1[c], idx2[c] of c1 / c2 and the bootstrap key
    // bsk_tab[key_of[c]], and reports identical operands in ct_status[c].
    // All null for a dense batch.
    const int64_t *idx1, *idx2;
    const int64_t *gidx;   // gate of ciphertext c at gates[gidx[c]] (null: gates[c])
    int64_t row_words;     // words from one input row to the next (n + 1 when dense)
    const int32_t *key_of;
    const W *const *bsk_tab;
    int *ct_status;
    int mode;  // FHEACC_EVAL_*: which residue-0 steps multiply by RGSW(1)
    int64_t c_base;  // ciphertext of workgroup 0 (a batch split over several launches)
    int64_t c_end;   // one past the last ciphertext of this launch
    // Tiled batches: progress words, one per ciphertext (null: no pacing),
    // cleared by the launcher before each tile; groups of workgroups w mod
    // sync_groups; a step waits for its group's step - sync_window, at most
    // spin_limit polls.
    uint32_t *sync;
    int sync_window, sync_groups, spin_limit;
};

template <typename W>
struct KSArgs {
    Arith<W> A;
    const W *ksk;
    const uint8_t *ksd;
    const uint64_t *ksb;
    int n, n_pad, S;
    uint32_t Bks;
    uint64_t Q;
    int lq1;
    uint64_t *out, *cap_ks;
    const int64_t *oidx;  // output row of ciphertext c (null: row c)
    int64_t out_row_words;  // words from one output row to the next (n + 1 when dense)
    int64_t B;
    int nt;  // staged form: threads per workgroup, the ciphertexts sharing one key stream (256, 512, 1024)
    // gather form: the S steps of a ciphertext split into `split` ranges, one
    // workgroup each, partial sums in part[B][split][n+1] (split > 1 only)
    int split;
    uint64_t *part;
};

template <typename W>
struct NTTArgs {
    Arith<W> A;
    Twiddles<W> tw;
    uint64_t *polys;
    int64_t count;
    int per_wg;
    int inverse;
};

template <typename W>
struct EPArgs {
    Arith<W> A;
    Twiddles<W> tw;
    const uint64_t *acc;
    uint64_t *out;
    const W *key;
    int64_t key_stride;
    int d, bits, pb;
    int64_t c_base;  // ciphertext of workgroup 0
    // words from one input (output) RLWE to the next: 2N for dense batches;
    // homo_expand reads and writes rows of RGSWs in place
    int64_t in_stride, out_stride;
};

template <typename W>
struct SubsArgs {
    Arith<W> A;
    Twiddles<W> tw;
    const uint64_t *in;
    uint64_t *out;
    const W *key;
    int64_t key_stride;
    int sub_factor;
    int d, bits, dch;
    int64_t c_base;  // workgroup index of workgroup 0
    // expansion slots: slot k of expansion b is the RLWE at in + b set_stride
    // + k slot_stride (dense [B][N][2][N]: N 2N and 2N; homo_expand: RGSW rows)
    int64_t set_stride, slot_stride;
};

// CMUX chains (RLWE.cpp:788-838): per ciphertext acc = c0; for each step
// acc = (X(acc) - acc) (x) C[step] + acc, where X(acc) is c1 (one plain
// CMUX) or acc times X^rot[step] (blind rotation).
template <typename W>
struct CmuxArgs {
    Arith<W> A;
    Twiddles<W> tw;
    const uint64_t *c0, *c1;  // ciphertext c at c0 + c in_stride; c1 null: rotation steps
    int64_t in_stride;        // 2N for dense inputs; the CMUX tree reads leaf pairs
    uint64_t *out;            // [B][2][N]
    const W *key;             // device-layout RGSW of (ciphertext c, step s) at
    int64_t key_stride;       //   key + c key_ct_stride + s key_stride
    int64_t key_ct_stride;
    const int32_t *rot;       // [steps] exponents in [0, 2N) (device)
    int steps;
    int pre_rot;              // public rotation applied first (exponent in [0, 2N)), 0 = none
    int d, bits, pb;
    int64_t c_base;           // ciphertext of workgroup 0
};

struct ShoupGen {
    uint64_t Q, a, b, bp;
    __host__ __device__ ShoupGen() : Q(0), a(0), b(0), bp(0) {}
    __device__ __forceinline__ uint64_t operator()(uint64_t k) const {
        uint64_t qh = __umul64hi(k, bp);
        uint64_t r = k * b - qh * Q;
        if (r >= Q) qh++;
        return k * a + qh;
    }
};

template <typename W>
struct KeygenArgs {
    Arith<W> A;
    Twiddles<W> tw;
    uint64_t seed;
    const uint64_t *cdt;
    int cdt_len;
    uint64_t E_rlwe;
    W *zn, *znp;
    W *bsk;              // bootstrap key, or subs key in subs mode
    size_t rgsw_elems;   // per device RGSW (subs mode: per (i, l) RLWE key)
    int d;               // digitG (subs mode: digitKS_R)
    int n;               // LWE n (subs mode: log2 N)
    int digitR;
    uint32_t Br, q;
    const uint64_t *lwe_sk;
    const uint64_t *z;                 // RLWE secret, time domain
    const uint64_t *bg_pow, *bg_powp;  // BG^h (subs mode: Bks_R^l) mod Q
    ShoupGen sg;
};

// RLWE / RGSW encryption of message polynomials and RLWE decryption
// (RLWE.cpp:324-376, 561-577); keygen fields as above (zn, znp: the secret's
// transform and its Shoup companions; bg_pow: BG^h mod Q).
template <typename W>
struct EncArgs {
    KeygenArgs<W> k;
    const uint64_t *msg;   // [count][N] time domain mod Q
    const uint64_t *ct;    // decrypt input [count][2][N]
    uint64_t *out;         // RLWE [count][2][N], RGSW [count][d][2][2][N], or messages [count][N]
    uint64_t first;        // stream index of the first ciphertext
    uint64_t scale;        // decrypt: Q / t
};

template <typename W>
struct KskGenArgs {
    Arith<W> A;
    uint64_t seed;
    const uint64_t *cdt;
    int cdt_len;
    uint64_t E_rlwe;
    W *ksk;
    int64_t rows;
    int n, n_pad, N, dks, Bks;
    const W *s, *sp;
    const uint64_t *z;
    const uint64_t *cij, *cijp;
};

// persist > 0: tiles of min(persist, occupancy) workgroups per CU over `cus`
// CUs, one launch each; 0: one launch of a workgroup per ciphertext.  *grid
// (may be null) receives the workgroups of the first launch.
template <typename W>
hipError_t launch_blind_rotate(const BRArgs<W> &a, int logN, int nt, int64_t B, size_t lds, hipStream_t st,
                               int persist, int cus, int64_t *grid);
template <typename W>
hipError_t launch_lwe_keyswitch(const KSArgs<W> &a, bool gather, hipStream_t st);
template <typename W>
size_t ks_staged_lds(uint32_t Bks, int nt);  // LDS bytes of the staged key switch at nt threads
template <typename W>
hipError_t launch_ntt(const NTTArgs<W> &a, int logN, hipStream_t st);
template <typename W>
hipError_t launch_external_product(const EPArgs<W> &a, int logN, int64_t count, hipStream_t st);
template <typename W>
hipError_t launch_subs_keyswitch(const SubsArgs<W> &a, int logN, int64_t count, hipStream_t st);
// expand_RLWE level sub_factor over the [B][N][2][N] slots in a.out (== a.in)
template <typename W>
hipError_t launch_expand_level(const SubsArgs<W> &a, int logN, int64_t B, hipStream_t st);
template <typename W>
hipError_t launch_cmux_chain(const CmuxArgs<W> &a, int logN, int nt, int64_t count, size_t lds, hipStream_t st);
// RGSW words between the reference's order (ddr = false) or the accelerator's
// DDR image order (ddr = true) and the device quad layout
template <typename W>
hipError_t launch_rgsw_to_device(const uint64_t *src, int64_t src_stride, W *dst, int64_t dst_stride, int d, int N,
                                 ShoupGen sg, const int64_t *map, int64_t count, hipStream_t st, bool ddr = false);
template <typename W>
hipError_t launch_rgsw_to_ref(const W *src, int64_t src_stride, uint64_t *dst, int64_t dst_stride, int d, int N,
                              const int64_t *map, int64_t count, hipStream_t st, bool ddr = false);
template <typename W>
hipError_t launch_ksk_convert(const uint64_t *src, W *dst, uint64_t *back, int Bks, int dks, int N, int n1,
                              int n_pad, hipStream_t st);
template <typename W>
hipError_t launch_keygen(const KeygenArgs<W> &a, const KskGenArgs<W> &ka, int logN, const uint64_t *z, hipStream_t st,
                         bool subs);
// which: 0 RLWE encrypt, 1 RGSW encrypt, 2 RLWE decrypt (k.zn must hold the
// secret's transform: launch_secret_ntt first)
template <typename W>
hipError_t launch_secret_ntt(const KeygenArgs<W> &a, int logN, const uint64_t *z, hipStream_t st);
template <typename W>
hipError_t launch_encrypt(const EncArgs<W> &a, int which, int logN, int64_t count, hipStream_t st);
template <typename W>
hipError_t launch_subs_convert(const void *src, void *dst, int N, int64_t total, bool to_ref, ShoupGen sg,
                               hipStream_t st);
hipError_t launch_digest(const uint64_t *w, int64_t count, uint64_t *partials, int blocks, hipStream_t st);

}  // namespace fheacc
