// kernels_impl.hpp -- HIP kernels of the gate-bootstrapping path for gfx950.
//
// Included by the kernels_*.hip translation units; each instantiates a share
// of the launchers so that the library's device code compiles in parallel.
//
// Device layouts (W = uint32_t for Q < 2^30, else uint64_t):
//   RGSW (external-product key)  [digitG][N] x {K0a, K0b, K1a, K1b}
//        K0 = c_text[h][0] (multiplies the digits of a), K1 = c_text[h][1]
//        (digits of b), RLWE.cpp:642-648.  Narrow: uint4 per (h, k).  Wide:
//        each value followed by its Shoup companion floor(K * 2^64 / Q).
//   bootstrap key  [n][digitR][Br] RGSW: the keys of one blind-rotation
//        step (i, j) are contiguous.  Residue 0 (RGSW(1)) is stored as the
//        reference's key holds it; the software evaluation skips it
//        (RLWE.cpp:759), the FPGA evaluation order multiplies by it
//        (myfpga.cpp:1037-1046,1190-1206).
//   LWE ks key     [N][digitKS][Bks][n_pad] W: the Bks candidate rows of one
//        key-switch step (coefficient i, digit j) are contiguous.
//   subs key       [log2 N][digitKS_R][N] x {Ka, Kb} (+ companions if wide).
//
// One ciphertext per workgroup for every polynomial kernel; polynomials live
// in LDS and are transformed by ntt_lds.hpp.  Nothing here is floating
// point: no MFMA.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <type_traits>

#include "fheacc_internal.hpp"
#include "kernels.hpp"
#include "ntt_lds.hpp"

namespace fheacc {

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed + stream * 0x9E3779B97F4A7C15ULL);
}
__device__ __forceinline__ uint64_t stream_word(uint64_t key, uint64_t ctr) {
    return mix64(key + (ctr + 1) * 0x9E3779B97F4A7C15ULL);
}
__device__ __forceinline__ int64_t gauss_dev(const uint64_t *cdt, int len, uint64_t w) {
    uint64_t v = w >> 1;
    int k = 0;
    while (k < len - 1 && v >= cdt[k]) k++;
    if (!k) return 0;
    return (w & 1) ? -(int64_t)k : (int64_t)k;
}

// floor((2 v q + Q) / (2 Q)), LWE.cpp:255-260 (not reduced mod q)
__device__ __forceinline__ uint64_t scale_round_dev(uint64_t v, int lq1, uint64_t Q) {
    uint64_t lo = v << lq1, hi = v >> (64 - lq1);
    uint64_t lo2 = lo + Q;
    hi += lo2 < lo;
    lo = lo2;
    const uint64_t twoQ = 2 * Q;
    double approx = ((double)v * (double)(1ULL << lq1) + (double)Q) / (double)twoQ;
    uint64_t r = approx > 0 ? (uint64_t)approx : 0;
    // exact fix-up: r * 2Q <= num < (r + 1) * 2Q
    for (int it = 0; it < 2; it++) {
        uint64_t ph = __umul64hi(r, twoQ), pl = r * twoQ;
        if (ph > hi || (ph == hi && pl > lo)) { r--; continue; }
        uint64_t r1 = r + 1;
        uint64_t qh = __umul64hi(r1, twoQ), ql = r1 * twoQ;
        if (qh < hi || (qh == hi && ql <= lo)) r = r1;
        else break;
    }
    return r;
}

// ShoupGen (kernels.hpp): k * 2^64 / Q for k < Q, from a = floor(2^64/Q),
// b = 2^64 mod Q and b's Shoup companion; the wide key layout stores it
// beside each key word.

template <typename W>
__device__ __forceinline__ W mulmod_generic(const Arith<W> &A, W x, W y, W yp);
template <>
__device__ __forceinline__ uint32_t mulmod_generic(const Arith<uint32_t> &A, uint32_t x, uint32_t y, uint32_t) {
    return A.reduce64((uint64_t)x * y);
}
template <>
__device__ __forceinline__ uint64_t mulmod_generic(const Arith<uint64_t> &A, uint64_t x, uint64_t y, uint64_t yp) {
    return Arith<uint64_t>::red(A.mul_shoup(x, y, yp), A.Q);
}

// x X^e in the NTT domain at output index idx, e in [0, 2N): NTT(X^e)[idx] =
// psi^((2 brev(idx) + 1) e) (bit-reversed output order, nbt.cpp:423-454), and
// psi^t = rou[brev(t)] for t < N, -psi^(t - N) above (psi^N = -1).
template <typename W, int LOGN>
__device__ __forceinline__ W mul_monomial(const Arith<W> &A, const Twiddles<W> &tw, W x, int idx, int e) {
    constexpr int N = 1 << LOGN;
    const int br = (int)(__brev((unsigned)idx) >> (32 - LOGN));
    const int t = (int)(((int64_t)(2 * br + 1) * e) & (2 * N - 1));
    const int j = (int)(__brev((unsigned)(t & (N - 1))) >> (32 - LOGN));
    const W v = Arith<W>::red(A.mul_shoup(x, tw.fw[j], tw.fwp[j]), A.Q);
    return t >= N ? (v ? A.Q - v : (W)0) : v;
}

// ---------------------------------------------------------------------------
// external product (RLWE.cpp:623-704) on an accumulator held in LDS
// ---------------------------------------------------------------------------

template <typename W>
struct MacSum;
// Digit words arrive lazily reduced (the forward transform's non-canonical
// bound): the 32-bit sum of 2 digitG such products stays below 2^64, and the
// 64-bit sum of products in [0, 5Q) below 2^62 (both checked when the context
// derives its parameters); Shoup products accept any 64-bit input.
template <>
struct MacSum<uint32_t> {
    static constexpr bool kCanonicalInputs = false;
    uint64_t v = 0;
    __device__ __forceinline__ void add(const Arith<uint32_t> &, uint32_t d, uint32_t k, uint32_t) {
        v += (uint64_t)d * k;
    }
    __device__ __forceinline__ uint32_t get(const Arith<uint32_t> &A) const { return A.reduce64(v); }
};
template <>
struct MacSum<uint64_t> {
    static constexpr bool kCanonicalInputs = false;
    uint64_t v = 0;
    __device__ __forceinline__ void add(const Arith<uint64_t> &A, uint64_t d, uint64_t k, uint64_t kp) {
        v += A.mul_shoup_lazy(d, k, kp);
    }
    __device__ __forceinline__ uint64_t get(const Arith<uint64_t> &A) const { return A.reduce_small(v); }
};

// Key words of row h at coefficient k: (K0a, K0b, K1a, K1b) and companions.
template <typename W>
struct KeyQuad {
    W v[4], p[4];
};
// Key pointers may come from a table in memory (per-client keys), where the
// compiler cannot see that they are global and would emit flat loads, which
// also count against the LDS wait counter; these go through global loads.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ T ld_global(const void *p, size_t i) {
    typedef const __attribute__((address_space(1))) T global_t;
    return ((global_t *)p)[i];
}

// A pointer every lane holds the same value of, moved to scalar registers
// (the compiler cannot see that a pointer read from memory is uniform).
template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T *)(((uint64_t)hi << 32) | lo);
}
// The key base is uniform across the workgroup and the quad index fits 32
// bits, so the load takes a scalar base and a 32-bit vector byte offset.
template <typename T>
__device__ __forceinline__ T ld_global_off(const void *p, uint32_t byte_off) {
    typedef const __attribute__((address_space(1))) T global_t;
    return *(global_t *)((const __attribute__((address_space(1))) char *)p + byte_off);
}
__device__ __forceinline__ KeyQuad<uint32_t> load_quad(const uint32_t *key, int idx) {
    const u32x4 q = ld_global_off<u32x4>(key, (uint32_t)idx * 16u);
    KeyQuad<uint32_t> r;
    r.v[0] = q.x; r.v[1] = q.y; r.v[2] = q.z; r.v[3] = q.w;
    r.p[0] = r.p[1] = r.p[2] = r.p[3] = 0;
    return r;
}
__device__ __forceinline__ KeyQuad<uint64_t> load_quad(const uint64_t *key, int idx) {
    KeyQuad<uint64_t> r;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const u64x2 t = ld_global_off<u64x2>(key, (uint32_t)idx * 64u + 16u * i);
        r.v[i] = t.x;
        r.p[i] = t.y;
    }
    return r;
}
// W words of one key quad (with companions when wide): the row h of an RGSW
// starts at key + (h << LOGN) * kQuadWords<W>
template <typename W>
constexpr int kQuadWords = sizeof(W) == 4 ? 4 : 8;

// Key quads through a buffer resource over one RGSW: the row's byte offset
// goes in the scalar offset and the thread's in a loop-invariant vector
// offset, so a load costs no VALU address arithmetic.  The range check is
// off in effect (num_records 2^32 - 1; every offset stays inside the RGSW).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t key_rsrc(const void *key) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(key), 0, 0x7fffffff, 0x00020000);
}
__device__ __forceinline__ KeyQuad<uint32_t> load_quad_rs(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff,
                                                          const uint32_t *) {
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    const i32x4 q = __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, 0);
    KeyQuad<uint32_t> r;
    r.v[0] = (uint32_t)q.x; r.v[1] = (uint32_t)q.y; r.v[2] = (uint32_t)q.z; r.v[3] = (uint32_t)q.w;
    r.p[0] = r.p[1] = r.p[2] = r.p[3] = 0;
    return r;
}
__device__ __forceinline__ KeyQuad<uint64_t> load_quad_rs(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff,
                                                          const uint64_t *) {
    typedef int i32x4 __attribute__((ext_vector_type(4)));
    KeyQuad<uint64_t> r;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const i32x4 t = __builtin_amdgcn_raw_buffer_load_b128(rs, voff + 16u * i, soff, 0);
        r.v[i] = (uint64_t)(uint32_t)t.x | ((uint64_t)(uint32_t)t.y << 32);
        r.p[i] = (uint64_t)(uint32_t)t.z | ((uint64_t)(uint32_t)t.w << 32);
    }
    return r;
}

// acc (2N words in LDS, NTT domain, canonical) <- acc (x) rgsw.
// The 2d digit polys are taken in the order (h, a), (h, b), h = 0..d-1, and
// transformed pb at a time in dig (pb polys of LDS scratch).
template <typename W, int LOGN, int NT>
__device__ void ext_prod(W *acc, W *dig, const W *key, const Arith<W> &A, const Twiddles<W> &tw, int d,
                         int bits, int pb) {
    constexpr int N = 1 << LOGN;
    constexpr int CPT = (N + NT - 1) / NT;
    // An opaque copy of the thread index: the caller runs this in a long
    // loop, and without it the compiler hoists every per-thread address of
    // every pass out of that loop and runs out of registers.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    // 64-bit words at 1024 threads: 3-stage passes and one set of key quads
    // keep within 128 registers
    constexpr bool kLean = sizeof(W) == 8 && NT >= 1024;
    constexpr int MAXR = kLean ? 3 : 4;
    ntt_inverse<LOGN, NT, MAXR, 2>(acc, 2, A, tw, tid);
    MacSum<W> sa[CPT], sb[CPT];
    const W mask = bits >= (int)(8 * sizeof(W)) ? ~(W)0 : (((W)1 << bits) - 1);
    const int total = 2 * d;
    for (int P0 = 0; P0 < total; P0 += pb) {
        const int np = min(pb, total - P0);
        for (int idx = tid; idx < (np << LOGN); idx += NT) {
            const int P = P0 + (idx >> LOGN), k = idx & (N - 1);
            const W x = acc[pad(((P & 1) << LOGN) + k)];
            dig[pad(idx)] = (x >> ((P >> 1) * bits)) & mask;
        }
        __syncthreads();
        // digit transforms stay lazy: the MAC sums reduce them
        ntt_forward<LOGN, NT, MAXR>(dig, np, A, tw, MacSum<W>::kCanonicalInputs, tid);
        if (!(P0 & 1) && !(np & 1) && N >= NT) {
            // whole (a, b) digit pairs: key quads of the next pair are loaded
            // while the current pair is accumulated (two sets in flight)
            // (pairs alternate between two register sets; each coefficient's
            // digits are immediate offsets from a pointer that advances two
            // padded polys per pair, as in ext_prod_inplace)
            constexpr int PS = N + N / 8;
            const int npairs = np >> 1, h0 = P0 >> 1;
            KeyQuad<W> cur[CPT], nxt[CPT];
            const W *dcol[CPT];
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                dcol[c] = dig + pad(tid + c * NT);
                cur[c] = load_quad(key, (h0 << LOGN) + tid + c * NT);
            }
            auto pair = [&](const KeyQuad<W> *q) {
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    const W da = dcol[c][0], db = dcol[c][PS];
                    sa[c].add(A, da, q[c].v[0], q[c].p[0]);
                    sb[c].add(A, da, q[c].v[1], q[c].p[1]);
                    sa[c].add(A, db, q[c].v[2], q[c].p[2]);
                    sb[c].add(A, db, q[c].v[3], q[c].p[3]);
                    dcol[c] += 2 * PS;
                }
            };
            auto fetch = [&](KeyQuad<W> *q, int hp) {
#pragma unroll
                for (int c = 0; c < CPT; c++)
                    q[c] = load_quad(key + ((size_t)(h0 + hp) << LOGN) * kQuadWords<W>, tid + c * NT);
            };
            if constexpr (kLean) {
#pragma unroll 1
                for (int hp = 0; hp < npairs; hp++) {
                    if (hp) fetch(cur, hp);
                    pair(cur);
                }
            } else {
#pragma unroll 1
                for (int hp = 0; hp < npairs; hp += 2) {
                    if (hp + 1 < npairs) fetch(nxt, hp + 1);
                    pair(cur);
                    if (hp + 1 < npairs) {
                        if (hp + 2 < npairs) fetch(cur, hp + 2);
                        pair(nxt);
                    }
                }
            }
            __syncthreads();
            continue;
        }
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int k = tid + c * NT;
            if (k < N) {
                for (int pp = 0; pp < np;) {
                    const int P = P0 + pp;
                    KeyQuad<W> kq = load_quad(key, ((P >> 1) << LOGN) + k);
                    if (!(P & 1)) {
                        const W da = dig[pad((pp << LOGN) + k)];
                        sa[c].add(A, da, kq.v[0], kq.p[0]);
                        sb[c].add(A, da, kq.v[1], kq.p[1]);
                        if (pp + 1 < np) {
                            const W db = dig[pad(((pp + 1) << LOGN) + k)];
                            sa[c].add(A, db, kq.v[2], kq.p[2]);
                            sb[c].add(A, db, kq.v[3], kq.p[3]);
                            pp += 2;
                        } else {
                            pp += 1;
                        }
                    } else {
                        const W db = dig[pad((pp << LOGN) + k)];
                        sa[c].add(A, db, kq.v[2], kq.p[2]);
                        sb[c].add(A, db, kq.v[3], kq.p[3]);
                        pp += 1;
                    }
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int k = tid + c * NT;
        if (k < N) {
            acc[pad(k)] = sa[c].get(A);
            acc[pad(N + k)] = sb[c].get(A);
        }
    }
    __syncthreads();
}

// The same product with every digit poly in one LDS batch, computed in place:
// buf holds 2d padded polys, the accumulator (a, b) in the first two.  Each
// thread owns coefficients k = tid + c NT: it reads its a[k], b[k] into
// registers before the digits overwrite them, and after the transforms it
// reads all 2d digits of k before writing the new a[k], b[k] -- slots no
// other thread touches -- so the accumulator needs no LDS of its own.  (At
// the shape whose multiply-accumulate takes the forward's last stages, the
// thread's coefficients there are CPT tid + c instead: still columns only
// that thread touches between the barriers.)
// D > 0: the digit count d is the compile-time D (the transforms, the digit
// split and the MAC unroll over the polys); D = 0: any d.
template <typename W, int LOGN, int NT, int D = 0>
__device__ void ext_prod_inplace(W *buf, const W *key, const Arith<W> &A, const Twiddles<W> &tw, int d_rt,
                                 int bits) {
    const int d = D > 0 ? D : d_rt;
    constexpr int N = 1 << LOGN;
    constexpr int CPT = (N + NT - 1) / NT;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));  // see ext_prod
    // with N a multiple of NT every k = tid + c NT is a coefficient; the
    // opaque tid hides that from the compiler, so say it here
    constexpr bool kFull = N % NT == 0;
    // 64-bit words at 1024 threads: 3-stage passes (a cluster and its
    // twiddles) and one set of key quads keep within 128 VGPRs
    constexpr bool kLean = sizeof(W) == 8 && NT >= 1024;
    constexpr int MAXR = kLean ? 3 : 4;
    // With N = 2^F NT a thread's words tid + c NT (stride NT = N / 2^F) are
    // the cluster of the forward transform's first F stages and of the
    // inverse's last F: those stages run in registers, between the digit
    // split and the LDS, when that saves an LDS pass of each transform
    constexpr int F = !kFull ? 0 : CPT >= 8 ? 3 : CPT >= 4 ? 2 : CPT >= 2 ? 1 : 0;
    constexpr bool kFuse = F > 0 && (1 << F) == CPT && pass_count<MAXR>(LOGN - F) < pass_count<MAXR>(LOGN);
    // The forward side too, except at N = 1024 on 32-bit words at 256
    // threads: there the unfused schedule 4,3,3 runs as fast as [2],4,4 and
    // its LDS sees fewer bank conflicts (16 % of cycles against 19 %), and
    // the forward's last stages run in the multiply-accumulate instead
    // (kMacTail); at N = 512 the fused [1],4,4 is 2.7 % faster than 4,3,2.
    constexpr bool kFuseFwd = kFuse && !(sizeof(W) == 4 && LOGN == 10 && NT == 256);
    // The inverse runs on 2 polys only: with 16-word clusters half the
    // threads would idle at every pass.  On 32-bit words, where 8-word
    // clusters give every thread one (2N / 8 = NT) and the last stages run
    // in registers, 3-stage passes are faster despite the extra pass
    // (N = 1024 at 256 threads: -1.6 %; at N = 512, where they would leave
    // half the threads idle too, +2.5 %).
    constexpr int IMAXR = (sizeof(W) == 4 && MAXR > 3 && 2 * N / 8 == NT && kFuse) ? 3 : MAXR;
    ntt_inverse<LOGN, NT, IMAXR, 2, kFuse ? LOGN - F : LOGN>(buf, 2, A, tw, tid);
    const W mask = bits >= (int)(8 * sizeof(W)) ? ~(W)0 : (((W)1 << bits) - 1);
    W xa[CPT], xb[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int k = tid + c * NT;
        if (kFull || k < N) {
            FHEACC_CHECK_LDS(&buf[pad(N + k)]);
            xa[c] = buf[pad(k)];
            xb[c] = buf[pad(N + k)];
        }
    }
    // digit h of a goes to poly 2h, of b to poly 2h + 1: a poly is a multiple
    // of 8 words, so pad(2hN + k) = pad(k) + 2h padded-poly strides, and the
    // unrolled writes become immediate LDS offsets.  A thread writes only the
    // columns it has just read, so no barrier separates the two.
    constexpr int PS = N + N / 8;  // padded words per poly
    constexpr int DU = 8;          // digit count unrolled (larger d loops)
    if constexpr (kFuse) {
        const ClusterTwiddles<F, W> ti = inv_twiddles<LOGN, LOGN - F, F>(0, tw);
        Words<F, W> ya, yb;
#pragma unroll
        for (int c = 0; c < CPT; c++) ya.x[c] = xa[c], yb.x[c] = xb[c];
        ya = inv_tail_regs<F>(ya, A, tw, ti);
        yb = inv_tail_regs<F>(yb, A, tw, ti);
#pragma unroll
        for (int c = 0; c < CPT; c++) xa[c] = ya.x[c], xb[c] = yb.x[c];
    }
    if constexpr (kFuseFwd) {
        const ClusterTwiddles<F, W> tf = fwd_twiddles<0, F>(0, tw);
        W *col = buf + pad(tid);
        // digit h of x through the first F stages, into poly p
        auto put = [&](auto hv, const W (&x)[CPT], int h, int p) {
            Words<F, W> y;
#pragma unroll
            for (int c = 0; c < CPT; c++) y.x[c] = (x[c] >> (h * bits)) & mask;
            y = fwd_regs<F, decltype(hv)::value, false>(y, false, A, tw, tf);
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                FHEACC_CHECK_LDS(&col[p * PS + pad(c * NT)]);
                col[p * PS + pad(c * NT)] = y.x[c];
            }
        };
        auto split = [&](auto hv) {
            if constexpr (D > 0) {
#pragma unroll
                for (int h = 0; h < D; h++) {
                    put(hv, xa, h, 2 * h);
                    put(hv, xb, h, 2 * h + 1);
                }
            } else {
#pragma unroll 1
                for (int h = 0; h < d; h++) {
                    put(hv, xa, h, 2 * h);
                    put(hv, xb, h, 2 * h + 1);
                }
            }
        };
        if constexpr (sizeof(W) == 4) {
            if (A.harvey) split(std::true_type{});
            else split(std::false_type{});
        } else {
            split(std::false_type{});
        }
    } else {
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int k = tid + c * NT;
            if (kFull || k < N) {
                W *col = buf + pad(k);
                if (d <= DU) {
#pragma unroll
                    for (int h = 0; h < DU; h++) {
                        if (h < d) {
                            FHEACC_CHECK_LDS(&col[(2 * h + 1) * PS]);
                            col[2 * h * PS] = (xa[c] >> (h * bits)) & mask;
                            col[(2 * h + 1) * PS] = (xb[c] >> (h * bits)) & mask;
                        }
                    }
                } else {
                    for (int h = 0; h < d; h++) {
                        FHEACC_CHECK_LDS(&col[(2 * h + 1) * PS]);
                        col[2 * h * PS] = (xa[c] >> (h * bits)) & mask;
                        col[(2 * h + 1) * PS] = (xb[c] >> (h * bits)) & mask;
                    }
                }
            }
        }
    }
    constexpr uint32_t QB = kQuadWords<W> * sizeof(W);  // bytes per key quad
    const __amdgpu_buffer_rsrc_t krs = key_rsrc(uniform_ptr(key));
    KeyQuad<W> cur[CPT], nxt[CPT];
    // With one register set of key quads (kLean) each row's loads would be
    // waited for in full; row 0's go out before the forward transform, which
    // hides them, where a compile-time digit count leaves the registers for
    // it (the 54-bit d = 3 gate: -2.1 %; d = 6 external product: -1.7 %)
    constexpr bool kEarlyRow0 = kLean && D > 0;
    if constexpr (kEarlyRow0) {
#pragma unroll
        for (int c = 0; c < CPT; c++)
            if (kFull || tid + c * NT < N) cur[c] = load_quad_rs(krs, (uint32_t)(tid + c * NT) * QB, 0u, key);
    }
    __syncthreads();
    // At N = 1024 on 32-bit words at 256 threads the forward transform's last
    // two stages run in the multiply-accumulate instead, on groups of 4
    // consecutive coefficients: a thread reads words 4 tid .. 4 tid + 3 of
    // each digit poly after stage 7, finishes stages 8 and 9 in registers
    // and accumulates those 4 coefficients.  Their key quads are adjacent, 64
    // bytes per lane.  One LDS round trip of the 2d polys fewer (-2.2 %).
    constexpr bool kMacTail = sizeof(W) == 4 && LOGN == 10 && NT == 256 && kFull && CPT == 4;
    if constexpr (kMacTail) {
        constexpr int T = 2;  // forward stages in the multiply-accumulate
        ntt_forward<LOGN, NT, MAXR, 2 * D, kFuseFwd ? F : 0, LOGN - T>(buf, 2 * d, A, tw, false, tid);
        const ClusterTwiddles<T, W> tt = fwd_twiddles<LOGN - T, T>(tid, tw);
        MacSum<W> sa[CPT], sb[CPT];
        W *g = buf + pad(CPT * tid);  // 4 consecutive words: one 8-word padding block
        auto mac_row = [&](auto hv, int h) {
            KeyQuad<W> q[CPT];
#pragma unroll
            for (int c = 0; c < CPT; c++)
                q[c] = load_quad_rs(krs, (uint32_t)(CPT * tid + c) * QB, ((uint32_t)h << LOGN) * QB, key);
            Words<T, W> ya, yb;
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                FHEACC_CHECK_LDS(&g[(2 * h + 1) * PS + c]);
                ya.x[c] = g[2 * h * PS + c], yb.x[c] = g[(2 * h + 1) * PS + c];
            }
            ya = fwd_regs<T, decltype(hv)::value, true>(ya, false, A, tw, tt);
            yb = fwd_regs<T, decltype(hv)::value, true>(yb, false, A, tw, tt);
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                sa[c].add(A, ya.x[c], q[c].v[0], q[c].p[0]);
                sb[c].add(A, ya.x[c], q[c].v[1], q[c].p[1]);
                sa[c].add(A, yb.x[c], q[c].v[2], q[c].p[2]);
                sb[c].add(A, yb.x[c], q[c].v[3], q[c].p[3]);
            }
        };
        auto mac = [&](auto hv) {
            if constexpr (D > 0) {
#pragma unroll
                for (int h = 0; h < D; h++) mac_row(hv, h);
            } else {
#pragma unroll 1
                for (int h = 0; h < d; h++) mac_row(hv, h);
            }
        };
        if (A.harvey) mac(std::true_type{});
        else mac(std::false_type{});
        // every digit word of these 4 columns has been read by this thread
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            g[c] = sa[c].get(A);
            g[PS + c] = sb[c].get(A);
        }
        __syncthreads();
        return;
    }
    ntt_forward<LOGN, NT, MAXR, 2 * D, kFuseFwd ? F : 0>(buf, 2 * d, A, tw, MacSum<W>::kCanonicalInputs, tid);
    // gadget rows outer, the thread's coefficients inner: row h+1's key quads
    // are in flight while row h is accumulated.  Rows alternate between two
    // register sets (no copies; one set when kLean), and each coefficient's
    // digit words are read through a pointer that moves two padded polys per
    // row, so the a and b digits are immediate offsets from it.
    MacSum<W> sa[CPT], sb[CPT];
    const W *dcol[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        dcol[c] = buf + pad(tid + c * NT);
        if (!kEarlyRow0 && (kFull || tid + c * NT < N))
            cur[c] = load_quad_rs(krs, (uint32_t)(tid + c * NT) * QB, 0u, key);
    }
    auto row = [&](const KeyQuad<W> *q) {
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            if (kFull || tid + c * NT < N) {
                FHEACC_CHECK_LDS(&dcol[c][PS]);
                const W da = dcol[c][0], db = dcol[c][PS];
                sa[c].add(A, da, q[c].v[0], q[c].p[0]);
                sb[c].add(A, da, q[c].v[1], q[c].p[1]);
                sa[c].add(A, db, q[c].v[2], q[c].p[2]);
                sb[c].add(A, db, q[c].v[3], q[c].p[3]);
            }
            dcol[c] += 2 * PS;
        }
    };
    auto fetch = [&](KeyQuad<W> *q, int h) {
#pragma unroll
        for (int c = 0; c < CPT; c++)
            if (kFull || tid + c * NT < N)
                q[c] = load_quad_rs(krs, (uint32_t)(tid + c * NT) * QB, ((uint32_t)h << LOGN) * QB, key);
    };
    if constexpr (kLean) {
        // one register set of key quads: two would not fit in 128 VGPRs, and
        // at 128 (the d = 3 kernel) the gate ran 12 % slower
#pragma unroll 1
        for (int h = 0; h < d; h++) {
            if (h) fetch(cur, h);
            row(cur);
        }
    } else if constexpr (D > 0) {
#pragma unroll
        for (int h = 0; h < D; h += 2) {
            if (h + 1 < D) fetch(nxt, h + 1);
            row(cur);
            if (h + 1 < D) {
                if (h + 2 < D) fetch(cur, h + 2);
                row(nxt);
            }
        }
    } else {
#pragma unroll 1
        for (int h = 0; h < d; h += 2) {
            if (h + 1 < d) fetch(nxt, h + 1);
            row(cur);
            if (h + 1 < d) {
                if (h + 2 < d) fetch(cur, h + 2);
                row(nxt);
            }
        }
    }
    // every digit of column k has been read by its owner: overwrite in place
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int k = tid + c * NT;
        if (kFull || k < N) {
            buf[pad(k)] = sa[c].get(A);
            buf[pad(N + k)] = sb[c].get(A);
        }
    }
    __syncthreads();
}

// Two key words of row h at coefficient k: (K0a, K0b) for the digits of a
// (half 0) or (K1a, K1b) for those of b (half 1), with companions when wide.
template <typename W>
struct KeyPair {
    W v[2], p[2];
};
__device__ __forceinline__ KeyPair<uint32_t> load_pair(const uint32_t *key, int idx, int half) {
    typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
    const u32x2 q = ld_global_off<u32x2>(key, (uint32_t)idx * 16u + 8u * half);
    KeyPair<uint32_t> r;
    r.v[0] = q.x; r.v[1] = q.y;
    r.p[0] = r.p[1] = 0;
    return r;
}
__device__ __forceinline__ KeyPair<uint64_t> load_pair(const uint64_t *key, int idx, int half) {
    KeyPair<uint64_t> r;
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const u64x2 t = ld_global_off<u64x2>(key, (uint32_t)idx * 64u + 32u * half + 16u * i);
        r.v[i] = t.x;
        r.p[i] = t.y;
    }
    return r;
}

// The in-place product in two batches of d digit polys, the digits of a and
// then those of b, so that buf holds max(2, d) padded polys instead of 2d.
// Each thread keeps its coefficients' a[k], b[k] in registers from the
// inverse transform on, and sums both halves' products for its own k
// before the new accumulator is written.  Half the LDS of ext_prod_inplace
// for the same transforms and products: a 54-bit N=2048 d=3 gate takes 55 KB
// instead of 110 KB, so two workgroups share a CU.
template <typename W, int LOGN, int NT>
__device__ void ext_prod_split(W *buf, const W *key, const Arith<W> &A, const Twiddles<W> &tw, int d, int bits) {
    constexpr int N = 1 << LOGN;
    constexpr int CPT = (N + NT - 1) / NT;
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));  // see ext_prod
    constexpr bool kFull = N % NT == 0;
    // 64-bit words: 3-stage inverse passes; the forward passes, which run
    // beside b's held words and the sums, take 2 stages to fit 128 registers
    // (3 spill at 512 threads)
    constexpr int MAXR = sizeof(W) == 8 ? 3 : 4;
    constexpr int MAXRF = sizeof(W) == 8 ? 2 : 4;
    constexpr int PS = N + N / 8;  // padded words per poly
    ntt_inverse<LOGN, NT, MAXR, 2>(buf, 2, A, tw, tid);
    const W mask = bits >= (int)(8 * sizeof(W)) ? ~(W)0 : (((W)1 << bits) - 1);
    W xa[CPT], xb[CPT];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int k = tid + c * NT;
        if (kFull || k < N) {
            FHEACC_CHECK_LDS(&buf[pad(N + k)]);
            xa[c] = buf[pad(k)];
            xb[c] = buf[pad(N + k)];
        }
    }
    __syncthreads();
    MacSum<W> sa[CPT], sb[CPT];
    // one half: its digit polys, their transforms and products (called twice,
    // so that a's words are dead once their digits are written)
    auto half_product = [&](const W *xs, const int half) {
        // digit h of a (half 0) or of b (half 1) goes to poly h
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            const int k = tid + c * NT;
            if (kFull || k < N) {
                const W x = xs[c];
                W *col = buf + pad(k);
                if (d <= 8) {
#pragma unroll
                    for (int h = 0; h < 8; h++)
                        if (h < d) col[h * PS] = (x >> (h * bits)) & mask;
                } else {
                    for (int h = 0; h < d; h++) col[h * PS] = (x >> (h * bits)) & mask;
                }
            }
        }
        __syncthreads();
        ntt_forward<LOGN, NT, MAXRF>(buf, d, A, tw, MacSum<W>::kCanonicalInputs, tid);
        // rows outer, the thread's coefficients inner; row h+1's key words are
        // in flight while row h is summed
        KeyPair<W> cur[CPT], nxt[CPT];
        const W *dcol[CPT];
#pragma unroll
        for (int c = 0; c < CPT; c++) {
            dcol[c] = buf + pad(tid + c * NT);
            if constexpr (sizeof(W) == 4)
                if (kFull || tid + c * NT < N) cur[c] = load_pair(key, tid + c * NT, half);
        }
        auto row = [&](const KeyPair<W> *q) {
#pragma unroll
            for (int c = 0; c < CPT; c++) {
                if (kFull || tid + c * NT < N) {
                    const W dg = dcol[c][0];
                    sa[c].add(A, dg, q[c].v[0], q[c].p[0]);
                    sb[c].add(A, dg, q[c].v[1], q[c].p[1]);
                }
                dcol[c] += PS;
            }
        };
        auto fetch = [&](KeyPair<W> *q, int h) {
#pragma unroll
            for (int c = 0; c < CPT; c++)
                if (kFull || tid + c * NT < N)
                    q[c] = load_pair(key + ((size_t)h << LOGN) * kQuadWords<W>, tid + c * NT, half);
        };
        if constexpr (sizeof(W) == 8) {
            // 64-bit words: each coefficient's key words are loaded just
            // before its products (the 16 waves of two workgroups per CU
            // cover the loads), so the products fit 128 registers
#pragma unroll 1
            for (int h = 0; h < d; h++) {
#pragma unroll
                for (int c = 0; c < CPT; c++) {
                    if (kFull || tid + c * NT < N) {
                        const KeyPair<W> q = load_pair(key + ((size_t)h << LOGN) * kQuadWords<W>, tid + c * NT, half);
                        const W dg = dcol[c][0];
                        sa[c].add(A, dg, q.v[0], q.p[0]);
                        sb[c].add(A, dg, q.v[1], q.p[1]);
                    }
                    dcol[c] += PS;
                }
            }
        } else {
#pragma unroll 1
            for (int h = 0; h < d; h += 2) {
                if (h + 1 < d) fetch(nxt, h + 1);
                row(cur);
                if (h + 1 < d) {
                    if (h + 2 < d) fetch(cur, h + 2);
                    row(nxt);
                }
            }
        }
        // every digit poly has been read: the next half (or the new
        // accumulator) overwrites it
        __syncthreads();
    };
    half_product(xa, 0);
    half_product(xb, 1);
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const int k = tid + c * NT;
        if (kFull || k < N) {
            buf[pad(k)] = sa[c].get(A);
            buf[pad(N + k)] = sb[c].get(A);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// K10 + K5 + blind rotation + KS preparation, one ciphertext per workgroup
// (RLWE.cpp:726-767, LWE.cpp:203-222,233-240, RLWE.cpp:218-226)
// LDS: acc (2 padded polys), dig (pb padded polys), the n evaluated a words.
// ---------------------------------------------------------------------------

// Lock-step pacing of a tile (one launch of resident workgroups): prog[w]
// holds how many steps workgroup w has finished.  Workgroups are grouped by
// w mod groups (one group per XCD when workgroups are dealt round robin); a
// workgroup starts step t once every member of its group has finished step
// t - window, or after a bounded spin.  Thread i reads member i's word, so a
// check is one parallel load and one barrier; the next step's words are read
// while this step computes.  The words only pace the work (every
// ciphertext's words depend on its own inputs alone), so a late or missing
// workgroup costs time, never a wrong word or a hang.
struct StepPacer {
    uint32_t *prog = nullptr;  // the tile's progress words (null: off)
    int first = 0;             // this workgroup's group: members first + groups i, i < members
    int groups = 1, members = 0, window = 0, spin_limit = 0;
    uint32_t next = 0;         // member threadIdx.x's word as read before the previous step
    bool have_next = false;
    __device__ __forceinline__ uint32_t read() const {
        return (int)threadIdx.x < members
                   ? __hip_atomic_load(prog + first + groups * (int)threadIdx.x, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT)
                   : ~0u;
    }
    // checkpoints: every max(1, window / 2) steps (a check costs a device-wide
    // read and a barrier); between them a workgroup runs free
    __device__ __forceinline__ bool checkpoint(uint32_t t) const { return t % (uint32_t)max(1, window >> 1) == 0; }
    // before step t (uniform across the workgroup)
    __device__ __forceinline__ void wait(uint32_t t) {
        if (!prog || !checkpoint(t)) return;
        const int64_t need = (int64_t)t - window + 1;  // steps every member must have finished
        if (need > 0) {
            uint32_t v = have_next ? next : read();
            int all = __syncthreads_and(v >= (uint32_t)need);
            for (int spins = 0; !all && spins < spin_limit; spins++) {
                __builtin_amdgcn_s_sleep(2);
                v = read();
                all = __syncthreads_and(v >= (uint32_t)need);
            }
        }
        next = read();
        have_next = true;
    }
    // after step t: every thread has finished it (the product ends on a barrier)
    __device__ __forceinline__ void done(uint32_t t) {
        if (prog && threadIdx.x == 0 && checkpoint(t))
            __hip_atomic_store(prog + blockIdx.x, (uint32_t)(t + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// One gate bootstrap's blind rotation and key-switch preparation on the
// workgroup's LDS: ciphertext c, thread index tid (an opaque copy in the
// persistent kernel), steps counted from step0 for pacing.
template <typename W, int LOGN, int NT, int FORM, bool PACE, int D = 0>
__device__ __forceinline__ void br_one(const BRArgs<W> &a, int64_t c, int tid, StepPacer &pace, uint32_t step0) {
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem[];
    // kInPlace: every digit poly in one batch, product computed in place in
    // 2d polys; kSplit: in place in d polys, digits of a then of b;
    // kSeparate: acc (2 polys) and a separate pb-poly digit buffer (the
    // launcher picks the form from pb)
    W *acc = reinterpret_cast<W *>(smem);
    W *dig = acc + padded(2 * N);
    uint16_t *av = reinterpret_cast<uint16_t *>(
        FORM == kInPlace ? acc + padded(max(2, 2 * a.d) * N)
                         : (FORM == kSplit ? acc + padded(max(2, a.d) * N) : dig + padded(a.pb * N)));
    __shared__ uint32_t s_b;
    const int n = a.n;
    const uint32_t q = a.q, qm = a.q - 1;
    FHEACC_CHECK(c >= a.c_base && c < a.c_end, kCheckCiphertext);
    FHEACC_CHECK_LDS(&av[n - 1]);
    // gates outside enum GATES [0, 5] are an error of the call: flagged like
    // identical operands, and the ciphertext skips its products
    int g = a.gates[a.gidx ? a.gidx[c] : c];
    const bool bad_gate = (unsigned)g > 5u;
    if (bad_gate) g = 0;
    const bool xr = (g == 4 || g == 5);
    const uint64_t *x1 = a.c1 + (a.idx1 ? a.idx1[c] : c) * a.row_words;
    const uint64_t *x2 = a.c2 + (a.idx2 ? a.idx2[c] : c) * a.row_words;
    int same = 1;
    for (int k = tid; k <= n; k += NT) {
        uint64_t u = x1[k], v = x2[k];
        same &= (u == v);
        uint32_t r = (uint32_t)(xr ? 2 * (u - v) : u + v) & qm;
        if (k < n) av[k] = (uint16_t)r;
        else s_b = r;
    }
    const int all_same = __syncthreads_and(same);
    if (tid == 0) {
        // bit 0: identical operands (RLWE.cpp:727-730); bit 1: gate out of range
        const int flags = (all_same ? FHEACC_STATUS_SAME : 0) | (bad_gate ? FHEACC_STATUS_GATE : 0);
        if (a.ct_status) a.ct_status[c] = flags;
        else if (flags) atomicOr(a.status, flags);
    }
    // accumulator init: only b carries the test polynomial
    const uint64_t Q8 = a.A.Q / 8 + 1, nQ8 = a.A.Q - Q8;
    const uint32_t b = s_b;
    const uint32_t b1 = (uint32_t)(a.gate_constant[g] & qm), b2 = (b1 + q / 2) & qm;
    const int ef = (2 * N) / (int)q;
    for (int idx = tid; idx < padded(2 * N); idx += NT) acc[idx] = 0;
    __syncthreads();
    for (int i = tid; i < (int)(q / 2); i += NT) {
        const uint32_t tmp = (b - (uint32_t)i) & qm;
        bool neg = (b1 < b2) ? (tmp >= b1 && tmp < b2) : !(tmp >= b2 && tmp < b1);
        acc[pad(N + i * ef)] = (W)(neg ? nQ8 : Q8);
    }
    __syncthreads();
    ntt_forward<LOGN, NT, 4, 1>(acc + pad(N), 1, a.A, a.tw, true, tid);
    // blind rotation: i outer, j inner.  Software order (RLWE.cpp:754-767)
    // skips residue 0; the accelerator's one-ciphertext form multiplies at
    // step (0, 0) whatever the residue, its batched forms at every step
    // (myfpga.cpp:1030-1071, 1180-1225), with btkey[0] = RGSW(1).  A skipped
    // step (and every step of a flagged ciphertext) still counts for pacing.
    const W *bsk = uniform_ptr(a.key_of ? a.bsk_tab[a.key_of[c]] : a.bsk);
    const int mode = a.mode;
    for (int i = 0, t = 0; i < n; i++) {
        // uniform across the workgroup: the key address of each step is
        // scalar, and its loads take a scalar base
        uint32_t nega = __builtin_amdgcn_readfirstlane((q - av[i]) & qm);
        for (int j = 0; j < a.digitR; j++, t++) {
            const uint32_t r = nega % a.Br;
            nega /= a.Br;
            if constexpr (PACE) pace.wait(step0 + t);
            if (!bad_gate && (r || mode == FHEACC_EVAL_FPGA_XN || (mode == FHEACC_EVAL_FPGA_X1 && i == 0 && j == 0))) {
                const W *key = bsk + ((size_t)(i * a.digitR + j) * a.Br + r) * a.rgsw_elems;
                if constexpr (FORM == kInPlace) ext_prod_inplace<W, LOGN, NT, D>(acc, key, a.A, a.tw, a.d, a.bits);
                else if constexpr (FORM == kSplit) ext_prod_split<W, LOGN, NT>(acc, key, a.A, a.tw, a.d, a.bits);
                else ext_prod<W, LOGN, NT>(acc, dig, key, a.A, a.tw, a.d, a.bits, a.pb);
            }
            if constexpr (PACE) pace.done(step0 + t);
        }
    }
    if (a.cap_acc) {
        uint64_t *dst = a.cap_acc + c * 2 * N;
        for (int idx = tid; idx < 2 * N; idx += NT) dst[idx] = acc[pad(idx)];
    }
    __syncthreads();
    // LWE_keyswitch preparation: NTT-domain reversal of a, inverse NTT,
    // base-Bks digits of every a coefficient, b0 + Q/8 + 1
    for (int idx = tid; idx < N / 2; idx += NT) {
        W t = acc[pad(idx)];
        acc[pad(idx)] = acc[pad(N - 1 - idx)];
        acc[pad(N - 1 - idx)] = t;
    }
    __syncthreads();
    ntt_inverse<LOGN, NT, 4, 2>(acc, 2, a.A, a.tw, tid);
    const int dks = a.dks;
    const uint32_t Bks = a.Bks;
    uint8_t *dst = a.ksd + c * (size_t)N * dks;
    for (int i = tid; i < N; i += NT) {
        uint64_t x = acc[pad(i)];
        for (int j = 0; j < dks; j++) {
            dst[i * dks + j] = (uint8_t)(x % Bks);
            x /= Bks;
        }
    }
    if (tid == 0) {
        uint64_t v = (uint64_t)acc[pad(N)] + Q8;
        a.ksb[c] = v >= a.A.Q ? v - a.A.Q : v;
    }
}

// One ciphertext per workgroup; the hardware dispatcher refills CUs as gates
// finish.
// Waves per SIMD the blind rotation and the CMUX chains are compiled for.  The
// split product on 64-bit words at 512 threads runs two workgroups per CU
// (55 KB of LDS each at N = 2048), 16 waves, so it must fit 128 registers; so
// must the in-place product on 32-bit words at 256 threads and N = 1024 (four
// workgroups of 8 polys per CU, 37 KB of LDS each).  Every other shape keeps
// the register budget its workgroup size gives.
template <typename W, int LOGN, int NT, int FORM>
constexpr int product_waves_per_eu() {
    return (FORM == kSplit && sizeof(W) == 8 && NT == 512) ||
                   (FORM == kInPlace && sizeof(W) == 4 && NT == 256 && LOGN == 10)
               ? 4
               : 1;
}

// D > 0: a kernel for digit count D alone (launch_br_nt picks it when a.d == D)
template <typename W, int LOGN, int NT, int FORM, int D = 0>
__global__ __launch_bounds__(NT, (product_waves_per_eu<W, LOGN, NT, FORM>())) void k_blind_rotate(BRArgs<W> a) {
    StepPacer none;
    br_one<W, LOGN, NT, FORM, false, D>(a, a.c_base + blockIdx.x, threadIdx.x, none, 0);
}

// The digit counts that get a kernel of their own at a launch shape: the
// shapes the autotuner picks for the benchmark set (27-bit Q, N = 1024,
// d = 4: 256 threads for batches, 1024 for a single gate), the mixed batch's
// 54-bit set (N = 2048, d = 3) and its N = 512 set (d = 5), and the
// reference's 256-bit application set (N = 2048, 29-bit Q, d = 3, 256
// threads).  Every other (shape, d) runs the general kernel.
template <typename W, int LOGN, int NT>
constexpr int specialized_digits() {
    if constexpr (sizeof(W) == 4 && LOGN == 10 && (NT == 256 || NT == 1024)) return 4;
    if constexpr (sizeof(W) == 8 && LOGN == 11 && NT == 1024) return 3;
    if constexpr (sizeof(W) == 4 && LOGN == 9 && NT == 256) return 5;
    if constexpr (sizeof(W) == 4 && LOGN == 11 && NT == 256) return 3;
    return 0;
}

// Persistent and paced: a resident grid of G workgroups, workgroup w taking
// ciphertexts c_base + w, + w + G, ...; its steps are counted across its
// ciphertexts, so a group stays within a window of steps from gate to gate
// without waiting at gate boundaries.  A workgroup out of work marks itself
// finished so that no one waits for it.  D > 0: the kernel for digit count D
// alone, as for k_blind_rotate.
template <typename W, int LOGN, int NT, int FORM, int D = 0>
__global__ __launch_bounds__(NT, (product_waves_per_eu<W, LOGN, NT, FORM>())) void k_blind_rotate_paced(BRArgs<W> a) {
    StepPacer pace;
    if (a.sync) {
        pace.prog = a.sync;
        pace.groups = a.sync_groups;
        pace.first = (int)(blockIdx.x % (unsigned)a.sync_groups);
        pace.members = ((int)gridDim.x - pace.first + a.sync_groups - 1) / a.sync_groups;
        pace.window = a.sync_window;
        pace.spin_limit = a.spin_limit;
    }
    const uint32_t steps = (uint32_t)a.n * a.digitR;
    uint32_t step0 = 0;
    for (int64_t c = a.c_base + blockIdx.x; c < a.c_end; c += gridDim.x, step0 += steps) {
        // an opaque thread index per gate, so that per-thread addresses are
        // recomputed for each gate instead of held across the whole loop
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        br_one<W, LOGN, NT, FORM, true, D>(a, c, tid, pace, step0);
        __syncthreads();  // the next gate rewrites the LDS
    }
    if (a.sync && threadIdx.x == 0)
        __hip_atomic_store(a.sync + blockIdx.x, ~0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// LWE key switch + modulus switch (LWE.cpp:241-269)
// ---------------------------------------------------------------------------

template <typename W>
__device__ __forceinline__ uint64_t ks_finish(const KSArgs<W> &a, int64_t c, int col, uint64_t s) {
    const uint64_t Q = a.Q;
    uint64_t v;
    if (col < a.n) v = s ? Q - s : 0;
    else {
        uint64_t b = a.ksb[c];
        v = b >= s ? b - s : b + Q - s;
    }
    if (a.cap_ks) a.cap_ks[c * (a.n + 1) + col] = v;
    return scale_round_dev(v, a.lq1, Q);
}

template <typename W>
__device__ __forceinline__ void ks_add(uint64_t &s, W v, uint64_t Q);
template <>
__device__ __forceinline__ void ks_add(uint64_t &s, uint32_t v, uint64_t) { s += v; }
template <>
__device__ __forceinline__ void ks_add(uint64_t &s, uint64_t v, uint64_t Q) {
    uint64_t t = s + v;
    s = t >= Q ? t - Q : t;
}
template <typename W>
__device__ __forceinline__ uint64_t ks_reduce(uint64_t s, const Arith<W> &A);
template <>
__device__ __forceinline__ uint64_t ks_reduce(uint64_t s, const Arith<uint32_t> &A) { return A.reduce64(s); }
template <>
__device__ __forceinline__ uint64_t ks_reduce(uint64_t s, const Arith<uint64_t> &) { return s; }

// Staged form: a workgroup owns KB output columns of NT ciphertexts, one
// ciphertext per thread.  The Bks candidate rows (KB words each) of 16 steps
// are staged in LDS per fill; each thread reads its own 16 digits with one
// 16-byte load and adds the rows they select.  The key is streamed once per
// NT ciphertexts.  Rows are padded to RS words to spread LDS banks.  When a
// fill is at most KS_PV vectors per thread (Bks <= 32) the fills are double
// buffered: fill k+1's key vectors and digits are in flight in registers
// while fill k is summed, and one barrier per fill suffices.
constexpr int KS_SB = 16;  // steps per LDS fill
constexpr int KS_PV = 8;   // staged 16-byte vectors per thread when double buffered

template <typename W, int KB>
__host__ __device__ constexpr bool ks_double_buffered(uint32_t Bks, int nt) {
    return KS_SB * (int)Bks * (KB / (16 / (int)sizeof(W))) <= KS_PV * nt;
}

template <typename W, int NT, int KB>
__global__ __launch_bounds__(NT) void k_lwe_keyswitch_staged(KSArgs<W> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr int SB = KS_SB;
    constexpr int RS = KB + 16 / (int)sizeof(W);       // padded row stride (words)
    constexpr int V = 16 / (int)sizeof(W);             // words per 16-byte vector
    const uint32_t Bks = a.Bks;
    W *kt0 = reinterpret_cast<W *>(smem);              // [SB][Bks][RS] (twice when double buffered)
    const int col0 = blockIdx.x * KB;
    const int64_t c = (int64_t)blockIdx.y * NT + threadIdx.x;
    const bool live = c < a.B;
    const int S = a.S;
    const uint8_t *dg = a.ksd + (live ? c : 0) * (size_t)S;
    uint64_t s[KB];
#pragma unroll
    for (int j = 0; j < KB; j++) s[j] = 0;
    const int vec_per_fill = SB * Bks * (KB / V);
    // staged vector e of the fill that starts at step s0 (zero past the key)
    auto fetch = [&](int s0, int e) {
        const int row = e / (KB / V), vv = e - row * (KB / V);  // row = st * Bks + dd
        const int st = row / Bks;
        const int gc = col0 + vv * V;
        u32x4 val = {0, 0, 0, 0};
        if (s0 + st < S && gc + V <= a.n_pad)
            val = ld_global<u32x4>(a.ksk + ((size_t)(s0 * Bks) + row) * a.n_pad + gc, 0);
        return val;
    };
    auto put = [&](W *kt, int e, u32x4 val) {
        const int row = e / (KB / V), vv = e - row * (KB / V);
        *reinterpret_cast<u32x4 *>(kt + row * RS + vv * V) = val;
    };
    // this thread's 16 digits of the fill at s0, packed, zero past S
    auto fetch_digits = [&](int s0) {
        u32x4 dv = {0, 0, 0, 0};
        if (s0 + SB <= S) {
            dv = ld_global<u32x4>(dg + s0, 0);
        } else {
            uint32_t w[4] = {0, 0, 0, 0};
            for (int t = 0; s0 + t < S; t++) w[t >> 2] |= (uint32_t)dg[s0 + t] << (8 * (t & 3));
            dv.x = w[0]; dv.y = w[1]; dv.z = w[2]; dv.w = w[3];
        }
        return dv;
    };
    // 32-bit keys with 16 Q < 2^32: a fill's 16 rows sum in 32 bits and are
    // added to the 64-bit sums once per fill
    const bool fold = sizeof(W) == 4 && (uint64_t)SB * a.Q < (1ULL << 32);
    auto accumulate = [&](const W *kt, int s0, u32x4 dv) {
        const uint32_t w4[4] = {dv.x, dv.y, dv.z, dv.w};
        const int nsb = min(SB, S - s0);
        if constexpr (sizeof(W) == 4) {
            if (fold) {
                uint32_t f[KB];
#pragma unroll
                for (int j = 0; j < KB; j++) f[j] = 0;
#pragma unroll
                for (int st = 0; st < SB; st++) {
                    if (st < nsb) {
                        const int d = (int)((w4[st >> 2] >> (8 * (st & 3))) & 0xFF);
                        const W *row = kt + (st * (int)Bks + d) * RS;
                        FHEACC_CHECK_LDS(row + KB - 1);
#pragma unroll
                        for (int vv = 0; vv < KB / V; vv++) {
                            uint4 q = *reinterpret_cast<const uint4 *>(row + vv * V);
                            f[vv * 4 + 0] += q.x;
                            f[vv * 4 + 1] += q.y;
                            f[vv * 4 + 2] += q.z;
                            f[vv * 4 + 3] += q.w;
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < KB; j++) s[j] += f[j];
                return;
            }
        }
#pragma unroll
        for (int st = 0; st < SB; st++) {
            if (st < nsb) {
                const int d = (int)((w4[st >> 2] >> (8 * (st & 3))) & 0xFF);
                const W *row = kt + (st * (int)Bks + d) * RS;
                FHEACC_CHECK_LDS(row + KB - 1);
#pragma unroll
                for (int vv = 0; vv < KB / V; vv++) {
                    uint4 q = *reinterpret_cast<const uint4 *>(row + vv * V);
                    if (sizeof(W) == 4) {
                        ks_add<W>(s[vv * 4 + 0], (W)q.x, a.Q);
                        ks_add<W>(s[vv * 4 + 1], (W)q.y, a.Q);
                        ks_add<W>(s[vv * 4 + 2], (W)q.z, a.Q);
                        ks_add<W>(s[vv * 4 + 3], (W)q.w, a.Q);
                    } else {
                        ks_add<W>(s[vv * 2 + 0], (W)(((uint64_t)q.y << 32) | q.x), a.Q);
                        ks_add<W>(s[vv * 2 + 1], (W)(((uint64_t)q.w << 32) | q.z), a.Q);
                    }
                }
            }
        }
    };
    if (ks_double_buffered<W, KB>(Bks, NT)) {
        W *kt1 = kt0 + SB * (int)Bks * RS;
#pragma unroll
        for (int i = 0; i < KS_PV; i++) {
            const int e = threadIdx.x + i * NT;
            if (e < vec_per_fill) put(kt0, e, fetch(0, e));
        }
        u32x4 dcur = fetch_digits(0);
        for (int s0 = 0, k = 0; s0 < S; s0 += SB, k++) {
            W *cur = (k & 1) ? kt1 : kt0, *nxt = (k & 1) ? kt0 : kt1;
            const bool more = s0 + SB < S;
            u32x4 pre[KS_PV];
            u32x4 dnext = {0, 0, 0, 0};
            if (more) {
#pragma unroll
                for (int i = 0; i < KS_PV; i++) {
                    const int e = threadIdx.x + i * NT;
                    if (e < vec_per_fill) pre[i] = fetch(s0 + SB, e);
                }
                dnext = fetch_digits(s0 + SB);
            }
            // cur was written before this barrier (by the previous pass), and
            // every read of nxt in the previous pass happened before it
            __syncthreads();
            accumulate(cur, s0, dcur);
            if (more) {
#pragma unroll
                for (int i = 0; i < KS_PV; i++) {
                    const int e = threadIdx.x + i * NT;
                    if (e < vec_per_fill) put(nxt, e, pre[i]);
                }
            }
            dcur = dnext;
        }
    } else {
        for (int s0 = 0; s0 < S; s0 += SB) {
            for (int e = threadIdx.x; e < vec_per_fill; e += NT) put(kt0, e, fetch(s0, e));
            const u32x4 dv = fetch_digits(s0);
            __syncthreads();
            accumulate(kt0, s0, dv);
            __syncthreads();
        }
    }
    if (!live) return;
    const int64_t orow = a.oidx ? a.oidx[c] : c;
#pragma unroll
    for (int j = 0; j < KB; j++) {
        const int col = col0 + j;
        if (col <= a.n) a.out[orow * a.out_row_words + col] = ks_finish(a, c, col, ks_reduce<W>(s[j], a.A));
    }
}

// Gather form for small batches: each ciphertext reads only its N*digitKS
// rows.  The steps are split into a.split ranges, a workgroup each, so that a
// handful of ciphertexts still spreads over the CUs instead of queueing
// thousands of dependent loads in two workgroups; k_lwe_keyswitch_finish adds
// the partial sums.  The order of a sum mod Q does not change its value.
template <typename W, int NT>
__global__ __launch_bounds__(NT) void k_lwe_keyswitch_gather(KSArgs<W> a) {
    const int cb = (a.n + 1 + NT - 1) / NT;  // column blocks; blockIdx.x = part * cb + column block
    const int part = blockIdx.x / cb;
    const int col = (blockIdx.x - part * cb) * NT + threadIdx.x;
    const int64_t c = blockIdx.y;
    const int st0 = (int)((int64_t)a.S * part / a.split), st1 = (int)((int64_t)a.S * (part + 1) / a.split);
    const uint8_t *dg = a.ksd + c * (size_t)a.S;
    uint64_t s = 0;
    if (col <= a.n) {
        for (int st = st0; st < st1; st++) ks_add<W>(s, a.ksk[((size_t)st * a.Bks + dg[st]) * a.n_pad + col], a.Q);
        if (a.split == 1)
            a.out[(a.oidx ? a.oidx[c] : c) * a.out_row_words + col] = ks_finish(a, c, col, ks_reduce<W>(s, a.A));
        else
            a.part[(c * a.split + part) * (a.n + 1) + col] = s;
    }
}

// Sum of the gather form's partial sums: 32-bit keys leave plain 64-bit sums
// (below S Q in all), 64-bit keys canonical ones.
template <typename W, int NT>
__global__ __launch_bounds__(NT) void k_lwe_keyswitch_finish(KSArgs<W> a) {
    const int col = blockIdx.x * NT + threadIdx.x;
    const int64_t c = blockIdx.y;
    if (col > a.n) return;
    uint64_t s = 0;
    for (int part = 0; part < a.split; part++) {
        const uint64_t v = a.part[(c * a.split + part) * (a.n + 1) + col];
        if constexpr (sizeof(W) == 4) {
            s += v;
        } else {
            const uint64_t t = s + v;
            s = t >= a.Q ? t - a.Q : t;
        }
    }
    a.out[(a.oidx ? a.oidx[c] : c) * a.out_row_words + col] = ks_finish(a, c, col, ks_reduce<W>(s, a.A));
}

// ---------------------------------------------------------------------------
// batched NTT / external product / subs + RLWE key switch
// ---------------------------------------------------------------------------

// NP > 0: every workgroup transforms NP polys (a compile-time count): the
// words move as 16-byte vectors, each thread's loads all in flight at once,
// and the passes unroll over the polys.  NP = 0: per_wg polys, the last
// workgroup fewer.
template <typename W, int LOGN, int NT, int NP = 0>
__global__ __launch_bounds__(NT) void k_ntt(NTTArgs<W> a) {
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem[];
    W *s = reinterpret_cast<W *>(smem);
    const int64_t p0 = (int64_t)blockIdx.x * (NP > 0 ? NP : a.per_wg);
    const int np = NP > 0 ? NP : (int)min((int64_t)a.per_wg, a.count - p0);
    uint64_t *g = a.polys + p0 * N;
    // 3-stage passes for 64-bit words here (fewer registers, more resident
    // waves; measured 15 % faster at N=1024); the fused kernels keep 4
    constexpr int MAXR = sizeof(W) == 8 ? 3 : 4;
    if constexpr (NP > 0) {
        constexpr int VEC = NP * N / 2;  // 16-byte vectors of two words
        constexpr int IT = (VEC + NT - 1) / NT;
        u64x2 *gv = reinterpret_cast<u64x2 *>(g);
        u64x2 v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int e = threadIdx.x + it * NT;
            if (VEC % NT == 0 || e < VEC) v[it] = gv[e];
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int e = threadIdx.x + it * NT;
            if (VEC % NT == 0 || e < VEC) {
                FHEACC_CHECK_LDS(&s[pad(2 * e) + 1]);
                s[pad(2 * e)] = (W)v[it].x;  // words 2e, 2e + 1 share an 8-word padding block
                s[pad(2 * e) + 1] = (W)v[it].y;
            }
        }
        __syncthreads();
        if (a.inverse) ntt_inverse<LOGN, NT, MAXR, NP>(s, NP, a.A, a.tw);
        else ntt_forward<LOGN, NT, MAXR, NP>(s, NP, a.A, a.tw);
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int e = threadIdx.x + it * NT;
            if (VEC % NT == 0 || e < VEC) {
                u64x2 o;
                o.x = s[pad(2 * e)];
                o.y = s[pad(2 * e) + 1];
                gv[e] = o;
            }
        }
        return;
    }
    for (int i = threadIdx.x; i < np * N; i += NT) s[pad(i)] = (W)g[i];
    __syncthreads();
    if (a.inverse) ntt_inverse<LOGN, NT, MAXR>(s, np, a.A, a.tw);
    else ntt_forward<LOGN, NT, MAXR>(s, np, a.A, a.tw);
    for (int i = threadIdx.x; i < np * N; i += NT) g[i] = s[pad(i)];
}

// INPLACE (pb >= 2d): the product runs in the 2d digit polys, the RLWE in
// the first two of them, as in the blind rotation; otherwise acc (2 polys)
// plus a pb-poly digit buffer.
template <typename W, int LOGN, int NT, bool INPLACE, int D = 0>
__global__ __launch_bounds__(NT) void k_external_product(EPArgs<W> a) {
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem[];
    W *acc = reinterpret_cast<W *>(smem);
    W *dig = acc + padded(2 * N);
    const int64_t c = a.c_base + blockIdx.x;
    const uint64_t *src = a.acc + c * a.in_stride;
    for (int i = threadIdx.x; i < 2 * N; i += NT) acc[pad(i)] = (W)src[i];
    __syncthreads();
    if constexpr (INPLACE) ext_prod_inplace<W, LOGN, NT, D>(acc, a.key + c * a.key_stride, a.A, a.tw, a.d, a.bits);
    else ext_prod<W, LOGN, NT>(acc, dig, a.key + c * a.key_stride, a.A, a.tw, a.d, a.bits, a.pb);
    uint64_t *dst = a.out + c * a.out_stride;
    for (int i = threadIdx.x; i < 2 * N; i += NT) dst[i] = acc[pad(i)];
}

// RLWE_substitute (RLWE.cpp:444-462, poly_substitute 23-37) followed by
// RLWE_keyswitch (RLWE.cpp:464-510).  The keyswitch's own iNTT of the
// substituted ciphertext is the identity on the time-domain words already
// in LDS, so the digits are taken from them directly.
//
// EXPAND: one level of expand_RLWE (app.cpp:26-97) over [B][N] slots of
// RLWEs: workgroup (b, j), j < 2^i, substitutes and key-switches slot j with
// factor i, then writes slot j = c + subs and slot j + 2^i =
// X^(-2^i) (c - subs) (RLWE_rotate_freq backward, RLWE.cpp:411-430; the
// rotation polynomial (Q - 1) X^(N - 2^i) is X^(2N - 2^i)).
template <typename W, int LOGN, int NT, bool EXPAND>
__global__ __launch_bounds__(NT) void k_subs_keyswitch(SubsArgs<W> a) {
    constexpr int N = 1 << LOGN;
    constexpr int CPT = (N + NT - 1) / NT;
    // at 1024 threads 3-stage NTT passes keep within 128 registers
    constexpr int MAXR = NT >= 1024 ? 3 : 4;
    extern __shared__ __align__(16) unsigned char smem[];
    W *x = reinterpret_cast<W *>(smem);  // input, time domain (2 padded polys)
    W *y = x + padded(2 * N);            // substituted (2 padded polys)
    W *dig = y + padded(2 * N);          // dch padded digit polys
    const int64_t c = a.c_base + blockIdx.x;
    // EXPAND: workgroup c takes slot c mod 2^i of expansion c / 2^i
    const int64_t off = EXPAND ? (c >> a.sub_factor) * a.set_stride + (c & ((1 << a.sub_factor) - 1)) * a.slot_stride
                               : c * 2 * N;
    const uint64_t *src = a.in + off;
    for (int i = threadIdx.x; i < 2 * N; i += NT) x[pad(i)] = (W)src[i];
    __syncthreads();
    ntt_inverse<LOGN, NT, MAXR>(x, 2, a.A, a.tw);
    const int power = (N >> a.sub_factor) + 1;
    for (int idx = threadIdx.x; idx < 2 * N; idx += NT) {
        const int half = idx >> LOGN, i = idx & (N - 1);
        const int64_t ip = (int64_t)i * power;
        const int io = (int)(ip & (N - 1));
        const W v = x[pad(idx)];
        FHEACC_CHECK_LDS(&y[pad((half << LOGN) + io)]);
        y[pad((half << LOGN) + io)] = ((ip >> LOGN) & 1) ? (v ? (W)a.A.Q - v : (W)0) : v;
    }
    __syncthreads();
    // output b starts from NTT(subs(b)); the digits come from subs(a)
    for (int i = threadIdx.x; i < N; i += NT) x[pad(N + i)] = y[pad(N + i)];
    __syncthreads();
    ntt_forward<LOGN, NT, MAXR>(x + pad(N), 1, a.A, a.tw);
    MacSum<W> sa[CPT], sb[CPT];
    const int d = a.d, bits = a.bits, dch = a.dch;
    const W mask = bits >= (int)(8 * sizeof(W)) ? ~(W)0 : (((W)1 << bits) - 1);
    const W *key = a.key + (size_t)a.sub_factor * a.key_stride;
    for (int h0 = 0; h0 < d; h0 += dch) {
        const int nd = min(dch, d - h0);
        for (int k = threadIdx.x; k < N; k += NT) {
            const W v = y[pad(k)];
            for (int hh = 0; hh < nd; hh++) {
                FHEACC_CHECK_LDS(&dig[pad((hh << LOGN) + k)]);
                dig[pad((hh << LOGN) + k)] = (v >> ((h0 + hh) * bits)) & mask;
            }
        }
        __syncthreads();
        ntt_forward<LOGN, NT, MAXR>(dig, nd, a.A, a.tw, MacSum<W>::kCanonicalInputs);
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) {
            const int k = threadIdx.x + cc * NT;
            if (k < N) {
                for (int hh = 0; hh < nd; hh++) {
                    const W dv = dig[pad((hh << LOGN) + k)];
                    const size_t e = ((size_t)(h0 + hh) << LOGN) + k;
                    if (sizeof(W) == 4) {
                        uint2 kv = reinterpret_cast<const uint2 *>(key)[e];
                        sa[cc].add(a.A, dv, (W)kv.x, 0);
                        sb[cc].add(a.A, dv, (W)kv.y, 0);
                    } else {
                        const ulonglong2 *kp = reinterpret_cast<const ulonglong2 *>(key) + 2 * e;
                        ulonglong2 ka = kp[0], kb = kp[1];
                        sa[cc].add(a.A, dv, (W)ka.x, (W)ka.y);
                        sb[cc].add(a.A, dv, (W)kb.x, (W)kb.y);
                    }
                }
            }
        }
        __syncthreads();
    }
    const uint64_t Q = a.A.Q;
    if constexpr (!EXPAND) {
        uint64_t *dst = a.out + c * 2 * N;
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) {
            const int k = threadIdx.x + cc * NT;
            if (k < N) {
                const uint64_t ma = sa[cc].get(a.A), mb = sb[cc].get(a.A);
                dst[k] = ma ? Q - ma : 0;
                const uint64_t bv = x[pad(N + k)];
                dst[N + k] = bv >= mb ? bv - mb : bv + Q - mb;
            }
        }
    } else {
        // slot j is re-read from memory (this workgroup alone writes it, and
        // only after every thread has read its own words)
        uint64_t *lo = a.out + off, *hi = lo + (a.slot_stride << a.sub_factor);
        const int e = 2 * N - (1 << a.sub_factor);
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) {
            const int k = threadIdx.x + cc * NT;
            if (k < N) {
                const uint64_t ma = sa[cc].get(a.A), mb = sb[cc].get(a.A);
                const uint64_t bv = x[pad(N + k)];
                const uint64_t s2[2] = {ma ? Q - ma : 0, bv >= mb ? bv - mb : bv + Q - mb};
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const uint64_t v = src[h * N + k], sv = s2[h];
                    const uint64_t sum = v + sv;
                    lo[h * N + k] = sum >= Q ? sum - Q : sum;
                    const uint64_t df = v >= sv ? v - sv : v + Q - sv;
                    hi[h * N + k] = mul_monomial<W, LOGN>(a.A, a.tw, (W)df, k, e);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// CMUX chains: CMUX (RLWE.cpp:788-793), blind_rotate (RLWE.cpp:822-838) and
// blind_rotate_full (RLWE.cpp:795-820), one ciphertext per workgroup.  The
// running accumulator of the thread's coefficients stays in registers; the
// difference X(acc) - acc goes through the external product in LDS.
// ---------------------------------------------------------------------------

template <typename W, int LOGN, int NT, int FORM>
__global__ __launch_bounds__(NT, (product_waves_per_eu<W, LOGN, NT, FORM>())) void k_cmux_chain(CmuxArgs<W> a) {
    constexpr int N = 1 << LOGN;
    constexpr int CPT = (N + NT - 1) / NT;
    constexpr bool kFull = N % NT == 0;
    extern __shared__ __align__(16) unsigned char smem[];
    W *acc = reinterpret_cast<W *>(smem);
    W *dig = acc + padded(2 * N);
    const int64_t c = a.c_base + blockIdx.x;
    const uint64_t *src = a.c0 + c * a.in_stride;
    const uint64_t Q = a.A.Q;
    W sav[2][CPT];
#pragma unroll
    for (int cc = 0; cc < CPT; cc++) {
        const int k = threadIdx.x + cc * NT;
        if (kFull || k < N) {
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const W v = (W)src[h * N + k];
                sav[h][cc] = a.pre_rot ? mul_monomial<W, LOGN>(a.A, a.tw, v, k, a.pre_rot) : v;
            }
        }
    }
    for (int st = 0; st < a.steps; st++) {
        const int e = a.c1 ? 0 : a.rot[st];
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) {
            const int k = threadIdx.x + cc * NT;
            if (kFull || k < N) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const W v = sav[h][cc];
                    const W o = a.c1 ? (W)a.c1[c * a.in_stride + h * N + k] : mul_monomial<W, LOGN>(a.A, a.tw, v, k, e);
                    acc[pad(h * N + k)] = o >= v ? o - v : o + (W)Q - v;
                }
            }
        }
        __syncthreads();
        const W *key = a.key + c * a.key_ct_stride + st * a.key_stride;
        if constexpr (FORM == kInPlace) ext_prod_inplace<W, LOGN, NT>(acc, key, a.A, a.tw, a.d, a.bits);
        else if constexpr (FORM == kSplit) ext_prod_split<W, LOGN, NT>(acc, key, a.A, a.tw, a.d, a.bits);
        else ext_prod<W, LOGN, NT>(acc, dig, key, a.A, a.tw, a.d, a.bits, a.pb);
#pragma unroll
        for (int cc = 0; cc < CPT; cc++) {
            const int k = threadIdx.x + cc * NT;
            if (kFull || k < N) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const W s = acc[pad(h * N + k)] + sav[h][cc];
                    sav[h][cc] = s >= (W)Q ? s - (W)Q : s;
                }
            }
        }
        // the next step rewrites acc: every thread has read its words
        __syncthreads();
    }
    uint64_t *dst = a.out + c * 2 * N;
#pragma unroll
    for (int cc = 0; cc < CPT; cc++) {
        const int k = threadIdx.x + cc * NT;
        if (kFull || k < N) {
            dst[k] = sav[0][cc];
            dst[N + k] = sav[1][cc];
        }
    }
}

// ---------------------------------------------------------------------------
// key layout conversion and seeded generation
// ---------------------------------------------------------------------------

// Offsets, inside one RGSW of u64 words, of the four key words of gadget row h
// and coefficient k (K0a, K0b, K1a, K1b).  Reference order [h][g][a, b][N]
// (RLWE.h:313-329); or the accelerator's DDR order: RLWEs c_text[h][g] g-major
// (myfpga.cpp:624-646), each as 4 words of a then 4 of b, N/4 times
// (convert_RLWE_to_buffer, myfpga.h:201-223).
__device__ __forceinline__ void rgsw_word_offsets(int h, int k, int d, int N, bool ddr, size_t o[4]) {
    if (!ddr) {
        const size_t b = (size_t)h * 4 * N + k;
        o[0] = b;
        o[1] = b + N;
        o[2] = b + 2 * (size_t)N;
        o[3] = b + 3 * (size_t)N;
    } else {
        const size_t in = (size_t)(k >> 2) * 8 + (k & 3);
        o[0] = (size_t)h * 2 * N + in;
        o[1] = o[0] + 4;
        o[2] = (size_t)(d + h) * 2 * N + in;
        o[3] = o[2] + 4;
    }
}

// reference (or DDR-image) RGSW u64 words -> device quad layout
template <typename W>
__global__ void k_rgsw_to_device(const uint64_t *src, int64_t src_stride, W *dst, int64_t dst_stride,
                                 int d, int N, ShoupGen sg, const int64_t *map, int64_t count, bool ddr) {
    const int64_t e = blockIdx.y;
    if (e >= count) return;
    const int64_t se = map ? map[e] : e;
    const uint64_t *s = src + se * src_stride;
    W *o = dst + e * dst_stride;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < d * N; idx += gridDim.x * blockDim.x) {
        const int h = idx / N, k = idx - h * N;
        size_t off[4];
        rgsw_word_offsets(h, k, d, N, ddr, off);
        uint64_t v0 = s[off[0]], v1 = s[off[1]], v2 = s[off[2]], v3 = s[off[3]];
        if (sizeof(W) == 4) {
            reinterpret_cast<uint4 *>(o)[idx] = make_uint4((uint32_t)v0, (uint32_t)v1, (uint32_t)v2, (uint32_t)v3);
        } else {
            ulonglong2 *p = reinterpret_cast<ulonglong2 *>(o) + 4 * (size_t)idx;
            p[0] = make_ulonglong2(v0, sg(v0));
            p[1] = make_ulonglong2(v1, sg(v1));
            p[2] = make_ulonglong2(v2, sg(v2));
            p[3] = make_ulonglong2(v3, sg(v3));
        }
    }
}

// device quad layout -> reference RGSW layout
template <typename W>
__global__ void k_rgsw_to_ref(const W *src, int64_t src_stride, uint64_t *dst, int64_t dst_stride, int d, int N,
                              const int64_t *map, int64_t count, bool ddr) {
    const int64_t e = blockIdx.y;
    if (e >= count) return;
    const W *s = src + e * src_stride;
    uint64_t *o = dst + (map ? map[e] : e) * dst_stride;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < d * N; idx += gridDim.x * blockDim.x) {
        const int h = idx / N, k = idx - h * N;
        size_t off[4];
        rgsw_word_offsets(h, k, d, N, ddr, off);
        if (sizeof(W) == 4) {
            uint4 q = reinterpret_cast<const uint4 *>(s)[idx];
            o[off[0]] = q.x; o[off[1]] = q.y; o[off[2]] = q.z; o[off[3]] = q.w;
        } else {
            const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(s) + 4 * (size_t)idx;
            o[off[0]] = p[0].x; o[off[1]] = p[1].x; o[off[2]] = p[2].x; o[off[3]] = p[3].x;
        }
    }
}

// reference KSK [Bks][dks][N][n+1] -> device [N][dks][Bks][n_pad] (and back)
template <typename W>
__global__ void k_ksk_convert(const uint64_t *src, W *dst, uint64_t *back, int Bks, int dks, int N, int n1,
                              int n_pad) {
    const int64_t row = blockIdx.x;  // device row index ((i * dks + j) * Bks + dd)
    const int dd = row % Bks;
    const int64_t ij = row / Bks;
    const int j = ij % dks, i = ij / dks;
    const int64_t rrow = ((int64_t)dd * dks + j) * N + i;
    for (int k = threadIdx.x; k < n_pad; k += blockDim.x) {
        if (back) {
            if (k < n1) back[rrow * n1 + k] = dst[row * n_pad + k];
        } else {
            dst[row * n_pad + k] = k < n1 ? (W)src[rrow * n1 + k] : (W)0;
        }
    }
}

// Forward NTT of the RLWE secret (and Shoup companions) for key generation.
template <typename W, int LOGN, int NT>
__global__ __launch_bounds__(NT) void k_secret_ntt(const uint64_t *z, W *zn, W *znp, Arith<W> A, Twiddles<W> tw,
                                                   ShoupGen sg) {
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem[];
    W *s = reinterpret_cast<W *>(smem);
    for (int i = threadIdx.x; i < N; i += NT) s[pad(i)] = (W)z[i];
    __syncthreads();
    ntt_forward<LOGN, NT>(s, 1, A, tw);
    for (int i = threadIdx.x; i < N; i += NT) {
        const W v = s[pad(i)];
        zn[i] = v;
        znp[i] = sizeof(W) == 4 ? (W)(((uint64_t)v << 32) / A.Q) : (W)sg(v);
    }
}

// RLWE encryption of zero into LDS (RLWE.cpp:324-344): a uniform and e
// Gaussian in the time domain, both transformed, b = a*s + e.  Word stream as
// the host sampler.  buf: 2 padded polys; leaves a in the first, b in the second.
template <typename W, int LOGN, int NT>
__device__ void rlwe_enc_zero(W *buf, const W *zn, const W *znp, uint64_t seed, uint64_t stream,
                              const KeygenArgs<W> &a) {
    constexpr int N = 1 << LOGN;
    const uint64_t key = stream_key(seed, stream);
    const uint64_t Q = a.A.Q;
    for (int k = threadIdx.x; k < N; k += NT) {
        buf[pad(k)] = (W)__umul64hi(stream_word(key, k), Q);
        int64_t e = gauss_dev(a.cdt, a.cdt_len, stream_word(key, N + k)) % (int64_t)a.E_rlwe;
        buf[pad(N + k)] = (W)(e < 0 ? (int64_t)Q + e : e);
    }
    __syncthreads();
    ntt_forward<LOGN, NT>(buf, 2, a.A, a.tw);
    for (int k = threadIdx.x; k < N; k += NT) {
        W p = mulmod_generic<W>(a.A, buf[pad(k)], zn[k], znp[k]);
        uint64_t t = (uint64_t)p + buf[pad(N + k)];
        buf[pad(N + k)] = (W)(t >= Q ? t - Q : t);
    }
    __syncthreads();
}

// Bootstrap key: one workgroup per (RGSW, row h); both columns g = 0, 1.
// RLWE.cpp:512-526 (X^mm, sign), 561-577 (rows), 579-606 (btkey[r][j][k]).
template <typename W, int LOGN, int NT>
__global__ __launch_bounds__(NT) void k_bsk_gen(KeygenArgs<W> a) {
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem[];
    W *e0 = reinterpret_cast<W *>(smem);  // one column's RLWE: a, b
    W *m = e0 + padded(2 * N);            // message monomial
    const int d = a.d;
    const int64_t wid = blockIdx.x;
    const int h = wid % d;
    const int64_t dev_rgsw = wid / d;  // (k * digitR + j) * Br + r
    const int rr = dev_rgsw % a.Br;
    const int64_t kj = dev_rgsw / a.Br;
    const int j = kj % a.digitR;
    const int k = kj / a.digitR;
    uint32_t powbr = 1;
    for (int t = 0; t < j; t++) powbr = (uint32_t)(((uint64_t)powbr * a.Br) % a.q);
    const uint32_t msg = (uint32_t)((((uint64_t)rr * powbr) % a.q * a.lwe_sk[k]) % a.q);
    uint32_t mm = msg * (uint32_t)(2 * N / a.q);
    bool neg = false;
    if (mm >= (uint32_t)N) { mm -= N; neg = true; }
    const uint64_t Q = a.A.Q;
    uint64_t coef = a.bg_pow[h];  // BG^h mod Q
    if (neg) coef = coef ? Q - coef : 0;
    for (int i = threadIdx.x; i < N; i += NT) m[pad(i)] = (W)(i == (int)mm ? coef : 0);
    const int64_t ref_e = ((int64_t)rr * a.digitR + j) * a.n + k;
    const uint64_t sbase = stream_id(TAG_BSK, (uint64_t)ref_e * 2 * d) + (uint64_t)h * 2;
    __syncthreads();
    ntt_forward<LOGN, NT>(m, 1, a.A, a.tw);
    W *o = a.bsk + (size_t)dev_rgsw * a.rgsw_elems;
    // column g = 0 carries m in a (K0a, K0b); g = 1 carries m in b (K1a, K1b)
    for (int g = 0; g < 2; g++) {
        rlwe_enc_zero<W, LOGN, NT>(e0, a.zn, a.znp, a.seed, sbase + g, a);
        for (int i = threadIdx.x; i < N; i += NT) {
            uint64_t va = e0[pad(i)], vb = e0[pad(N + i)];
            uint64_t t = (g == 0 ? va : vb) + (uint64_t)m[pad(i)];
            t = t >= Q ? t - Q : t;
            if (g == 0) va = t;
            else vb = t;
            const int idx = (h << LOGN) + i;
            if (sizeof(W) == 4) {
                reinterpret_cast<uint2 *>(o)[2 * idx + g] = make_uint2((uint32_t)va, (uint32_t)vb);
            } else {
                ulonglong2 *p = reinterpret_cast<ulonglong2 *>(o) + 4 * (size_t)idx + 2 * g;
                p[0] = make_ulonglong2(va, a.sg(va));
                p[1] = make_ulonglong2(vb, a.sg(vb));
            }
        }
        __syncthreads();  // e0 is reused by the next column
    }
}

// LWE key-switch key (LWE.cpp:55-106): one wave per row.
template <typename W>
__global__ __launch_bounds__(256) void k_ksk_gen(KskGenArgs<W> a) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + wave;  // device row ((kk * dks + j) * Bks + dd)
    if (row >= a.rows) return;
    const int dd = row % a.Bks;
    const int64_t kj = row / a.Bks;
    const int j = kj % a.dks, kk = kj / a.dks;
    const int64_t ref_e = ((int64_t)dd * a.dks + j) * a.N + kk;
    const uint64_t key = stream_key(a.seed, stream_id(TAG_KSK, (uint64_t)ref_e));
    const uint64_t Q = a.A.Q;
    W *o = a.ksk + row * a.n_pad;
    uint64_t part = 0;
    for (int h = lane; h < a.n; h += 64) {
        W av = (W)__umul64hi(stream_word(key, h), Q);
        o[h] = av;
        uint64_t p = mulmod_generic<W>(a.A, av, a.s[h], a.sp[h]);
        part += p;
        if (part >= Q) part -= Q;
    }
    for (int h = a.n + 1 + lane; h < a.n_pad; h += 64) o[h] = 0;
    // wave reduction mod Q
    for (int off = 32; off > 0; off >>= 1) {
        uint64_t o2 = __shfl_down(part, off, 64);
        part += o2;
        if (part >= Q) part -= Q;
    }
    if (lane == 0) {
        const int ci = dd * a.dks + j;
        uint64_t mult = mulmod_generic<W>(a.A, (W)a.z[kk], (W)a.cij[ci], (W)a.cijp[ci]);
        int64_t e = gauss_dev(a.cdt, a.cdt_len, stream_word(key, a.n)) % (int64_t)a.E_rlwe;
        uint64_t b = e < 0 ? mult + Q + e : mult + e;  // (e + mult) mod Q
        b = b >= Q ? b - Q : b;
        b += part;
        b = b >= Q ? b - Q : b;
        o[a.n] = (W)b;
    }
}

// Subs key (app.cpp:4-24 with RLWE.cpp:293-322): key[i][l] = RLWE_0 with
// NTT(z(X^(N/2^i+1))) * Bks_R^l added to b.  One workgroup per (i, l).
template <typename W, int LOGN, int NT>
__global__ __launch_bounds__(NT) void k_subs_gen(KeygenArgs<W> a) {
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem[];
    W *buf = reinterpret_cast<W *>(smem);
    W *old = buf + padded(2 * N);
    const int l = blockIdx.x % a.d, i = blockIdx.x / a.d;
    const uint64_t Q = a.A.Q;
    // substituted secret (time domain) scaled by Bks_R^l, then transformed
    const int power = (N >> i) + 1;
    for (int k = threadIdx.x; k < N; k += NT) {
        const int64_t ip = (int64_t)k * power;
        const int io = (int)(ip & (N - 1));
        uint64_t v = a.z[k];
        old[pad(io)] = (W)(((ip >> LOGN) & 1) ? (v ? Q - v : 0) : v);
    }
    __syncthreads();
    ntt_forward<LOGN, NT>(old, 1, a.A, a.tw);
    rlwe_enc_zero<W, LOGN, NT>(buf, a.zn, a.znp, a.seed, stream_id(TAG_SUBS, (uint64_t)i * 64) + l, a);
    W *o = a.bsk + (size_t)(i * a.d + l) * a.rgsw_elems;
    for (int k = threadIdx.x; k < N; k += NT) {
        // the reference scales the NTT-domain old key by Bks_R once per l
        W sc = mulmod_generic<W>(a.A, old[pad(k)], (W)a.bg_pow[l], (W)a.bg_powp[l]);
        uint64_t t = (uint64_t)buf[pad(N + k)] + sc;
        uint64_t vb = t >= Q ? t - Q : t, va = buf[pad(k)];
        if (sizeof(W) == 4) {
            reinterpret_cast<uint2 *>(o)[k] = make_uint2((uint32_t)va, (uint32_t)vb);
        } else {
            ulonglong2 *p = reinterpret_cast<ulonglong2 *>(o) + 2 * (size_t)k;
            p[0] = make_ulonglong2(va, a.sg(va));
            p[1] = make_ulonglong2(vb, a.sg(vb));
        }
    }
}

// RLWE_encrypt (RLWE.cpp:324-344) of message polynomials given in the time
// domain: the encryption of zero of stream (TAG_RLWE_ENC, first + c), then
// NTT(m) added to b.  One workgroup per ciphertext.
template <typename W, int LOGN, int NT>
__global__ __launch_bounds__(NT) void k_rlwe_encrypt(EncArgs<W> a) {
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem[];
    W *buf = reinterpret_cast<W *>(smem);  // the encryption of zero (a, b)
    W *m = buf + padded(2 * N);            // the message
    const int64_t c = blockIdx.x;
    const uint64_t Q = a.k.A.Q;
    for (int k = threadIdx.x; k < N; k += NT) m[pad(k)] = (W)(a.msg[c * N + k] % Q);
    __syncthreads();
    ntt_forward<LOGN, NT>(m, 1, a.k.A, a.k.tw);
    rlwe_enc_zero<W, LOGN, NT>(buf, a.k.zn, a.k.znp, a.k.seed, stream_id(TAG_RLWE_ENC, a.first + c), a.k);
    uint64_t *o = a.out + c * 2 * N;
    for (int k = threadIdx.x; k < N; k += NT) {
        const uint64_t t = (uint64_t)buf[pad(N + k)] + m[pad(k)];
        o[k] = buf[pad(k)];
        o[N + k] = t >= Q ? t - Q : t;
    }
}

// RGSW encryption of message polynomials (RLWE.cpp:561-577), written in the
// reference RGSW layout: workgroup (e, h) writes row h of RGSW e, column g =
// the encryption of zero of stream (TAG_RGSW_ENC, (first + e) 2d) + 2h + g
// with NTT(m) BG^h added to a (g = 0) or b (g = 1).
template <typename W, int LOGN, int NT>
__global__ __launch_bounds__(NT) void k_rgsw_encrypt(EncArgs<W> a) {
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem[];
    W *buf = reinterpret_cast<W *>(smem);
    W *m = buf + padded(2 * N);
    const int d = a.k.d;
    const int64_t e = blockIdx.x / d;
    const int h = blockIdx.x % d;
    const uint64_t Q = a.k.A.Q;
    for (int k = threadIdx.x; k < N; k += NT) m[pad(k)] = (W)(a.msg[e * N + k] % Q);
    __syncthreads();
    ntt_forward<LOGN, NT>(m, 1, a.k.A, a.k.tw);
    for (int k = threadIdx.x; k < N; k += NT)
        m[pad(k)] = mulmod_generic<W>(a.k.A, m[pad(k)], (W)a.k.bg_pow[h], (W)a.k.bg_powp[h]);
    const uint64_t sbase = stream_id(TAG_RGSW_ENC, (a.first + e) * 2 * d) + 2 * (uint64_t)h;
    uint64_t *o = a.out + ((size_t)e * d + h) * 4 * N;
    for (int g = 0; g < 2; g++) {
        __syncthreads();  // buf is rewritten; m is complete
        rlwe_enc_zero<W, LOGN, NT>(buf, a.k.zn, a.k.znp, a.k.seed, sbase + g, a.k);
        for (int k = threadIdx.x; k < N; k += NT) {
            uint64_t va = buf[pad(k)], vb = buf[pad(N + k)];
            uint64_t &t = g == 0 ? va : vb;
            t += m[pad(k)];
            t = t >= Q ? t - Q : t;
            o[(size_t)g * 2 * N + k] = va;
            o[(size_t)g * 2 * N + N + k] = vb;
        }
    }
}

// RLWE_decrypt(s, c, t) (RLWE.cpp:362-376): b - a s in the NTT domain, the
// inverse transform, then denoise(t) (RLWE.cpp:138-155): (x + scale/2) mod Q
// divided by scale = Q / t.
template <typename W, int LOGN, int NT>
__global__ __launch_bounds__(NT) void k_rlwe_decrypt(EncArgs<W> a) {
    constexpr int N = 1 << LOGN;
    extern __shared__ __align__(16) unsigned char smem[];
    W *m = reinterpret_cast<W *>(smem);
    const int64_t c = blockIdx.x;
    const uint64_t Q = a.k.A.Q;
    const uint64_t *x = a.ct + c * 2 * N;
    for (int k = threadIdx.x; k < N; k += NT) {
        const W as = mulmod_generic<W>(a.k.A, (W)x[k], a.k.zn[k], a.k.znp[k]);
        const W b = (W)x[N + k];
        m[pad(k)] = b >= as ? b - as : (W)(b + (Q - as));
    }
    __syncthreads();
    ntt_inverse<LOGN, NT>(m, 1, a.k.A, a.k.tw);
    const uint64_t half = a.scale / 2;
    for (int k = threadIdx.x; k < N; k += NT) {
        uint64_t v = (uint64_t)m[pad(k)] + half;
        v = v >= Q ? v - Q : v;
        a.out[c * N + k] = v / a.scale;
    }
}

// subs key device -> reference [logN][dks_R][2][N]
template <typename W>
__global__ void k_subs_to_ref(const W *src, uint64_t *dst, int N, int64_t total) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t il = e / N;
        const int k = e - il * N;
        uint64_t va, vb;
        if (sizeof(W) == 4) {
            uint2 v = reinterpret_cast<const uint2 *>(src)[e];
            va = v.x; vb = v.y;
        } else {
            const ulonglong2 *p = reinterpret_cast<const ulonglong2 *>(src) + 2 * e;
            va = p[0].x; vb = p[1].x;
        }
        dst[il * 2 * N + k] = va;
        dst[il * 2 * N + N + k] = vb;
    }
}

template <typename W>
__global__ void k_subs_from_ref(const uint64_t *src, W *dst, int N, int64_t total, ShoupGen sg) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t il = e / N;
        const int k = e - il * N;
        uint64_t va = src[il * 2 * N + k], vb = src[il * 2 * N + N + k];
        if (sizeof(W) == 4) {
            reinterpret_cast<uint2 *>(dst)[e] = make_uint2((uint32_t)va, (uint32_t)vb);
        } else {
            ulonglong2 *p = reinterpret_cast<ulonglong2 *>(dst) + 2 * e;
            p[0] = make_ulonglong2(va, sg(va));
            p[1] = make_ulonglong2(vb, sg(vb));
        }
    }
}

// ---------------------------------------------------------------------------
// host launchers (template dispatch on word width, log2 N, workgroup size)
// ---------------------------------------------------------------------------

#define FHEACC_LOGN_SWITCH(logN, ...)                   \
    switch (logN) {                                      \
        case 4: { constexpr int LOGN = 4; __VA_ARGS__; } break;   \
        case 5: { constexpr int LOGN = 5; __VA_ARGS__; } break;   \
        case 6: { constexpr int LOGN = 6; __VA_ARGS__; } break;   \
        case 7: { constexpr int LOGN = 7; __VA_ARGS__; } break;   \
        case 8: { constexpr int LOGN = 8; __VA_ARGS__; } break;   \
        case 9: { constexpr int LOGN = 9; __VA_ARGS__; } break;   \
        case 10: { constexpr int LOGN = 10; __VA_ARGS__; } break; \
        case 11: { constexpr int LOGN = 11; __VA_ARGS__; } break; \
        default: return hipErrorInvalidValue;            \
    }

// Workgroups of one launch: a grid dimension must stay below 2^32 work-items,
// so larger batches go out as several launches (the kernels add a base to
// their workgroup index).
template <int NT>
constexpr int64_t max_wgs_per_launch() {
    return ((int64_t)1 << 31) / NT;
}

// Clears the lock-step counters of a tile.
template <typename W>
__global__ void k_clear_pacing(uint32_t *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

template <typename W, int LOGN, int NT>
static hipError_t launch_br_nt(const BRArgs<W> &a, int64_t B, size_t lds, hipStream_t st, int persist, int cus,
                               int64_t *grid) {
    // the paced persistent kernel exists for the throughput shapes (256 and
    // 512 threads, N >= 512)
    constexpr bool kPaceable = LOGN >= 9 && (NT == 256 || NT == 512);
    const int form = product_form(a.pb, a.d);
    auto k = form == kInPlace ? k_blind_rotate<W, LOGN, NT, kInPlace>
                              : (form == kSplit ? k_blind_rotate<W, LOGN, NT, kSplit>
                                                : k_blind_rotate<W, LOGN, NT, kSeparate>);
    constexpr int kSpecD = specialized_digits<W, LOGN, NT>();
    if constexpr (kSpecD > 0)
        if (form == kInPlace && a.d == kSpecD) k = k_blind_rotate<W, LOGN, NT, kInPlace, kSpecD>;
    if constexpr (kPaceable) {
        if (persist > 0 && cus > 0 && a.sync) {
            auto kp = form == kInPlace ? k_blind_rotate_paced<W, LOGN, NT, kInPlace>
                                       : (form == kSplit ? k_blind_rotate_paced<W, LOGN, NT, kSplit>
                                                         : k_blind_rotate_paced<W, LOGN, NT, kSeparate>);
            if constexpr (kSpecD > 0)
                if (form == kInPlace && a.d == kSpecD) kp = k_blind_rotate_paced<W, LOGN, NT, kInPlace, kSpecD>;
            hipError_t e = hipFuncSetAttribute((const void *)kp, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            int occ = 0;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)kp, NT, lds);
            if (e != hipSuccess) return e;
            const int64_t G = std::min<int64_t>(B, (int64_t)std::max(1, std::min(occ, persist)) * cus);
            // a group's members are read one per thread
            if (G <= (int64_t)a.sync_groups * NT) {
                BRArgs<W> part = a;
                part.c_end = a.c_base + B;
                // progress words, one per workgroup, cleared by a kernel (not a
                // memset) so that the launch can be captured
                hipLaunchKernelGGL(k_clear_pacing<W>, dim3((unsigned)std::min<int64_t>((G + 255) / 256, 1024)),
                                   dim3(256), 0, st, part.sync, G);
                hipLaunchKernelGGL(kp, dim3((unsigned)G), dim3(NT), lds, st, part);
                if (grid) *grid = G;
                return hipGetLastError();
            }
        }
    }
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    // tiles: launches of one resident wave of workgroups each (G of them, at
    // most the occupancy per CU), back to back; otherwise one launch (split
    // only at the 2^31 work-item limit) that the dispatcher refills as gates end
    int64_t G = max_wgs_per_launch<NT>();
    if (persist > 0 && cus > 0) {
        int occ = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void *)k, NT, lds);
        if (e != hipSuccess) return e;
        G = (int64_t)std::max(1, std::min(occ, persist)) * cus;
    }
    if (grid) *grid = std::min(B, G);
    for (int64_t off = 0; off < B; off += G) {
        BRArgs<W> part = a;
        const int64_t cnt = std::min(B - off, G);
        part.c_base = a.c_base + off;
        part.c_end = part.c_base + cnt;
        part.sync = nullptr;
        hipLaunchKernelGGL(k, dim3((unsigned)cnt), dim3(NT), lds, st, part);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename W>
hipError_t launch_blind_rotate(const BRArgs<W> &a, int logN, int nt, int64_t B, size_t lds, hipStream_t st,
                               int persist, int cus, int64_t *grid) {
    FHEACC_LOGN_SWITCH(logN, {
        if (nt == 128 && LOGN >= 7) return launch_br_nt<W, LOGN, 128>(a, B, lds, st, persist, cus, grid);
        if (nt == 512 && LOGN >= 9) return launch_br_nt<W, LOGN, 512>(a, B, lds, st, persist, cus, grid);
        if constexpr (LOGN >= 10)
            if (nt == 1024) return launch_br_nt<W, LOGN, 1024>(a, B, lds, st, persist, cus, grid);
        return launch_br_nt<W, LOGN, 256>(a, B, lds, st, persist, cus, grid);
    })
    return hipSuccess;
}

template <typename W>
hipError_t launch_lwe_keyswitch(const KSArgs<W> &a, bool gather, hipStream_t st) {
    if (gather) {
        const unsigned cb = (unsigned)((a.n + 1 + 255) / 256);
        hipLaunchKernelGGL((k_lwe_keyswitch_gather<W, 256>), dim3(cb * (unsigned)a.split, (unsigned)a.B), dim3(256), 0,
                           st, a);
        if (a.split > 1)
            hipLaunchKernelGGL((k_lwe_keyswitch_finish<W, 256>), dim3(cb, (unsigned)a.B), dim3(256), 0, st, a);
        return hipGetLastError();
    }
    constexpr int KB = sizeof(W) == 4 ? 16 : 8;
    // a.nt ciphertexts share each staged key stream (one per thread); 1024
    // threads only on 64-bit words (32-bit words need 16 columns of sums per
    // thread, which spill at 128 registers)
    const int nt = (a.nt == 1024 && sizeof(W) == 8) || a.nt == 512 ? a.nt : 256;
    dim3 grid((a.n + 1 + KB - 1) / KB, (unsigned)((a.B + nt - 1) / nt));
    const size_t lds = ks_staged_lds<W>(a.Bks, nt);
    auto k = nt == 512 ? k_lwe_keyswitch_staged<W, 512, KB> : k_lwe_keyswitch_staged<W, 256, KB>;
    if constexpr (sizeof(W) == 8)
        if (nt == 1024) k = k_lwe_keyswitch_staged<W, 1024, KB>;
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, grid, dim3(nt), lds, st, a);
    return hipGetLastError();
}

template <typename W>
size_t ks_staged_lds(uint32_t Bks, int nt) {
    constexpr int KB = sizeof(W) == 4 ? 16 : 8;
    const size_t fill = (size_t)KS_SB * Bks * (KB + 16 / sizeof(W)) * sizeof(W);
    return ks_double_buffered<W, KB>(Bks, nt) ? 2 * fill : fill;
}

template <typename W>
hipError_t launch_ntt(const NTTArgs<W> &a, int logN, hipStream_t st) {
    const int64_t wgs = (a.count + a.per_wg - 1) / a.per_wg;
    const size_t lds = (size_t)padded(a.per_wg << logN) * sizeof(W);
    FHEACC_LOGN_SWITCH(logN, {
        // N >= 1024: two polys per workgroup; a compile-time count when
        // every workgroup has both
        if constexpr (LOGN >= 10) {
            if (a.per_wg == 2 && a.count % 2 == 0) {
                hipLaunchKernelGGL((k_ntt<W, LOGN, 256, 2>), dim3((unsigned)wgs), dim3(256), lds, st, a);
                return hipGetLastError();
            }
        }
        hipLaunchKernelGGL((k_ntt<W, LOGN, 256>), dim3((unsigned)wgs), dim3(256), lds, st, a);
    })
    return hipGetLastError();
}

template <typename W, int LOGN, int NT, bool INPLACE>
static hipError_t launch_ep_nt(const EPArgs<W> &a, int64_t count, size_t lds, hipStream_t st) {
    auto k = k_external_product<W, LOGN, NT, INPLACE>;
    // the standalone product's own digit-count kernel: 54-bit N = 1024, d = 6
    // (BG = 2^9), the external-product microbenchmark's shape
    if constexpr (INPLACE && sizeof(W) == 8 && LOGN == 10 && NT == 1024)
        if (a.d == 6) k = k_external_product<W, LOGN, NT, INPLACE, 6>;
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    for (int64_t off = 0; off < count; off += max_wgs_per_launch<NT>()) {
        EPArgs<W> part = a;
        part.c_base = off;
        hipLaunchKernelGGL(k, dim3((unsigned)std::min(count - off, max_wgs_per_launch<NT>())), dim3(NT), lds, st,
                           part);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// One workgroup per product.  All 2d digit polys fit the LDS at every
// supported shape but one workgroup per CU at large N, so the in-place
// product runs 1024 threads from N = 1024 up (16 waves per CU instead of 4).
template <typename W>
hipError_t launch_external_product(const EPArgs<W> &a, int logN, int64_t count, hipStream_t st) {
    const bool inplace = a.pb >= 2 * a.d;
    const size_t lds = inplace ? (size_t)padded(std::max(2, 2 * a.d) << logN) * sizeof(W)
                               : (size_t)(padded(2 << logN) + padded(a.pb << logN)) * sizeof(W);
    FHEACC_LOGN_SWITCH(logN, {
        if (!inplace) return launch_ep_nt<W, LOGN, 256, false>(a, count, lds, st);
        if constexpr (LOGN >= 10) return launch_ep_nt<W, LOGN, 1024, true>(a, count, lds, st);
        else return launch_ep_nt<W, LOGN, 256, true>(a, count, lds, st);
    })
    return hipSuccess;
}

// One workgroup per RLWE.  From N = 1024 up the LDS (4 polys plus the digit
// batch) leaves one workgroup per CU, so it runs 1024 threads (16 waves per
// CU instead of 4).
template <typename W, int LOGN, bool EXPAND>
static hipError_t launch_subs_nt(const SubsArgs<W> &a, int64_t count, size_t lds, hipStream_t st) {
    constexpr int NT = LOGN >= 10 ? 1024 : 256;
    auto k = k_subs_keyswitch<W, LOGN, NT, EXPAND>;
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    for (int64_t off = 0; off < count; off += max_wgs_per_launch<NT>()) {
        SubsArgs<W> part = a;
        part.c_base = off;
        hipLaunchKernelGGL(k, dim3((unsigned)std::min(count - off, max_wgs_per_launch<NT>())), dim3(NT), lds, st,
                           part);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename W>
hipError_t launch_subs_keyswitch(const SubsArgs<W> &a, int logN, int64_t count, hipStream_t st) {
    const size_t lds = (size_t)(2 * padded(2 << logN) + padded(a.dch << logN)) * sizeof(W);
    FHEACC_LOGN_SWITCH(logN, { return launch_subs_nt<W, LOGN, false>(a, count, lds, st); })
    return hipSuccess;
}

template <typename W>
hipError_t launch_expand_level(const SubsArgs<W> &a, int logN, int64_t B, hipStream_t st) {
    const size_t lds = (size_t)(2 * padded(2 << logN) + padded(a.dch << logN)) * sizeof(W);
    const int64_t count = B << a.sub_factor;
    FHEACC_LOGN_SWITCH(logN, { return launch_subs_nt<W, LOGN, true>(a, count, lds, st); })
    return hipSuccess;
}

template <typename W, int LOGN, int NT>
static hipError_t launch_cmux_nt(const CmuxArgs<W> &a, int64_t count, size_t lds, hipStream_t st) {
    const int form = product_form(a.pb, a.d);
    auto k = form == kInPlace ? k_cmux_chain<W, LOGN, NT, kInPlace>
                              : (form == kSplit ? k_cmux_chain<W, LOGN, NT, kSplit> : k_cmux_chain<W, LOGN, NT, kSeparate>);
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    for (int64_t off = 0; off < count; off += max_wgs_per_launch<NT>()) {
        CmuxArgs<W> part = a;
        part.c_base = off;
        hipLaunchKernelGGL(k, dim3((unsigned)std::min(count - off, max_wgs_per_launch<NT>())), dim3(NT), lds, st,
                           part);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <typename W>
hipError_t launch_cmux_chain(const CmuxArgs<W> &a, int logN, int nt, int64_t count, size_t lds, hipStream_t st) {
    FHEACC_LOGN_SWITCH(logN, {
        if (nt == 128 && LOGN >= 7) return launch_cmux_nt<W, LOGN, 128>(a, count, lds, st);
        if (nt == 512 && LOGN >= 9) return launch_cmux_nt<W, LOGN, 512>(a, count, lds, st);
        if constexpr (LOGN >= 10)
            if (nt == 1024) return launch_cmux_nt<W, LOGN, 1024>(a, count, lds, st);
        return launch_cmux_nt<W, LOGN, 256>(a, count, lds, st);
    })
    return hipGetLastError();
}

template <typename W>
hipError_t launch_rgsw_to_device(const uint64_t *src, int64_t src_stride, W *dst, int64_t dst_stride, int d, int N,
                                 ShoupGen sg, const int64_t *map, int64_t count, hipStream_t st, bool ddr) {
    int bx = (d * N + 255) / 256;
    for (int64_t off = 0; off < count; off += 65535) {
        int64_t cnt = count - off < 65535 ? count - off : 65535;
        hipLaunchKernelGGL(k_rgsw_to_device<W>, dim3(bx, (unsigned)cnt), dim3(256), 0, st,
                           map ? src : src + off * src_stride, src_stride, dst + off * dst_stride, dst_stride, d, N,
                           sg, map ? map + off : nullptr, cnt, ddr);
    }
    return hipGetLastError();
}

template <typename W>
hipError_t launch_rgsw_to_ref(const W *src, int64_t src_stride, uint64_t *dst, int64_t dst_stride, int d, int N,
                              const int64_t *map, int64_t count, hipStream_t st, bool ddr) {
    int bx = (d * N + 255) / 256;
    for (int64_t off = 0; off < count; off += 65535) {
        int64_t cnt = count - off < 65535 ? count - off : 65535;
        hipLaunchKernelGGL(k_rgsw_to_ref<W>, dim3(bx, (unsigned)cnt), dim3(256), 0, st, src + off * src_stride,
                           src_stride, dst, dst_stride, d, N, map ? map + off : nullptr, cnt, ddr);
    }
    return hipGetLastError();
}

template <typename W>
hipError_t launch_ksk_convert(const uint64_t *src, W *dst, uint64_t *back, int Bks, int dks, int N, int n1,
                              int n_pad, hipStream_t st) {
    const int64_t rows = (int64_t)N * dks * Bks;
    hipLaunchKernelGGL(k_ksk_convert<W>, dim3((unsigned)rows), dim3(128), 0, st, src, dst, back, Bks, dks, N, n1,
                       n_pad);
    return hipGetLastError();
}

template <typename W>
hipError_t launch_keygen(const KeygenArgs<W> &a, const KskGenArgs<W> &ka, int logN, const uint64_t *z, hipStream_t st,
                         bool subs) {
    FHEACC_LOGN_SWITCH(logN, {
        constexpr int N = 1 << LOGN;
        hipLaunchKernelGGL((k_secret_ntt<W, LOGN, 256>), dim3(1), dim3(256), padded(N) * sizeof(W), st, z, (W *)a.zn,
                           (W *)a.znp, a.A, a.tw, a.sg);
        if (subs) {
            size_t lds = (padded(2 * N) + padded(N)) * sizeof(W);
            hipLaunchKernelGGL((k_subs_gen<W, LOGN, 256>), dim3((unsigned)(a.n * a.d)), dim3(256), lds, st, a);
        } else {
            size_t lds = (padded(2 * N) + padded(N)) * sizeof(W);
            auto k = k_bsk_gen<W, LOGN, 256>;
            hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
            const int64_t wgs = (int64_t)a.n * a.digitR * a.Br * a.d;
            hipLaunchKernelGGL(k, dim3((unsigned)wgs), dim3(256), lds, st, a);
            hipLaunchKernelGGL(k_ksk_gen<W>, dim3((unsigned)((ka.rows + 3) / 4)), dim3(256), 0, st, ka);
        }
    })
    return hipGetLastError();
}

template <typename W>
hipError_t launch_secret_ntt(const KeygenArgs<W> &a, int logN, const uint64_t *z, hipStream_t st) {
    FHEACC_LOGN_SWITCH(logN, {
        constexpr int N = 1 << LOGN;
        hipLaunchKernelGGL((k_secret_ntt<W, LOGN, 256>), dim3(1), dim3(256), padded(N) * sizeof(W), st, z, (W *)a.zn,
                           (W *)a.znp, a.A, a.tw, a.sg);
    })
    return hipGetLastError();
}

template <typename W>
hipError_t launch_encrypt(const EncArgs<W> &a, int which, int logN, int64_t count, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    if (count * (which == 1 ? a.k.d : 1) > 0x7fffffff) return hipErrorInvalidValue;
    FHEACC_LOGN_SWITCH(logN, {
        constexpr int N = 1 << LOGN;
        constexpr int NT = 256;
        if (which == 0) {
            hipLaunchKernelGGL((k_rlwe_encrypt<W, LOGN, NT>), dim3((unsigned)count), dim3(NT),
                               (padded(2 * N) + padded(N)) * sizeof(W), st, a);
        } else if (which == 1) {
            hipLaunchKernelGGL((k_rgsw_encrypt<W, LOGN, NT>), dim3((unsigned)(count * a.k.d)), dim3(NT),
                               (padded(2 * N) + padded(N)) * sizeof(W), st, a);
        } else {
            hipLaunchKernelGGL((k_rlwe_decrypt<W, LOGN, NT>), dim3((unsigned)count), dim3(NT), padded(N) * sizeof(W),
                               st, a);
        }
    })
    return hipGetLastError();
}

template <typename W>
hipError_t launch_subs_convert(const void *src, void *dst, int N, int64_t total, bool to_ref, ShoupGen sg,
                               hipStream_t st) {
    int blocks = (int)((total + 255) / 256);
    if (blocks > 65535) blocks = 65535;
    if (to_ref) hipLaunchKernelGGL(k_subs_to_ref<W>, dim3(blocks), dim3(256), 0, st, (const W *)src, (uint64_t *)dst, N, total);
    else hipLaunchKernelGGL(k_subs_from_ref<W>, dim3(blocks), dim3(256), 0, st, (const uint64_t *)src, (W *)dst, N, total, sg);
    return hipGetLastError();
}


}  // namespace fheacc
