// modarith.hpp -- modular arithmetic on the integer VALU of gfx950.
//
// Two word widths share every kernel through the Arith<W> policy:
//   Arith<uint32_t>  "narrow", Q < 2^30 (27- to 30-bit sets): one VGPR per
//                    word, Shoup products from v_mul_hi_u32/v_mul_lo_u32 and
//                    64-bit lazy sums from v_mad_u64_u32.
//   Arith<uint64_t>  "wide", Q < 2^62 (54-bit sets): Shoup products built on
//                    __umul64hi.
// Every result that leaves a kernel is canonical in [0, Q), the reference's
// representation (FHE_ACC_runtime/include/nbt.h:59-101); intermediate values
// are kept lazily below small multiples of Q (each transform states its
// bounds), which is exact arithmetic mod Q and so does not change any
// canonical output.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fheacc {

template <typename W>
struct Arith;

template <>
struct Arith<uint32_t> {
    using W = uint32_t;
    W Q, Q2;
    uint64_t mu;  // floor(2^64 / Q), Barrett constant for 64-bit sums
    // Harvey's forward butterflies (words below 4Q) instead of fully lazy
    // ones: Q too large for (2 log2 N + 1) Q < 2^32, small enough for 4Q
    bool harvey;
    // inverse passes of up to lzr stages leave their sums unreduced:
    // 2^(lzr + 1) Q <= 2^32
    int lzr;

    // [0, 2m) -> [0, m): if x < m the subtraction wraps above x
    __device__ __forceinline__ static W red(W x, W m) {
        W y = x - m;
        return y < x ? y : x;
    }
    // x * w mod Q in [0, 2Q) for any 32-bit x; wp = floor(w * 2^32 / Q)
    __device__ __forceinline__ W mul_shoup(W x, W w, W wp) const {
        W qh = __umulhi(x, wp);
        return x * w - qh * Q;
    }
    // Lazy Cooley-Tukey butterfly (a, b) -> (a + wb, a - wb + 2Q), wb = b w
    // mod Q in [0, 2Q).  Written so the VALU needs three multiplies and three
    // adds: a' = a + b w - qh Q in one 3-input add (the -Q factor folded into
    // the multiply), b' = 2a + 2Q - a' in one shift-add and one subtract.
    // Sums wrap mod 2^32 on the way but the true results fit in 32 bits.
    __device__ __forceinline__ void ct_butterfly(W &a, W &b, W w, W wp) const {
        const W qh = __umulhi(b, wp);
        const W a0 = a;
        a = a0 + b * w + qh * (0u - Q);
        b = (a0 << 1) + Q2 - a;
    }
    // 64-bit sum -> [0, Q)
    __device__ __forceinline__ W reduce64(uint64_t x) const {
        uint64_t qh = __umul64hi(x, mu);
        W r = (W)x - (W)qh * Q;  // true remainder in [0, 2Q), exact mod 2^32
        return red(r, Q);
    }
    __device__ __forceinline__ static uint64_t mac(uint64_t acc, W a, W b) {
        return acc + (uint64_t)a * b;
    }
};

template <>
struct Arith<uint64_t> {
    using W = uint64_t;
    W Q, Q2;
    uint64_t mu;  // unused for the wide path (sums are reduced from [0, 10dQ))
    W Q5;         // 5Q: the bound of mul_shoup_lazy
    W NQ;         // 2^64 - Q: the remainder x w - qh Q is x w + qh NQ mod 2^64
    double invQ;  // 1 / Q rounded to double, for reduce_small's quotient
    int lzr;      // as for 32-bit words: 2^(lzr + 1) Q <= 2^64

    __device__ __forceinline__ static W red(W x, W m) {
        W y = x - m;
        return y < x ? y : x;
    }
    // x * w mod Q in [0, 2Q) for any 64-bit x; wp = floor(w * 2^64 / Q).
    // The remainder is written x w + qh (2^64 - Q) so that the second low
    // product accumulates into the first one's v_mad_u64_u32 (two VALU
    // instructions fewer than a product and a 64-bit subtraction).
    __device__ __forceinline__ W mul_shoup(W x, W w, W wp) const {
        W qh = __umul64hi(x, wp);
        return x * w + qh * NQ;
    }
    // x * w mod Q in [0, 5Q) for any 64-bit x.  The Shoup quotient
    // floor(x wp / 2^64) is taken from three 32-bit partial products: the
    // low-by-low product and the carries out of the low word are dropped,
    // each worth less than one unit, so the estimate is at most 3 short and
    // the remainder below 5Q.  One half-rate v_mad_u64_u32 and two
    // v_mul_hi_u32 instead of the full 64x64 high product.
    __device__ __forceinline__ W mul_shoup_lazy(W x, W w, W wp) const {
        const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
        const uint32_t pl = (uint32_t)wp, ph = (uint32_t)(wp >> 32);
        const uint64_t mid = (uint64_t)__umulhi(xl, ph) + __umulhi(xh, pl);
        const uint64_t qh = (uint64_t)xh * ph + mid;
        return x * w + qh * NQ;
    }
    // (a, b) -> (a + wb, a - wb + 5Q), wb = b w mod Q in [0, 5Q): a word
    // entering stage s is below (5s + 1) Q
    __device__ __forceinline__ void ct_butterfly(W &a, W &b, W w, W wp) const {
        const W v = mul_shoup_lazy(b, w, wp);
        const W a0 = a;
        a = a0 + v;
        b = a0 - v + Q5;
    }
    // x < 2^62, any multiple of Q below that -> [0, Q).  The quotient
    // x / Q < 2^8 comes from a double multiply by 1/Q: three roundings of
    // relative size 2^-53 each move it by far less than one unit, so the
    // truncated estimate is the true quotient or one off, and the remainder
    // lies in [-Q, 2Q), which the corrections below bring to [0, Q).  (A
    // double division instead costs a dozen VALU instructions.)
    __device__ __forceinline__ W reduce_small(uint64_t x) const {
        uint64_t qh = (uint64_t)((double)x * invQ);
        int64_t r = (int64_t)(x - qh * Q);
        if (r < 0) r += (int64_t)Q;
        if (r >= (int64_t)Q) r -= (int64_t)Q;
        if (r < 0) r += (int64_t)Q;
        return (W)r;
    }
};

// Host-side constant derivation (Shoup companions, Barrett mu).
static inline uint32_t shoup32(uint32_t w, uint32_t Q) {
    return (uint32_t)(((unsigned __int128)w << 32) / Q);
}
static inline uint64_t shoup64(uint64_t w, uint64_t Q) {
    return (uint64_t)(((unsigned __int128)w << 64) / Q);
}

}  // namespace fheacc
