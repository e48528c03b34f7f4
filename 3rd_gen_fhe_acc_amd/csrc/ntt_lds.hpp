// ntt_lds.hpp -- workgroup-cooperative negacyclic NTT / inverse NTT on
// polynomials held in LDS.
//
// The transform is the reference's (FHE_ACC_runtime/src/nbt.cpp:423-491):
// forward Cooley-Tukey, normal order in, bit-reversed order out, stage with
// half-size t uses twiddle rou[m + j/(2t)] (m = N/(2t)); inverse
// Gentleman-Sande with irou[m + j/(2t)], then a scaling by N^-1.  The tables
// are the reference's ROU_table/iROU_table (params.cpp:149-175) plus Shoup
// companions.
//
// Stages are grouped into passes of up to four (a compile-time schedule per
// log2 N): each thread loads a cluster of 2^R words at stride S, runs R
// butterfly stages in registers and writes the cluster back, so a
// length-1024 transform is three LDS round trips (4+3+3 stages).
//
// LDS layout: logical word i of a run of polynomials lives at i + i/8.  One
// pad word per eight spreads the strided cluster accesses (strides 1..8
// words) over distinct banks; without it a stride-8 pass hits 4 of 32 banks.
//
// Forward butterflies are fully lazy: no reduction between stages.  A value
// entering stage s is below (2s+1)Q, so after log2 N stages every word is
// below (2 log2 N + 1) Q; the context only selects 32-bit words when that
// bound fits in 32 bits.  The last pass makes the words canonical when asked.
#pragma once
#include "modarith.hpp"

namespace fheacc {

template <typename W>
struct Twiddles {
    const W *fw, *fwp;   // rou, Shoup companions   [N]
    const W *iw, *iwp;   // irou, Shoup companions  [N]
    W iN, iNp;
    W onep;              // Shoup companion of 1: floor(2^w / Q)
};

// physical LDS index of logical word i
__host__ __device__ __forceinline__ constexpr int pad(int i) { return i + (i >> 3); }
// LDS words occupied by `words` logical words
__host__ __device__ __forceinline__ constexpr int padded(int words) { return words + (words >> 3) + 1; }

// stages per pass for `rem` remaining stages: 4s, never a lone 1-stage pass
// unless the transform has a single stage
__host__ __device__ constexpr int pass_stages(int rem) { return rem <= 4 ? rem : (rem <= 6 ? 3 : 4); }

// Physical offset of cluster element e (stride S) from the cluster's first
// word b0: pad(b0 + e S) - pad(b0).  When S 2^R >= 8, b0 = 8m + off with
// off < S, and the offset is a constant pad(e S) (the low bits never carry).
template <int S, int R>
struct ClusterOffsets {
    static constexpr bool kConst = S * (1 << R) >= 8;
    __device__ __forceinline__ static int at(int b0, int e) {
        if constexpr (kConst) return pad(e * S);
        else return pad(b0 + e * S) - pad(b0);
    }
};

// Forward pass: stages S0 .. S0+R-1 over np polys of 2^LOGN words each.
// When last, outputs are made canonical.
template <int LOGN, int S0, int R, int NT, typename W>
__device__ __forceinline__ void fwd_pass(W *s, int np, bool last, const Arith<W> &A, const Twiddles<W> &tw,
                                         int tid) {
    constexpr int N = 1 << LOGN;
    constexpr int lS = LOGN - S0 - R;  // log2 of the cluster stride
    constexpr int S = 1 << lS;
    constexpr int per_poly = N >> R;
    const int clusters = np * per_poly;
    for (int cg = tid; cg < clusters; cg += NT) {
        const int p = cg / per_poly, c = cg - p * per_poly;
        const int blk = c >> lS, off = c & (S - 1);
        const int b0 = p * N + (blk << (lS + R)) + off;
        W *base = s + pad(b0);
        W x[1 << R];
#pragma unroll
        for (int e = 0; e < (1 << R); e++) x[e] = base[ClusterOffsets<S, R>::at(b0, e)];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int h = 1 << (R - 1 - k);
            const int st = S0 + k;
#pragma unroll
            for (int e = 0; e < (1 << R); e++) {
                if (e & h) continue;
                const int ti = (1 << st) + (blk << k) + (e >> (R - k));
                const W v = A.mul_shoup(x[e + h], tw.fw[ti], tw.fwp[ti]);
                const W u = x[e];
                x[e] = u + v;
                x[e + h] = u - v + A.Q2;
            }
        }
        if (last) {
#pragma unroll
            for (int e = 0; e < (1 << R); e++) x[e] = Arith<W>::red(A.mul_shoup(x[e], (W)1, tw.onep), A.Q);
        }
#pragma unroll
        for (int e = 0; e < (1 << R); e++) base[ClusterOffsets<S, R>::at(b0, e)] = x[e];
    }
}

// Inverse pass: GS stages U0 .. U0+R-1 (half-size 2^u).  Inputs and outputs in
// [0, 2Q).  When last, the words are scaled by N^-1 and made canonical.
template <int LOGN, int U0, int R, int NT, typename W>
__device__ __forceinline__ void inv_pass(W *s, int np, const Arith<W> &A, const Twiddles<W> &tw, int tid) {
    constexpr int N = 1 << LOGN;
    constexpr int S = 1 << U0;
    constexpr int per_poly = N >> R;
    constexpr bool last = U0 + R == LOGN;
    const int clusters = np * per_poly;
    for (int cg = tid; cg < clusters; cg += NT) {
        const int p = cg / per_poly, c = cg - p * per_poly;
        const int blk = c >> U0, off = c & (S - 1);
        const int b0 = p * N + (blk << (U0 + R)) + off;
        W *base = s + pad(b0);
        W x[1 << R];
#pragma unroll
        for (int e = 0; e < (1 << R); e++) x[e] = base[ClusterOffsets<S, R>::at(b0, e)];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int h = 1 << k;
#pragma unroll
            for (int e = 0; e < (1 << R); e++) {
                if (e & h) continue;
                const int ti = (N >> (U0 + k + 1)) + (blk << (R - k - 1)) + (e >> (k + 1));
                const W u = x[e], v = x[e + h];
                x[e] = Arith<W>::red(u + v, A.Q2);
                x[e + h] = A.mul_shoup(u - v + A.Q2, tw.iw[ti], tw.iwp[ti]);
            }
        }
        if constexpr (last) {
#pragma unroll
            for (int e = 0; e < (1 << R); e++) x[e] = Arith<W>::red(A.mul_shoup(x[e], tw.iN, tw.iNp), A.Q);
        }
#pragma unroll
        for (int e = 0; e < (1 << R); e++) base[ClusterOffsets<S, R>::at(b0, e)] = x[e];
    }
}

template <int LOGN, int ST, int NT, typename W>
__device__ __forceinline__ void fwd_passes(W *s, int np, bool canonical, const Arith<W> &A, const Twiddles<W> &tw,
                                           int tid) {
    if constexpr (ST < LOGN) {
        constexpr int R = pass_stages(LOGN - ST);
        fwd_pass<LOGN, ST, R, NT>(s, np, canonical && ST + R == LOGN, A, tw, tid);
        __syncthreads();
        fwd_passes<LOGN, ST + R, NT>(s, np, canonical, A, tw, tid);
    }
}

template <int LOGN, int ST, int NT, typename W>
__device__ __forceinline__ void inv_passes(W *s, int np, const Arith<W> &A, const Twiddles<W> &tw, int tid) {
    if constexpr (ST < LOGN) {
        constexpr int R = pass_stages(LOGN - ST);
        inv_pass<LOGN, ST, R, NT>(s, np, A, tw, tid);
        __syncthreads();
        inv_passes<LOGN, ST + R, NT>(s, np, A, tw, tid);
    }
}

// Forward transform of np padded polys in LDS (canonical inputs).  Output
// canonical in [0, Q), or, when canonical is false, below (2 logN + 1) Q and
// congruent to the canonical word (for consumers that reduce anyway).
template <int LOGN, int NT, typename W>
__device__ void ntt_forward(W *s, int np, const Arith<W> &A, const Twiddles<W> &tw, bool canonical = true,
                            int tid = -1) {
    fwd_passes<LOGN, 0, NT>(s, np, canonical, A, tw, tid < 0 ? (int)threadIdx.x : tid);
}

// Inverse transform of np padded polys in LDS (inputs in [0, 2Q)); output
// canonical in [0, Q).
template <int LOGN, int NT, typename W>
__device__ void ntt_inverse(W *s, int np, const Arith<W> &A, const Twiddles<W> &tw, int tid = -1) {
    inv_passes<LOGN, 0, NT>(s, np, A, tw, tid < 0 ? (int)threadIdx.x : tid);
}

}  // namespace fheacc
