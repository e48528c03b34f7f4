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
// Stages are grouped into passes of up to four: each thread loads a cluster
// of 2^R words at stride S, runs R butterfly stages in registers and writes
// the cluster back, so a length-1024 transform is three LDS round trips
// (4+3+3 stages) and three barriers.
//
// Forward butterflies are fully lazy: no reduction between stages.  A value
// entering stage s is below (2s+1)Q, so after log2 N stages every word is
// below (2 log2 N + 1) Q; the context only selects 32-bit words when that
// bound fits in 32 bits.  The last pass makes the words canonical.
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

// Stages per pass for `rem` remaining stages: 4s, never a lone 1-stage pass
// unless the transform has a single stage.
__device__ __forceinline__ int pass_stages(int rem) {
    if (rem <= 4) return rem;
    if (rem <= 6) return 3;
    return 4;
}

// Forward pass: stages s0 .. s0+R-1 over np polys of 2^logN words each.
// When last, outputs are made canonical.
template <int R, int NT, typename W>
__device__ __forceinline__ void fwd_pass(W *s, int np, int logN, int s0, bool last, const Arith<W> &A,
                                         const Twiddles<W> &tw) {
    const int N = 1 << logN;
    const int lS = logN - s0 - R;  // log2 of the cluster stride
    const int S = 1 << lS;
    const int per_poly = N >> R;
    const int clusters = np * per_poly;
    for (int cg = threadIdx.x; cg < clusters; cg += NT) {
        int p = cg / per_poly, c = cg - p * per_poly;
        int blk = c >> lS, off = c & (S - 1);
        W *base = s + p * N + (blk << (lS + R)) + off;
        W x[1 << R];
#pragma unroll
        for (int e = 0; e < (1 << R); e++) x[e] = base[e << lS];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int h = 1 << (R - 1 - k);
            const int st = s0 + k;
#pragma unroll
            for (int e = 0; e < (1 << R); e++) {
                if (e & h) continue;
                int ti = (1 << st) + (blk << k) + (e >> (R - k));
                W v = A.mul_shoup(x[e + h], tw.fw[ti], tw.fwp[ti]);
                W u = x[e];
                x[e] = u + v;
                x[e + h] = u - v + A.Q2;
            }
        }
        if (last) {
#pragma unroll
            for (int e = 0; e < (1 << R); e++) x[e] = Arith<W>::red(A.mul_shoup(x[e], (W)1, tw.onep), A.Q);
        }
#pragma unroll
        for (int e = 0; e < (1 << R); e++) base[e << lS] = x[e];
    }
}

// Inverse pass: GS stages u0 .. u0+R-1 (half-size 2^u).  Inputs and outputs in
// [0, 2Q).  When last, the words are scaled by N^-1 and made canonical.
template <int R, int NT, typename W>
__device__ __forceinline__ void inv_pass(W *s, int np, int logN, int u0, bool last, const Arith<W> &A,
                                         const Twiddles<W> &tw) {
    const int N = 1 << logN;
    const int S = 1 << u0;
    const int per_poly = N >> R;
    const int clusters = np * per_poly;
    for (int cg = threadIdx.x; cg < clusters; cg += NT) {
        int p = cg / per_poly, c = cg - p * per_poly;
        int blk = c >> u0, off = c & (S - 1);
        W *base = s + p * N + (blk << (u0 + R)) + off;
        W x[1 << R];
#pragma unroll
        for (int e = 0; e < (1 << R); e++) x[e] = base[e << u0];
#pragma unroll
        for (int k = 0; k < R; k++) {
            const int h = 1 << k;
#pragma unroll
            for (int e = 0; e < (1 << R); e++) {
                if (e & h) continue;
                int ti = (N >> (u0 + k + 1)) + (blk << (R - k - 1)) + (e >> (k + 1));
                W u = x[e], v = x[e + h];
                x[e] = Arith<W>::red(u + v, A.Q2);
                x[e + h] = A.mul_shoup(u - v + A.Q2, tw.iw[ti], tw.iwp[ti]);
            }
        }
        if (last) {
#pragma unroll
            for (int e = 0; e < (1 << R); e++) x[e] = Arith<W>::red(A.mul_shoup(x[e], tw.iN, tw.iNp), A.Q);
        }
#pragma unroll
        for (int e = 0; e < (1 << R); e++) base[e << u0] = x[e];
    }
}

// Full forward transform of np polys in LDS (inputs below 2^w / (2 logN + 1)
// ... canonical inputs always qualify); output canonical in [0, Q).
template <int NT, typename W>
__device__ void ntt_forward(W *s, int np, int logN, const Arith<W> &A, const Twiddles<W> &tw) {
    int st = 0;
    while (st < logN) {
        const int r = pass_stages(logN - st);
        const bool last = st + r == logN;
        switch (r) {
            case 4: fwd_pass<4, NT>(s, np, logN, st, last, A, tw); break;
            case 3: fwd_pass<3, NT>(s, np, logN, st, last, A, tw); break;
            case 2: fwd_pass<2, NT>(s, np, logN, st, last, A, tw); break;
            default: fwd_pass<1, NT>(s, np, logN, st, last, A, tw); break;
        }
        st += r;
        __syncthreads();
    }
}

// Full inverse transform of np polys in LDS (inputs in [0, 2Q)); output
// canonical in [0, Q).
template <int NT, typename W>
__device__ void ntt_inverse(W *s, int np, int logN, const Arith<W> &A, const Twiddles<W> &tw) {
    int st = 0;
    while (st < logN) {
        const int r = pass_stages(logN - st);
        const bool last = st + r == logN;
        switch (r) {
            case 4: inv_pass<4, NT>(s, np, logN, st, last, A, tw); break;
            case 3: inv_pass<3, NT>(s, np, logN, st, last, A, tw); break;
            case 2: inv_pass<2, NT>(s, np, logN, st, last, A, tw); break;
            default: inv_pass<1, NT>(s, np, logN, st, last, A, tw); break;
        }
        st += r;
        __syncthreads();
    }
}

}  // namespace fheacc
