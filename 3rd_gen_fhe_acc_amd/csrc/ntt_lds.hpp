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
// length-1024 transform is three LDS round trips (4+3+3 stages).  A caller
// whose threads already hold a cluster in registers can run its stages there
// (fwd_regs, inv_regs, inv_tail_regs) and have the LDS passes start or end at
// another stage (ntt_forward's ST0 and END, ntt_inverse's END).
//
// LDS layout: logical word i of a run of polynomials lives at i + i/8.  One
// pad word per eight spreads the strided cluster accesses (strides 1..8
// words) over distinct banks; without it a stride-8 pass hits 4 of 32 banks.
//
// Forward butterflies are fully lazy: no reduction between stages.  A value
// entering stage s is below (2s+1)Q, so after log2 N stages every word is
// below (2 log2 N + 1) Q; 64-bit words, whose products land in [0, 5Q), stay
// below (5 log2 N + 1) Q.  When the 32-bit bound does not fit but 4Q
// does (28- to 30-bit Q), the context selects Harvey's butterflies instead
// (Arith::harvey): the added input is first brought below 2Q, so every word
// stays below 4Q, and the last pass leaves its words below 2Q.  The last
// pass makes the words canonical when asked.
#pragma once
#include "device_checks.hpp"
#include "modarith.hpp"

namespace fheacc {

template <typename W>
struct Twiddles {
    const W *fw, *fwp;   // rou, Shoup companions   [N]
    const W *iw, *iwp;   // irou, Shoup companions  [N]
    W iN, iNp;
    W iNw, iNwp;         // N^-1 irou[1] (the inverse's last twiddle) and its companion
    W onep;              // Shoup companion of 1: floor(2^w / Q)
};

// physical LDS index of logical word i
__host__ __device__ __forceinline__ constexpr int pad(int i) { return i + (i >> 3); }
// LDS words occupied by `words` logical words
__host__ __device__ __forceinline__ constexpr int padded(int words) { return words + (words >> 3) + 1; }

// stages per pass for `rem` remaining stages, at most MAXR (4 by default;
// the standalone 64-bit NTT uses 3, where 16-word clusters and 15 twiddle
// pairs of 64-bit words cost more registers than the saved LDS round trip
// is worth): MAXR-stage passes, never a lone 1-stage pass unless the
// transform has a single stage
template <int MAXR>
__host__ __device__ constexpr int pass_stages(int rem) {
    return rem <= MAXR ? rem : (rem <= MAXR + 2 ? MAXR - 1 : MAXR);
}
// the last pass of that schedule over `rem` stages; the inverse transform
// runs the forward schedule mirrored (3,3,4 for 1024 points), so that its
// stride-1 first pass has 8-word clusters, which the padding keeps
// conflict-free
template <int MAXR>
__host__ __device__ constexpr int last_pass_stages(int rem) {
    return rem <= pass_stages<MAXR>(rem) ? rem : last_pass_stages<MAXR>(rem - pass_stages<MAXR>(rem));
}
// LDS passes of that schedule over `rem` stages
template <int MAXR>
__host__ __device__ constexpr int pass_count(int rem) {
    return rem <= 0 ? 0 : 1 + pass_count<MAXR>(rem - pass_stages<MAXR>(rem));
}


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

// Twiddles of one cluster, loaded once and reused for every poly the thread
// handles at the same in-poly cluster index (they depend on blk only).
template <int R, typename W>
struct ClusterTwiddles {
    W w[(1 << R) - 1], wp[(1 << R) - 1];
};

// Forward twiddle j of stage k (j < 2^k): rou[2^st + blk 2^k + j]
template <int S0, int R, typename W>
__device__ __forceinline__ ClusterTwiddles<R, W> fwd_twiddles(int blk, const Twiddles<W> &tw) {
    ClusterTwiddles<R, W> t;
#pragma unroll
    for (int k = 0; k < R; k++)
#pragma unroll
        for (int j = 0; j < (1 << k); j++) {
            const int ti = (1 << (S0 + k)) + (blk << k) + j;
            t.w[(1 << k) - 1 + j] = tw.fw[ti];
            t.wp[(1 << k) - 1 + j] = tw.fwp[ti];
        }
    return t;
}

// Inverse twiddle j of stage k (j < 2^(R-k-1)): irou[N/2^(U0+k+1) + blk 2^(R-k-1) + j]
template <int LOGN, int U0, int R, typename W>
__device__ __forceinline__ ClusterTwiddles<R, W> inv_twiddles(int blk, const Twiddles<W> &tw) {
    constexpr int N = 1 << LOGN;
    ClusterTwiddles<R, W> t;
    int o = 0;
#pragma unroll
    for (int k = 0; k < R; k++)
#pragma unroll
        for (int j = 0; j < (1 << (R - k - 1)); j++) {
            const int ti = (N >> (U0 + k + 1)) + (blk << (R - k - 1)) + j;
            t.w[o] = tw.iw[ti];
            t.wp[o] = tw.iwp[ti];
            o++;
        }
    return t;
}

// A cluster of 2^R words in registers.  The stage functions below take and
// return it by value: passed by reference, the 64-bit transforms compile to
// some 250 more register moves.
template <int R, typename W>
struct Words {
    W x[1 << R];
};

// R forward stages on a cluster held in registers (x[e] is word b0 + e S).
// HV: Harvey's butterflies (words below 4Q); FINAL: the transform's last pass
template <int R, bool HV, bool FINAL, typename W>
__device__ __forceinline__ Words<R, W> fwd_regs(Words<R, W> cl, bool last, const Arith<W> &A,
                                                const Twiddles<W> &tw, const ClusterTwiddles<R, W> &t) {
    W (&x)[1 << R] = cl.x;
#pragma unroll
    for (int k = 0; k < R; k++) {
        const int h = 1 << (R - 1 - k);
#pragma unroll
        for (int e = 0; e < (1 << R); e++) {
            if (e & h) continue;
            const int j = (1 << k) - 1 + (e >> (R - k));
            if constexpr (HV) x[e] = Arith<W>::red(x[e], A.Q2);
            A.ct_butterfly(x[e], x[e + h], t.w[j], t.wp[j]);
        }
    }
    if (last) {
#pragma unroll
        for (int e = 0; e < (1 << R); e++) x[e] = Arith<W>::red(A.mul_shoup(x[e], (W)1, tw.onep), A.Q);
    } else if constexpr (HV && FINAL) {
#pragma unroll
        for (int e = 0; e < (1 << R); e++) x[e] = Arith<W>::red(x[e], A.Q2);
    }
    return cl;
}

template <int S, int R, bool HV, bool FINAL, typename W>
__device__ __forceinline__ void fwd_cluster(W *s, int b0, bool last, const Arith<W> &A, const Twiddles<W> &tw,
                                            const ClusterTwiddles<R, W> &t) {
    W *base = s + pad(b0);
    Words<R, W> v;
#pragma unroll
    for (int e = 0; e < (1 << R); e++) {
        FHEACC_CHECK_LDS(&base[ClusterOffsets<S, R>::at(b0, e)]);
        v.x[e] = base[ClusterOffsets<S, R>::at(b0, e)];
    }
    v = fwd_regs<R, HV, FINAL>(v, last, A, tw, t);
#pragma unroll
    for (int e = 0; e < (1 << R); e++) base[ClusterOffsets<S, R>::at(b0, e)] = v.x[e];
}

// Lazy inverse passes leave the butterfly sums unreduced.  Within a pass the
// bound of every register is known at compile time: with inputs below 2Q,
// element e after stage k is below 2^(1 + z) Q, z = lazy_run(e, k), the run
// of zero bits of e from bit k down (a set bit means the word came out of a
// Shoup product, below 2Q; a zero bit means a sum, twice the bound of its
// operands, which share the bits below k and so their bound).
__host__ __device__ constexpr int lazy_run(int e, int k) {
    int z = 0;
    for (int b = k; b >= 0 && !((e >> b) & 1); b--) z++;
    return z;
}

// LZ = 1: the lazy form (needs 2^(R+1) Q within the word); LZ = 0: every sum
// is reduced to [0, 2Q) at once; LZ = 2 (64-bit words): the lazy form with
// approximate-quotient products, below.
template <int R, bool LAST, int LZ, typename W>
__device__ __forceinline__ Words<R, W> inv_regs(Words<R, W> cl, const Arith<W> &A, const Twiddles<W> &tw,
                                                const ClusterTwiddles<R, W> &t) {
    W (&x)[1 << R] = cl.x;
    int o = 0;
    if constexpr (LZ == 2 && sizeof(W) == 8) {
        // Products by mul_shoup_lazy land in [0, 5Q) (two 32-bit high
        // products and one multiply-add for the quotient, not a full 64x64
        // high product).  Inputs are below 5Q, so the bounds of the lazy form
        // hold with 5Q in place of 2Q: element e after stage k is below
        // 5Q 2^z, z = lazy_run(e, k), and so below 5Q 2^R <= 80Q, which the
        // context's parameter check keeps within 64 bits.  The pass hands the next one
        // words below 5Q; the transform's last stage multiplies exactly and
        // ends canonical.
        constexpr int KS = LAST ? R - 1 : R;
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const int h = 1 << k;
#pragma unroll
            for (int e = 0; e < (1 << R); e++) {
                if (e & h) continue;
                const int j = o + (e >> (k + 1));
                const W c = A.Q5 << lazy_run(e, k - 1);  // u, v < c
                const W u = x[e], v = x[e + h];
                x[e] = u + v;
                x[e + h] = A.mul_shoup_lazy(u - v + c, t.w[j], t.wp[j]);
            }
            o += 1 << (R - k - 1);
        }
        if constexpr (LAST) {
            constexpr int h = 1 << (R - 1);
#pragma unroll
            for (int e = 0; e < h; e++) {
                const W c = A.Q5 << lazy_run(e, R - 2);
                const W u = x[e], v = x[e + h];
                x[e] = Arith<W>::red(A.mul_shoup(u + v, tw.iN, tw.iNp), A.Q);
                x[e + h] = Arith<W>::red(A.mul_shoup(u - v + c, tw.iNw, tw.iNwp), A.Q);
            }
        } else {
            // back below 5Q for the next pass: one conditional subtraction
            // from [0, 10Q), a lazy product by 1 above that
#pragma unroll
            for (int e = 0; e < (1 << R); e++) {
                const int z = lazy_run(e, R - 1);
                if (z == 1) x[e] = Arith<W>::red(x[e], A.Q5);
                else if (z > 1) x[e] = A.mul_shoup_lazy(x[e], (W)1, tw.onep);
            }
        }
        return cl;
    }
    if constexpr (LZ == 1) {
        constexpr int KS = LAST ? R - 1 : R;
#pragma unroll
        for (int k = 0; k < KS; k++) {
            const int h = 1 << k;
#pragma unroll
            for (int e = 0; e < (1 << R); e++) {
                if (e & h) continue;
                const int j = o + (e >> (k + 1));
                const W c = A.Q << (1 + lazy_run(e, k - 1));  // u, v < c
                const W u = x[e], v = x[e + h];
                x[e] = u + v;
                x[e + h] = A.mul_shoup(u - v + c, t.w[j], t.wp[j]);
            }
            o += 1 << (R - k - 1);
        }
        if constexpr (LAST) {
            constexpr int h = 1 << (R - 1);
#pragma unroll
            for (int e = 0; e < h; e++) {
                const W c = A.Q << (1 + lazy_run(e, R - 2));
                const W u = x[e], v = x[e + h];
                x[e] = Arith<W>::red(A.mul_shoup(u + v, tw.iN, tw.iNp), A.Q);
                x[e + h] = Arith<W>::red(A.mul_shoup(u - v + c, tw.iNw, tw.iNwp), A.Q);
            }
        } else {
            // back below 2Q for the next pass
#pragma unroll
            for (int e = 0; e < (1 << R); e++) {
                const int z = lazy_run(e, R - 1);
                if (z == 1) x[e] = Arith<W>::red(x[e], A.Q2);
                else if (z > 1) x[e] = A.mul_shoup(x[e], (W)1, tw.onep);
            }
        }
        return cl;
    }
    // the transform's last stage (one twiddle, irou[1]) carries the N^-1
    // scaling: its sum goes straight into the product by N^-1 (a Shoup
    // product takes any 32- or 64-bit input), its difference into one
    // product by N^-1 irou[1]
    constexpr int KS = LAST ? R - 1 : R;
#pragma unroll
    for (int k = 0; k < KS; k++) {
        const int h = 1 << k;
#pragma unroll
        for (int e = 0; e < (1 << R); e++) {
            if (e & h) continue;
            const int j = o + (e >> (k + 1));
            const W u = x[e], v = x[e + h];
            x[e] = Arith<W>::red(u + v, A.Q2);
            x[e + h] = A.mul_shoup(u - v + A.Q2, t.w[j], t.wp[j]);
        }
        o += 1 << (R - k - 1);
    }
    if constexpr (LAST) {
        constexpr int h = 1 << (R - 1);
#pragma unroll
        for (int e = 0; e < h; e++) {
            const W u = x[e], v = x[e + h];
            x[e] = Arith<W>::red(A.mul_shoup(u + v, tw.iN, tw.iNp), A.Q);
            x[e + h] = Arith<W>::red(A.mul_shoup(u - v + A.Q2, tw.iNw, tw.iNwp), A.Q);
        }
    }
    return cl;
}

// The transform's last F stages (with the N^-1 scaling) on words a thread
// holds in registers: x[e] is word b0 + e N / 2^F.  Inputs below 2Q (5Q on
// 64-bit words), outputs canonical.
template <int F, typename W>
__device__ __forceinline__ Words<F, W> inv_tail_regs(Words<F, W> x, const Arith<W> &A, const Twiddles<W> &tw,
                                                     const ClusterTwiddles<F, W> &t) {
    if constexpr (sizeof(W) == 8) return inv_regs<F, true, 2>(x, A, tw, t);
    else if (F <= A.lzr) return inv_regs<F, true, 1>(x, A, tw, t);
    else return inv_regs<F, true, 0>(x, A, tw, t);
}

template <int S, int R, bool LAST, int LZ, typename W>
__device__ __forceinline__ void inv_cluster(W *s, int b0, const Arith<W> &A, const Twiddles<W> &tw,
                                            const ClusterTwiddles<R, W> &t) {
    W *base = s + pad(b0);
    Words<R, W> v;
#pragma unroll
    for (int e = 0; e < (1 << R); e++) {
        FHEACC_CHECK_LDS(&base[ClusterOffsets<S, R>::at(b0, e)]);
        v.x[e] = base[ClusterOffsets<S, R>::at(b0, e)];
    }
    v = inv_regs<R, LAST, LZ>(v, A, tw, t);
#pragma unroll
    for (int e = 0; e < (1 << R); e++) base[ClusterOffsets<S, R>::at(b0, e)] = v.x[e];
}

// Forward pass: stages S0 .. S0+R-1 over np polys of 2^LOGN words each.
// When last, outputs are made canonical.  When NT is a multiple of the
// clusters per poly, a thread meets the same in-poly cluster in every poly
// it handles, and loads that cluster's twiddles once.
template <int LOGN, int S0, int R, int NT, bool HV, int NP, typename W>
__device__ __forceinline__ void fwd_pass(W *s, int np, bool last, const Arith<W> &A, const Twiddles<W> &tw,
                                         int tid) {
    constexpr int N = 1 << LOGN;
    constexpr int lS = LOGN - S0 - R;  // log2 of the cluster stride
    constexpr int S = 1 << lS;
    constexpr int per_poly = N >> R;
    constexpr bool FINAL = S0 + R == LOGN;
    if constexpr (NT % per_poly == 0) {
        const int c = tid % per_poly;
        const int blk = c >> lS, off = c & (S - 1);
        const ClusterTwiddles<R, W> t = fwd_twiddles<S0, R>(blk, tw);
        // a trip count every thread shares (np is uniform), so the loop is
        // a scalar one; the guard only masks lanes when G does not divide np
        constexpr int G = NT / per_poly;
        constexpr int PSW = N + N / 8;  // padded words per poly (pad(p N + i) = p PSW + pad(i))
        const int p0 = tid / per_poly;
        if constexpr (NP > 0) {
            // a poly count known at compile time: unrolled, constant offsets
#pragma unroll
            for (int it = 0; it < (NP + G - 1) / G; it++)
                if (NP % G == 0 || p0 + it * G < NP)
                    fwd_cluster<S, R, HV, FINAL>(s + (p0 + it * G) * PSW, (blk << (lS + R)) + off, last, A, tw, t);
        } else {
            const int iters = (np + G - 1) / G;
            W *sp = s + p0 * PSW;
            for (int it = 0; it < iters; it++, sp += G * PSW)
                if (p0 + it * G < np) fwd_cluster<S, R, HV, FINAL>(sp, (blk << (lS + R)) + off, last, A, tw, t);
        }
    } else {
        const int clusters = np * per_poly;
        for (int cg = tid; cg < clusters; cg += NT) {
            const int p = cg / per_poly, c = cg - p * per_poly;
            const int blk = c >> lS, off = c & (S - 1);
            fwd_cluster<S, R, HV, FINAL>(s, p * N + (blk << (lS + R)) + off, last, A, tw,
                                         fwd_twiddles<S0, R>(blk, tw));
        }
    }
}

// Inverse pass: GS stages U0 .. U0+R-1 (half-size 2^u).  Inputs and outputs in
// [0, 2Q).  When last, the words are scaled by N^-1 and made canonical.
template <int LOGN, int U0, int R, int NT, int LZ, int NP, typename W>
__device__ __forceinline__ void inv_pass(W *s, int np, const Arith<W> &A, const Twiddles<W> &tw, int tid) {
    constexpr int N = 1 << LOGN;
    constexpr int S = 1 << U0;
    constexpr int per_poly = N >> R;
    constexpr bool last = U0 + R == LOGN;
    if constexpr (NT % per_poly == 0) {
        const int c = tid % per_poly;
        const int blk = c >> U0, off = c & (S - 1);
        const ClusterTwiddles<R, W> t = inv_twiddles<LOGN, U0, R>(blk, tw);
        constexpr int G = NT / per_poly;
        constexpr int PSW = N + N / 8;
        const int p0 = tid / per_poly;
        if constexpr (NP > 0) {
#pragma unroll
            for (int it = 0; it < (NP + G - 1) / G; it++)
                if (NP % G == 0 || p0 + it * G < NP)
                    inv_cluster<S, R, last, LZ>(s + (p0 + it * G) * PSW, (blk << (U0 + R)) + off, A, tw, t);
        } else {
            const int iters = (np + G - 1) / G;
            W *sp = s + p0 * PSW;
            for (int it = 0; it < iters; it++, sp += G * PSW)
                if (p0 + it * G < np) inv_cluster<S, R, last, LZ>(sp, (blk << (U0 + R)) + off, A, tw, t);
        }
    } else {
        const int clusters = np * per_poly;
        for (int cg = tid; cg < clusters; cg += NT) {
            const int p = cg / per_poly, c = cg - p * per_poly;
            const int blk = c >> U0, off = c & (S - 1);
            inv_cluster<S, R, last, LZ>(s, p * N + (blk << (U0 + R)) + off, A, tw,
                                        inv_twiddles<LOGN, U0, R>(blk, tw));
        }
    }
}

template <int LOGN, int ST, int NT, int MAXR, bool HV, int NP, int END = LOGN, typename W>
__device__ __forceinline__ void fwd_passes(W *s, int np, bool canonical, const Arith<W> &A, const Twiddles<W> &tw,
                                           int tid) {
    if constexpr (ST < END) {
        constexpr int R = pass_stages<MAXR>(END - ST);
        fwd_pass<LOGN, ST, R, NT, HV, NP>(s, np, canonical && ST + R == LOGN, A, tw, tid);
        __syncthreads();
        fwd_passes<LOGN, ST + R, NT, MAXR, HV, NP, END>(s, np, canonical, A, tw, tid);
    }
}

// Stages ST .. END-1.  LP: every pass in the lazy-product form (its 5Q bound
// between passes holds only when all passes use it); otherwise each pass is
// lazy or strict by its own stage count.
template <int LOGN, int ST, int END, int NT, int MAXR, bool LP, int NP, typename W>
__device__ __forceinline__ void inv_passes(W *s, int np, const Arith<W> &A, const Twiddles<W> &tw, int tid) {
    if constexpr (ST < END) {
        constexpr int R = last_pass_stages<MAXR>(END - ST);
        if constexpr (LP) inv_pass<LOGN, ST, R, NT, 2, NP>(s, np, A, tw, tid);
        else if (R <= A.lzr) inv_pass<LOGN, ST, R, NT, 1, NP>(s, np, A, tw, tid);
        else inv_pass<LOGN, ST, R, NT, 0, NP>(s, np, A, tw, tid);
        __syncthreads();
        inv_passes<LOGN, ST + R, END, NT, MAXR, LP, NP>(s, np, A, tw, tid);
    }
}

// Forward transform of np padded polys in LDS (canonical inputs).  Output
// canonical in [0, Q), or, when canonical is false, below (2 logN + 1) Q
// (below 2Q with Harvey's butterflies, below (5 logN + 1) Q on 64-bit words)
// and congruent to the canonical word (for consumers that reduce anyway).
// NP > 0: np is the compile-time constant NP (the passes unroll over the polys).
// ST0 > 0: stages 0 .. ST0-1 have already run (in registers, fwd_regs, with
// the same butterflies as this transform's: Harvey's when A.harvey).
// END < LOGN: only stages up to END-1; the caller runs the rest in registers.
template <int LOGN, int NT, int MAXR = 4, int NP = 0, int ST0 = 0, int END = LOGN, typename W>
__device__ void ntt_forward(W *s, int np, const Arith<W> &A, const Twiddles<W> &tw, bool canonical = true,
                            int tid = -1) {
    const int t = tid < 0 ? (int)threadIdx.x : tid;
    if constexpr (sizeof(W) == 4) {
        if (A.harvey) {
            fwd_passes<LOGN, ST0, NT, MAXR, true, NP, END>(s, np, canonical, A, tw, t);
            return;
        }
    }
    fwd_passes<LOGN, ST0, NT, MAXR, false, NP, END>(s, np, canonical, A, tw, t);
}

// Inverse transform of np padded polys in LDS (inputs in [0, 2Q)); output
// canonical in [0, Q).  END < LOGN: only stages 0 .. END-1, leaving words
// below 2Q (5Q on 64-bit words) for inv_tail_regs to finish in registers.
template <int LOGN, int NT, int MAXR = 4, int NP = 0, int END = LOGN, typename W>
__device__ void ntt_inverse(W *s, int np, const Arith<W> &A, const Twiddles<W> &tw, int tid = -1) {
    const int t = tid < 0 ? (int)threadIdx.x : tid;
    // 64-bit words: every pass takes lazy products.  Words stay below
    // 5Q 2^R <= 80Q, which the context's parameter check keeps within 64 bits.
    static_assert(sizeof(W) == 4 || MAXR <= 4, "lazy-product inverse passes need 80Q < 2^64");
    inv_passes<LOGN, 0, END, NT, MAXR, sizeof(W) == 8, NP>(s, np, A, tw, t);
}

}  // namespace fheacc
