// params.cpp -- host-side parameter derivation and the seeded sampler.
//
// Q, N^-1, psi and the twiddle tables are derived exactly as the reference
// apps derive them: Q = previous_prime(first_prime(nbit, N), N)
// (apps/fhe_wo_fpga.cpp:89, nbt.cpp:209-266), psi = the smallest primitive
// 2N-th root of unity (nbt.cpp:300-346), tables psi^i / psi^-i at bitrev(i)
// (params.cpp:149-175), digit counts ceil(log2 M / log2 B) in double
// (fhe_wo_fpga.cpp:122,128,133,138), gate constants {5,7,1,3,5,1}*(q>>3)
// (fhe_wo_fpga.cpp:141-146).
#include <math.h>

#include <algorithm>
#include <string>

#include "fheacc_internal.hpp"

namespace fheacc {

typedef unsigned __int128 u128;

static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
    uint64_t r = 1 % m;
    b %= m;
    while (e) {
        if (e & 1) r = mulmod(r, b, m);
        b = mulmod(b, b, m);
        e >>= 1;
    }
    return r;
}

// Deterministic Miller-Rabin, exact for 64-bit inputs.
static bool is_prime(uint64_t x) {
    if (x < 2) return false;
    for (uint64_t p : {2ULL, 3ULL, 5ULL, 7ULL, 11ULL, 13ULL, 17ULL, 19ULL, 23ULL, 29ULL, 31ULL, 37ULL}) {
        if (x == p) return true;
        if (x % p == 0) return false;
    }
    uint64_t d = x - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; s++; }
    for (uint64_t a : {2ULL, 325ULL, 9375ULL, 28178ULL, 450775ULL, 9780504ULL, 1795265022ULL}) {
        a %= x;
        if (!a) continue;
        uint64_t y = powmod(a, d, x);
        if (y == 1 || y == x - 1) continue;
        bool comp = true;
        for (int r = 1; r < s && comp; r++) {
            y = mulmod(y, y, x);
            if (y == x - 1) comp = false;
        }
        if (comp) return false;
    }
    return true;
}

static uint64_t prime_from_nbit(int nbit, int N) {
    uint64_t m = 2ULL * N, base = 1ULL << nbit;
    uint64_t c = base + m - (base % m) + 1;  // first candidate = 1 mod 2N above 2^nbit
    while (!is_prime(c)) c += m;
    c -= m;                                  // then the previous such prime
    while (!is_prime(c)) c -= m;
    return c;
}

static uint64_t bitrev(uint64_t x, int bits) {
    uint64_t r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1ULL) << (bits - 1 - i);
    return r;
}

static uint64_t smallest_root(int N, uint64_t Q) {
    uint64_t m = 2ULL * N;
    uint64_t r = 0;
    for (uint64_t g = 2; g < Q && !r; g++) {
        uint64_t c = powmod(g, (Q - 1) / m, Q);
        if (powmod(c, (uint64_t)N, Q) == Q - 1) r = c;
    }
    if (!r) return 0;
    uint64_t r2 = mulmod(r, r, Q), cur = r, best = r;
    for (uint64_t k = 1; k < m; k += 2) {
        if (cur < best && cur != 1) best = cur;
        cur = mulmod(cur, r2, Q);
    }
    return best;
}

static uint32_t digits(double lg, double base) { return (uint32_t)ceil(lg / log2(base)); }

int derive_params(const fheacc_params *in, HostParams *p, std::string *err) {
    auto bad = [&](const char *m) { *err = m; return -1; };
    if (!in) return bad("null parameter block");
    int N = in->N;
    if (N < 16 || N > 2048 || (N & (N - 1))) return bad("N must be a power of two in [16, 2048]");
    if (in->q < 4 || (in->q & (in->q - 1)) || (2ULL * N) % in->q || in->q > 65536)
        return bad("q must be a power of two dividing 2N (and <= 65536)");
    if (in->n < 1 || in->n > 4096) return bad("n must be in [1, 4096]");
    uint64_t Q = in->Q;
    if (Q == 0) {
        if (in->nbit < 12 || in->nbit > 61) return bad("nbit must be in [12, 61]");
        Q = prime_from_nbit(in->nbit, N);
    }
    if (Q < 3 || (Q - 1) % (2ULL * N) || Q >= (1ULL << 62) || !is_prime(Q))
        return bad("Q must be a prime = 1 mod 2N below 2^62");
    if (in->BGbits < 1 || in->BGbits > 32) return bad("BGbits must be in [1, 32]");
    if (in->Br < 2 || in->Br > in->q) return bad("Br must be in [2, q]");
    if (in->Bks < 2 || in->Bks > 256) return bad("Bks must be in [2, 256]");
    if (in->Bks_R_bits < 1 || in->Bks_R_bits > 32) return bad("Bks_R_bits must be in [1, 32]");
    p->n = in->n;
    p->q = in->q;
    p->N = N;
    p->logN = 0;
    while ((1 << p->logN) < N) p->logN++;
    p->Q = Q;
    p->iN = powmod((uint64_t)N, Q - 2, Q);
    p->ROU = smallest_root(N, Q);
    if (!p->ROU) return bad("no primitive 2N-th root of unity");
    p->BGbits = in->BGbits;
    double lgQ = log2((double)Q);
    p->digitG = digits(lgQ, (double)(1ULL << in->BGbits));
    p->Br = in->Br;
    p->digitR = digits(log2((double)in->q), (double)in->Br);
    p->Bks = in->Bks;
    p->digitKS = digits(lgQ, (double)in->Bks);
    p->Bks_R_bits = in->Bks_R_bits;
    p->digitKS_R = digits(lgQ, (double)(1ULL << in->Bks_R_bits));
    if (p->digitG < 1 || p->digitG > 16) return bad("digitG out of range");
    p->E = in->q / 16;
    p->E_rlwe = in->E_rlwe ? in->E_rlwe : 18;
    p->stddev = in->stddev > 0 ? in->stddev : 3.19;
    static const uint64_t gc[6] = {5, 7, 1, 3, 5, 1};
    for (int g = 0; g < 6; g++) p->gate_constant[g] = gc[g] * (in->q >> 3);
    p->rou.assign(N, 0);
    p->irou.assign(N, 0);
    uint64_t ir = powmod(p->ROU, Q - 2, Q), t = 1, it = 1;
    p->rou[0] = p->irou[0] = 1;
    for (int i = 1; i < N; i++) {
        t = mulmod(t, p->ROU, Q);
        it = mulmod(it, ir, Q);
        uint64_t r = bitrev((uint64_t)i, p->logN);
        p->rou[r] = t;
        p->irou[r] = it;
    }
    // Gaussian magnitude table: P(|trunc(Z sigma)| <= k) = erf((k+1)/(sigma sqrt 2))
    p->cdt.clear();
    for (int k = 0; k < 64; k++) {
        double c = erf((double)(k + 1) / (p->stddev * sqrt(2.0)));
        uint64_t v = (c >= 1.0) ? (1ULL << 63) : (uint64_t)ldexp(c, 63);
        p->cdt.push_back(v);
        if (v == (1ULL << 63)) break;
    }
    p->cdt.back() = 1ULL << 63;
    // Lazy forward butterflies leave 32-bit words below (2 log2 N + 1) Q,
    // which must be below 2^32.  Harvey's butterflies keep 32-bit words below
    // 4Q and hand the MAC words below 2Q.  The transformed digits reach the
    // MAC at that bound, and a sum of max(2 digitG, digitKS_R) products with
    // keys below Q must stay below 2^64.  64-bit words take products in
    // [0, 5Q) (Arith<uint64_t>::mul_shoup_lazy): forward words stay below
    // (5 log2 N + 1) Q < 2^64, and MAC sums below 5Q per term < 2^62.
    const u128 terms = (u128)std::max<uint32_t>(2 * p->digitG, p->digitKS_R);
    const u128 two64 = (u128)1 << 64, two32 = (u128)1 << 32;
    if ((u128)(5 * p->logN + 1) * Q >= two64 || terms * 5 * (u128)Q >= ((u128)1 << 62))
        return bad("Q too large for lazy 64-bit butterflies");
    const u128 lazy = (u128)(2 * p->logN + 1) * Q;
    p->harvey = false;
    p->narrow = lazy < two32 && terms * lazy * (Q - 1) < two64;
    if (!p->narrow && 4 * (u128)Q < two32 && terms * 2 * (u128)Q * (Q - 1) < two64) {
        p->narrow = true;
        p->harvey = true;
    }
    // 64-bit inverse passes of up to 4 stages take lazy products (below 5Q)
    // and leave their sums unreduced: every word stays below 80Q
    if (!p->narrow && 80 * (u128)Q >= two64) return bad("Q too large for lazy 64-bit inverse passes");
    return 0;
}

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Counter-based stream: word = mix64(mix64(seed + stream*G) + (ctr+1)*G).
uint64_t rng_word(uint64_t seed, uint64_t stream, uint64_t ctr) {
    const uint64_t G = 0x9E3779B97F4A7C15ULL;
    return mix64(mix64(seed + stream * G) + (ctr + 1) * G);
}

int64_t gauss_sample(const HostParams &p, uint64_t word) {
    uint64_t v = word >> 1;
    int k = 0, len = (int)p.cdt.size();
    while (k < len - 1 && v >= p.cdt[k]) k++;
    if (!k) return 0;
    return (word & 1) ? -(int64_t)k : (int64_t)k;
}

}  // namespace fheacc
