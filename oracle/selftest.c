/*
 * selftest.c -- drives the CPU oracle end to end under AddressSanitizer and
 * UndefinedBehaviorSanitizer (oracle/Makefile, target selftest-asan).
 *
 * TEST INFRASTRUCTURE ONLY.  Every routine of fhe_oracle.c that the parity
 * tests use runs here on small parameter sets, so an out-of-bounds access,
 * a use after free, a leak or undefined integer behaviour in the oracle
 * shows up as a sanitizer report.  The checks are the invariants the
 * routines must keep: iNTT(NTT(x)) = x, decrypted gate outputs equal the
 * truth table in every evaluation order, reruns give the same words.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fhe_oracle.h"

static int failures = 0;
#define CHECK(cond, ...)                           \
    do {                                           \
        if (!(cond)) {                             \
            fprintf(stderr, "FAIL: " __VA_ARGS__); \
            fprintf(stderr, "\n");                 \
            failures++;                            \
        }                                          \
    } while (0)

static uint64_t *alloc_words(size_t n) {
    uint64_t *p = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

static void fill(uint64_t *a, size_t n, uint64_t mod, uint64_t seed) {
    for (size_t i = 0; i < n; i++) a[i] = orc_rng_word(seed, 99, i) % mod;
}

static int truth(int gate, int a, int b) {
    switch (gate) {
        case 0: return a | b;
        case 1: return a & b;
        case 2: return !(a | b);
        case 3: return !(a & b);
        case 4: return a ^ b;
        default: return !(a ^ b);
    }
}

static void run_set(const char *name, int n, uint64_t q, int N, uint64_t Q, int nbit, uint32_t BGbits, uint64_t Br,
                    uint64_t Bks, uint32_t Bks_R_bits) {
    orc_params p;
    if (orc_params_init(&p, n, q, N, Q, nbit, BGbits, Br, Bks, Bks_R_bits, 3.19, 18) != 0) {
        CHECK(0, "%s: params", name);
        return;
    }
    const size_t rl = 2 * (size_t)N, lw = (size_t)n + 1;
    int logN = 0;
    while ((1 << logN) < N) logN++;
    const size_t rw = (size_t)p.digitG * 4 * N;
    /* transforms */
    uint64_t *x = alloc_words(4 * (size_t)N), *y = alloc_words(4 * (size_t)N);
    fill(x, 4 * (size_t)N, p.Q, 1);
    memcpy(y, x, 4 * (size_t)N * 8);
    orc_ntt_batch(&p, y, 4);
    orc_intt_batch(&p, y, 4);
    CHECK(memcmp(x, y, 4 * (size_t)N * 8) == 0, "%s: NTT round trip", name);
    /* external product on random words */
    uint64_t *rgsw = alloc_words(rw), *out = alloc_words(rl), *out2 = alloc_words(rl);
    fill(rgsw, rw, p.Q, 2);
    orc_rlwe_mult_rgsw(&p, x, rgsw, out);
    orc_rlwe_mult_rgsw_batch(&p, x, rgsw, 0, out2, 1);
    CHECK(memcmp(out, out2, rl * 8) == 0, "%s: external product batch", name);
    /* keys */
    uint64_t *s = alloc_words((size_t)n), *z = alloc_words((size_t)N);
    orc_lwe_secret(&p, 5, 1, s);
    orc_rlwe_secret(&p, 5, z);
    const size_t bsk_words = (size_t)p.Br * p.digitR * n * rw;
    const size_t ksk_words = (size_t)p.Bks * p.digitKS * N * lw;
    uint64_t *bsk = alloc_words(bsk_words), *ksk = alloc_words(ksk_words);
    orc_bsk_gen(&p, s, z, 6, bsk, 2);
    orc_ksk_gen(&p, z, s, 6, ksk, 2);
    /* every gate on every input pair, in every evaluation order */
    uint64_t *c1 = alloc_words(4 * lw), *c2 = alloc_words(4 * lw), *res = alloc_words(4 * lw);
    uint64_t *acc = alloc_words(4 * rl), *ks = alloc_words(4 * lw);
    for (int m = 0; m < 4; m++) {
        orc_lwe_encrypt(&p, s, (uint64_t)(m >> 1), 7, (uint64_t)m, c1 + m * lw);
        orc_lwe_encrypt(&p, s, (uint64_t)(m & 1), 7, (uint64_t)(100 + m), c2 + m * lw);
    }
    for (int mode = ORC_EVAL_SOFTWARE; mode <= ORC_EVAL_FPGA_XN; mode++) {
        p.eval_mode = mode;
        for (int gate = 0; gate < 6; gate++) {
            int32_t gates[4] = {gate, gate, gate, gate};
            int rc = orc_eval_bootstrap_batch(&p, c1, c2, gates, 4, bsk, ksk, res, acc, ks, 2);
            CHECK(rc == 0, "%s: bootstrap rc", name);
            for (int m = 0; m < 4; m++)
                CHECK((int)orc_lwe_decrypt(&p, s, res + m * lw) == truth(gate, m >> 1, m & 1),
                      "%s: mode %d gate %d inputs %d", name, mode, gate, m);
        }
    }
    p.eval_mode = ORC_EVAL_SOFTWARE;
    /* identical operands are refused, not evaluated */
    CHECK(orc_eval_bootstrap(&p, c1, c1, 3, bsk, ksk, res, NULL, NULL) == -1, "%s: identical operands", name);
    /* RLWE expansion, CMUX, leveled blind rotations */
    uint64_t *subs = alloc_words((size_t)logN * p.digitKS_R * rl), *exp = alloc_words((size_t)N * rl);
    uint64_t *exp2 = alloc_words((size_t)N * rl);
    orc_subs_key_gen(&p, z, 8, subs);
    orc_expand_rlwe(&p, x, subs, exp);
    orc_expand_rlwe(&p, x, subs, exp2);
    CHECK(memcmp(exp, exp2, (size_t)N * rl * 8) == 0, "%s: expansion rerun", name);
    orc_cmux(&p, x, x + rl, rgsw, out);
    const int steps = logN;
    uint64_t *rg = alloc_words((size_t)steps * rw);
    fill(rg, (size_t)steps * rw, p.Q, 9);
    orc_blind_rotate_rlwe(&p, x, 1, rg, steps, out);
    orc_blind_rotate_rlwe(&p, x, 0, rg, steps, out2);
    int32_t a[3] = {1, -3, N + 5};
    orc_blind_rotate_full(&p, x, -7, a, 3, rg, out);
    /* scale_round's edge: the top of [0, Q) rounds up to q itself */
    CHECK(orc_scale_round(p.Q - 1, p.q, p.Q) == p.q, "%s: scale_round edge", name);
    free(x); free(y); free(rgsw); free(out); free(out2); free(s); free(z); free(bsk); free(ksk);
    free(c1); free(c2); free(res); free(acc); free(ks); free(subs); free(exp); free(exp2); free(rg);
    orc_params_free(&p);
}

int main(void) {
    /* 27-bit Q (32-bit device words), Br = 23, Bks = 25 */
    run_set("n16_N256_q27", 16, 512, 256, 0, 27, 9, 23, 25, 7);
    /* 54-bit Q (64-bit device words), Br = 4 (a quarter of the steps at residue 0) */
    run_set("n6_N128_q54", 6, 256, 128, 0x3FFFFFFFFED001ULL, 54, 18, 4, 32, 9);
    if (failures) {
        fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    printf("ORACLE SELFTEST OK\n");
    return 0;
}
