/*
 * fhe_oracle.h -- CPU oracle for the FHEW/TFHE gate-bootstrapping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * software path (FHE_ACC_runtime/src/{nbt,params,LWE,RLWE,app}.cpp) used as the
 * parity reference for the HIP library.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * Arithmetic follows the reference literally: every modular product is an
 * unsigned __int128 '%' (FHE_ACC_runtime/include/nbt.h:59-101).
 *
 * Layouts (all words uint64, canonical in [0, modulus)):
 *   poly            : [N]
 *   RLWE            : [2][N]            (a, then b)
 *   RGSW            : [digitG][2][2][N] (row h, column g; g=0 carries m*BG^h in a,
 *                                        g=1 in b; each an RLWE a,b)
 *   bootstrap key   : [Br][digitR][n] RGSW   (RLWE.cpp:579-606 btkey[r][j][k])
 *   LWE             : [n+1]             (a[0..n-1], then b)
 *   LWE key-switch  : [Bks][digitKS][N] LWE(n) mod Q   (LWE.cpp:55-106)
 *   RLWE key-switch : [digitKS_R] RLWE   (RLWE.cpp:293-322)
 *
 * Parity status per function is recorded in the header of fhe_oracle.c.
 */
#ifndef FHE_ORACLE_H
#define FHE_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum orc_gate { ORC_OR = 0, ORC_AND, ORC_NOR, ORC_NAND, ORC_XOR, ORC_XNOR };

typedef struct orc_params {
    int32_t  n;          /* LWE dimension */
    uint64_t q;          /* LWE modulus (power of two, divides 2N) */
    int32_t  N;          /* ring degree (power of two) */
    uint64_t Q;          /* RLWE prime, Q = 1 mod 2N */
    uint64_t iN;         /* N^-1 mod Q */
    uint64_t ROU;        /* smallest primitive 2N-th root of unity mod Q */
    uint64_t BG;         /* gadget base (power of two) */
    uint32_t BGbits;
    uint32_t digitG;     /* ceil(log2 Q / log2 BG) */
    uint64_t Br;         /* blind-rotation radix */
    uint32_t digitR;     /* ceil(log2 q / log2 Br) */
    uint64_t Bks;        /* RLWE->LWE key-switch base (any integer >= 2) */
    uint32_t digitKS;    /* ceil(log2 Q / log2 Bks) */
    uint64_t Bks_R;      /* RLWE->RLWE key-switch base (power of two) */
    uint32_t Bks_R_bits;
    uint32_t digitKS_R;
    uint64_t E;          /* LWE noise bound (q/16 in the reference apps) */
    uint64_t E_rlwe;     /* RLWE noise bound (18 in the reference apps) */
    double   stddev;     /* Gaussian sigma (3.19) */
    uint64_t gate_constant[6];
    uint64_t *rou;       /* [N] psi^i at bitrev(i) */
    uint64_t *irou;      /* [N] psi^-i at bitrev(i) */
    uint64_t cdt[64];    /* Gaussian magnitude table, scaled to 2^63 */
    int32_t  cdt_len;
    int32_t  eval_mode;  /* blind-rotation order, ORC_EVAL_* (0 after init) */
} orc_params;

/* Blind-rotation orders: the software loop skips residue-0 steps
 * (RLWE.cpp:754-767); the accelerator's one-ciphertext form multiplies at step
 * (0,0) whatever the residue (myfpga.cpp:1030-1071); its batched forms
 * multiply at every step, by btkey[0] = RGSW(1) for residue 0
 * (myfpga.cpp:1180-1225). */
#define ORC_EVAL_SOFTWARE 0
#define ORC_EVAL_FPGA_X1 1
#define ORC_EVAL_FPGA_XN 2

/* ---------- number theory (nbt.cpp) ---------- */
uint64_t orc_reverse_bit(uint64_t order, uint64_t input);
uint64_t orc_power_mod(uint64_t base, uint64_t exp, uint64_t mod);
uint64_t orc_inverse_mod(uint64_t x, uint64_t mod);
int      orc_is_prime(uint64_t x);
uint64_t orc_first_prime(int nbit, int order);
uint64_t orc_previous_prime(uint64_t prime, int order);
uint64_t orc_root_of_unity(int order, uint64_t mod);
void     orc_rou_tables(uint64_t rou, uint64_t mod, int order, uint64_t *tbl, uint64_t *itbl);

/* Fill a parameter block.  Q == 0 derives Q from nbit as the reference apps do
 * (previous_prime(first_prime(nbit, N), N), fhe_wo_fpga.cpp:89).  Returns 0 on
 * success, -1 on invalid parameters.  The tables are malloc'ed. */
int  orc_params_init(orc_params *p, int n, uint64_t q, int N, uint64_t Q, int nbit,
                     uint32_t BGbits, uint64_t Br, uint64_t Bks, uint32_t Bks_R_bits,
                     double stddev, uint64_t E_rlwe);
void orc_params_free(orc_params *p);

/* ---------- transforms (nbt.cpp:423-491) ---------- */
void orc_ntt(uint64_t *a, int N, uint64_t Q, const uint64_t *rou);
void orc_intt(uint64_t *a, int N, uint64_t Q, const uint64_t *irou, uint64_t iN);
void orc_ntt_batch(const orc_params *p, uint64_t *a, int64_t count);
void orc_intt_batch(const orc_params *p, uint64_t *a, int64_t count);

/* ---------- RLWE / RGSW (RLWE.cpp) ---------- */
void orc_poly_substitute(const uint64_t *in, int sub_factor, int N, uint64_t Q, uint64_t *out);
void orc_digit_decompose(const orc_params *p, const uint64_t *rlwe, uint32_t d, uint32_t bits,
                         uint64_t *out /* [d][2][N], NTT domain */);
void orc_rlwe_mult_rgsw(const orc_params *p, const uint64_t *acc, const uint64_t *rgsw, uint64_t *out);
void orc_rlwe_mult_rgsw_batch(const orc_params *p, const uint64_t *acc, const uint64_t *rgsw,
                              int64_t rgsw_stride, uint64_t *out, int64_t count);
void orc_acc_init(const orc_params *p, uint64_t bound1, uint64_t bound2, uint64_t value, uint64_t *acc);
void orc_rlwe_substitute(const orc_params *p, const uint64_t *in, int sub_factor, uint64_t *out);
void orc_rlwe_keyswitch(const orc_params *p, const uint64_t *in, const uint64_t *kskey, uint64_t *out);

/* ---------- RLWE expansion and leveled ops (params.cpp:177-210, RLWE.cpp:377-400,411-430,788-838,
 *            app.cpp:26-97) ---------- */
/* NTT of X^(2^i) (forward) or of (Q-1) X^(N - 2^i) (backward), i < log2 N */
void orc_rotate_poly(const orc_params *p, int log_factor, int forward, uint64_t *out);
void orc_rlwe_add(const orc_params *p, const uint64_t *c1, const uint64_t *c2, uint64_t *out);
void orc_rlwe_sub(const orc_params *p, const uint64_t *c1, const uint64_t *c2, uint64_t *out);
void orc_rlwe_rotate_freq(const orc_params *p, const uint64_t *in, int log_factor, int forward, uint64_t *out);
/* expand_RLWE: out [N][2][N]; subs key [log2 N][digitKS_R][2][N] */
void orc_expand_rlwe(const orc_params *p, const uint64_t *in, const uint64_t *subs, uint64_t *out);
/* CMUX(c0, c1, SEL) = (c1 - c0) (x) SEL + c0 */
void orc_cmux(const orc_params *p, const uint64_t *c0, const uint64_t *c1, const uint64_t *rgsw, uint64_t *out);
/* poly_rotate_time (RLWE.cpp:49-68): a <- a X^(rf), rf = rotate_factor % N (C '%' sign) */
void orc_poly_rotate_time(uint64_t *a, int rotate_factor, int N, uint64_t Q);
/* blind_rotate_full (RLWE.cpp:795-820): X^b, then CMUX(acc, acc X^a[i], C[i]) in the time domain */
void orc_blind_rotate_full(const orc_params *p, const uint64_t *in, int b, const int32_t *a, int steps,
                           const uint64_t *rgsw, uint64_t *out);
/* blind_rotate: count RGSWs (count <= log2 N), acc = CMUX(acc, rotate_freq(acc, i), C[i]) */
void orc_blind_rotate_rlwe(const orc_params *p, const uint64_t *in, int forward, const uint64_t *rgsw,
                           int count, uint64_t *out);

/* ---------- LWE (LWE.cpp) ---------- */
int  orc_lwe_evaluate(const orc_params *p, const uint64_t *c1, const uint64_t *c2, int gate, uint64_t *out);
void orc_lwe_eval_not(const orc_params *p, const uint64_t *c, uint64_t *out);
void orc_lwe_keyswitch(const orc_params *p, const uint64_t *acc, const uint64_t *ksk, uint64_t *out);
uint64_t orc_scale_round(uint64_t x, uint64_t new_mod, uint64_t old_mod);
void orc_lwe_modswitch(const orc_params *p, const uint64_t *in, uint64_t *out);
uint64_t orc_lwe_decrypt(const orc_params *p, const uint64_t *sk, const uint64_t *c);

/* ---------- gate bootstrap (RLWE.cpp:726-786) ---------- */
/* Blind rotation only (acc in NTT domain after the loop), in the order
 * p->eval_mode names.  Returns the number of external products applied. */
int64_t orc_blind_rotate(const orc_params *p, const uint64_t *lwe, const uint64_t *bsk, uint64_t *acc);
/* Full gate; any capture pointer may be NULL.  Returns 0, or -1 when c1 == c2
 * (the reference exits the process there, RLWE.cpp:727-730). */
int  orc_eval_bootstrap(const orc_params *p, const uint64_t *c1, const uint64_t *c2, int gate,
                        const uint64_t *bsk, const uint64_t *ksk, uint64_t *out,
                        uint64_t *cap_acc, uint64_t *cap_ks);
/* Batched over independent ciphertexts with OpenMP-free pthreads (threads<=0:
 * all cores).  gates[i] per ciphertext.  Returns 0 or -1. */
int  orc_eval_bootstrap_batch(const orc_params *p, const uint64_t *c1, const uint64_t *c2,
                              const int32_t *gates, int64_t count, const uint64_t *bsk,
                              const uint64_t *ksk, uint64_t *out, uint64_t *cap_acc,
                              uint64_t *cap_ks, int threads);

/* ---------- seeded key generation (this library's own sampler) ---------- */
uint64_t orc_rng_word(uint64_t seed, uint64_t stream, uint64_t ctr);
int64_t  orc_gauss(const orc_params *p, uint64_t word);
void orc_lwe_secret(const orc_params *p, uint64_t seed, int binary, uint64_t *sk /* [n] mod q */);
void orc_rlwe_secret(const orc_params *p, uint64_t seed, uint64_t *sk_time /* [N] mod Q */);
void orc_rlwe_encrypt_zero(const orc_params *p, const uint64_t *sk_ntt, uint64_t seed, uint64_t stream,
                           uint64_t *out /* [2][N] NTT domain */);
void orc_rgsw_encrypt_monomial(const orc_params *p, const uint64_t *sk_ntt, uint64_t m_in,
                               uint64_t seed, uint64_t stream_base, uint64_t *out);
void orc_bsk_gen(const orc_params *p, const uint64_t *lwe_sk, const uint64_t *rlwe_sk_time,
                 uint64_t seed, uint64_t *bsk, int threads);
void orc_ksk_gen(const orc_params *p, const uint64_t *rlwe_sk_time, const uint64_t *lwe_sk,
                 uint64_t seed, uint64_t *ksk, int threads);
void orc_rlwe_ksk_gen(const orc_params *p, const uint64_t *old_sk_ntt, const uint64_t *new_sk_ntt,
                      uint64_t seed, uint64_t stream_base, uint64_t *ks);
void orc_subs_key_gen(const orc_params *p, const uint64_t *rlwe_sk_time, uint64_t seed, uint64_t *subs);
void orc_lwe_encrypt(const orc_params *p, const uint64_t *sk, uint64_t m, uint64_t seed, uint64_t index,
                     uint64_t *out);
void orc_rlwe_encrypt(const orc_params *p, const uint64_t *sk_ntt, const uint64_t *m_time, uint64_t seed,
                      uint64_t stream, uint64_t *out /* [2][N] NTT domain */);
void orc_rlwe_decrypt(const orc_params *p, const uint64_t *sk_ntt, const uint64_t *c, uint64_t t,
                      uint64_t *m_out /* [N] time domain */);
void orc_rgsw_encrypt(const orc_params *p, const uint64_t *sk_ntt, const uint64_t *m_time, uint64_t seed,
                      uint64_t stream_base, uint64_t *out /* [digitG][2][2][N] */);
/* application::homo_expand (app.cpp:78-97) and the tree selection's CMUX tree
 * (test.cpp:1576-1582) */
void orc_homo_expand(const orc_params *p, const uint64_t *packed, const uint64_t *enc_sec, const uint64_t *subs,
                     uint64_t *out /* [N][digitG][2][2][N] */);
void orc_cmux_tree(const orc_params *p, const uint64_t *leaves, int levels, const uint64_t *rgsw,
                   size_t rgsw_stride, uint64_t *out /* [2][N] */);

/* stream tags shared with the device key generator */
#define ORC_TAG_LWE_SK   1ULL
#define ORC_TAG_RLWE_SK  2ULL
#define ORC_TAG_BSK      3ULL
#define ORC_TAG_KSK      4ULL
#define ORC_TAG_SUBS     5ULL
#define ORC_TAG_LWE_ENC  6ULL
#define ORC_TAG_RLWE_ENC 7ULL
#define ORC_TAG_RGSW_ENC 8ULL
#define ORC_STREAM(tag, idx) (((uint64_t)(tag) << 56) | (uint64_t)(idx))

#ifdef __cplusplus
}
#endif
#endif
