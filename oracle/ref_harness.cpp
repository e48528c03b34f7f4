// ref_harness.cpp -- explicit-input driver around the reference's own CPU path.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (oracle/Makefile, target "ref") together
// with the reference sources FHE_ACC_runtime/src/{nbt,params,RNG,LWE,RLWE,app}.cpp
// where they lie, into oracle/_ref/libfheref.so.  It sets the reference's
// global parameter block (params.h:13-134) the way the apps do
// (apps/fhe_wo_fpga.cpp:35-162), fills the reference's key and ciphertext
// objects from flat arrays, and calls the reference functions unchanged.  Key
// generation never goes through the reference (its RNG is seeded from
// std::random_device, RNG.cpp:5-6); keys come from the caller.
//
// Flat layouts match oracle/fhe_oracle.h.
#include "LWE.h"
#include "RLWE.h"
#include "app.h"
#include "nbt.h"
#include "params.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

namespace {

void fill_rlwe(RLWE::RLWE_ciphertext &c, const uint64_t *src) {
    c.a.assign(src, src + N);
    c.b.assign(src + N, src + 2 * N);
    c.NTT_form = true;
}

void dump_rlwe(const RLWE::RLWE_ciphertext &c, uint64_t *dst) {
    std::memcpy(dst, c.a.data(), sizeof(uint64_t) * N);
    std::memcpy(dst + N, c.b.data(), sizeof(uint64_t) * N);
}

RLWE::RGSW_ciphertext make_rgsw(const uint64_t *src) {
    std::vector<std::vector<RLWE::RLWE_ciphertext>> rows(digitG);
    for (uint32_t h = 0; h < digitG; h++) {
        rows[h].resize(2);
        for (int g = 0; g < 2; g++) fill_rlwe(rows[h][g], src + ((size_t)h * 2 + g) * 2 * N);
    }
    return RLWE::RGSW_ciphertext(rows);
}

void fill_lwe(LWE::LWE_ciphertext &c, const uint64_t *src) {
    c.a.assign(src, src + n);
    c.b = src[n];
}

void dump_lwe(const LWE::LWE_ciphertext &c, uint64_t *dst) {
    std::memcpy(dst, c.a.data(), sizeof(uint64_t) * n);
    dst[n] = c.b;
}

std::unique_ptr<RLWE::bootstrap_key> g_bsk;
std::unique_ptr<LWE::keyswitch_key> g_ksk;

}  // namespace

extern "C" {

// Parameter block, as apps/fhe_wo_fpga.cpp:35-162 assigns it.
int ref_set_params(int n_, uint64_t q_, int N_, uint64_t Q_, uint32_t BGbits_, uint64_t Br_,
                   uint64_t Bks_, uint32_t Bks_R_bits_) {
    n = n_;
    t = 4;
    q = q_;
    q2 = q / 2;
    E = q / 16;
    N = N_;
    N2 = N / 2 + 1;
    digitN = (uint16_t)std::log2((double)N);
    E_rlwe = 18;
    Q = Q_;
    iN = inverse_mod(N, Q);
    ROU = root_of_unity(N, Q);
    free(ROU_table);
    free(iROU_table);
    ROU_table = (ciphertext_t *)malloc(N * sizeof(ciphertext_t));
    iROU_table = (ciphertext_t *)malloc(N * sizeof(ciphertext_t));
    root_of_unity_table(ROU, Q, N, ROU_table, iROU_table);
    BG = 1ULL << BGbits_;
    BGbits = (ciphertext_t)std::log2(BG);
    BG_mask = BG - 1;
    digitG = (ciphertext_t)std::ceil(std::log2((double)Q) / std::log2((double)BG));
    digitG2 = 2 * digitG;
    Br = Br_;
    digitR = (ciphertext_t)std::ceil(std::log2((double)q) / std::log2((double)Br));
    Bks = Bks_;
    digitKS = (ciphertext_t)std::ceil(std::log2((double)Q) / std::log2((double)Bks));
    Bks_R = 1ULL << Bks_R_bits_;
    Bks_R_bits = (ciphertext_t)std::log2(Bks_R);
    digitKS_R = (ciphertext_t)std::ceil(std::log2((double)Q) / std::log2((double)Bks_R));
    Bks_R_mask = Bks_R - 1;
    gate_constant[0] = 5 * (q >> 3);
    gate_constant[1] = 7 * (q >> 3);
    gate_constant[2] = 1 * (q >> 3);
    gate_constant[3] = 3 * (q >> 3);
    gate_constant[4] = 5 * (q >> 3);
    gate_constant[5] = 1 * (q >> 3);
    // rotate_poly_forward/backward (params.cpp:130-134), built as
    // rotate_vector_gen (params.cpp:177-210) builds them but for the
    // log2 N factors that exist at this N (the reference fixes
    // num_rotate_poly = 11, which only fits N = 2048)
    rotate_poly_forward.assign(digitN, std::vector<ciphertext_t>(N, 0));
    rotate_poly_backward.assign(digitN, std::vector<ciphertext_t>(N, 0));
    for (int i = 0; i < digitN; i++) {
        rotate_poly_forward[i][1 << i] = 1;
        rotate_poly_backward[i][N - (1 << i)] = Q - 1;
        NTT(rotate_poly_forward[i].data(), N, Q, ROU_table);
        NTT(rotate_poly_backward[i].data(), N, Q, ROU_table);
    }
    g_bsk.reset();
    g_ksk.reset();
    return 0;
}

void ref_get_params(uint64_t *out) {
    // ROU, iN, digitG, digitR, digitKS, digitKS_R
    out[0] = ROU;
    out[1] = iN;
    out[2] = digitG;
    out[3] = digitR;
    out[4] = digitKS;
    out[5] = digitKS_R;
}

void ref_ntt(uint64_t *a, int64_t count) {
    for (int64_t c = 0; c < count; c++) NTT(a + c * N, N, Q, ROU_table);
}

void ref_intt(uint64_t *a, int64_t count) {
    for (int64_t c = 0; c < count; c++) iNTT(a + c * N, N, Q, iROU_table, iN);
}

void ref_mult_rgsw(const uint64_t *acc, const uint64_t *rgsw, uint64_t *out) {
    RLWE::RLWE_ciphertext a;
    fill_rlwe(a, acc);
    RLWE::RLWE_ciphertext r = RLWE::RLWE_mult_RGSW(a, make_rgsw(rgsw));
    dump_rlwe(r, out);
}

void ref_acc_init(uint64_t bound1, uint64_t bound2, uint64_t value, uint64_t *out) {
    RLWE::RLWE_ciphertext acc = RLWE::acc_initialization(2 * N / q, bound1, bound2, value);
    dump_rlwe(acc, out);
}

void ref_lwe_evaluate(const uint64_t *c1, const uint64_t *c2, int gate, uint64_t *out) {
    LWE::LWE_ciphertext a, b;
    fill_lwe(a, c1);
    fill_lwe(b, c2);
    dump_lwe(LWE::LWE_evaluate(a, b, (GATES)gate), out);
}

void ref_modswitch(const uint64_t *in, uint64_t *out) {
    LWE::LWE_ciphertext a;
    fill_lwe(a, in);
    dump_lwe(LWE::LWE_modswitch(a), out);
}

uint64_t ref_scale_round(uint64_t x, uint64_t new_mod, uint64_t old_mod) {
    return LWE::scale_round(x, new_mod, old_mod);
}

// Keys: flat bootstrap key [Br][digitR][n] RGSW and key-switch key
// [Bks][digitKS][N] LWE(n).  Kept resident between calls.
void ref_load_keys(const uint64_t *bsk, const uint64_t *ksk) {
    size_t rw = (size_t)digitG * 4 * N;
    g_bsk.reset(new RLWE::bootstrap_key());
    g_bsk->btkey.resize(Br);
    for (uint64_t r = 0; r < Br; r++) {
        g_bsk->btkey[r].resize(digitR);
        for (uint64_t j = 0; j < digitR; j++) {
            g_bsk->btkey[r][j].reserve(n);
            for (int k = 0; k < n; k++)
                g_bsk->btkey[r][j].push_back(make_rgsw(bsk + ((r * digitR + j) * n + k) * rw));
        }
    }
    g_ksk.reset(new LWE::keyswitch_key());
    g_ksk->kskey.resize(Bks);
    size_t row = (size_t)n + 1;
    for (uint64_t d = 0; d < Bks; d++) {
        g_ksk->kskey[d].resize(digitKS);
        for (uint64_t j = 0; j < digitKS; j++) {
            g_ksk->kskey[d][j].resize(N);
            for (int i = 0; i < N; i++) fill_lwe(g_ksk->kskey[d][j][i], ksk + ((d * digitKS + j) * N + i) * row);
        }
    }
}

void ref_lwe_keyswitch(const uint64_t *acc, uint64_t *out) {
    RLWE::RLWE_ciphertext a;
    fill_rlwe(a, acc);
    dump_lwe(LWE::LWE_keyswitch(a, *g_ksk), out);
}

// RLWE::eval_bootstrap on each pair (RLWE.cpp:726-786).  Returns the wall time
// in seconds.  Equal operands make the reference exit the process, so the
// caller must not pass them.
double ref_eval_bootstrap(const uint64_t *c1, const uint64_t *c2, const int32_t *gates, int64_t count,
                          uint64_t *out) {
    size_t row = (size_t)n + 1;
    auto t0 = std::chrono::steady_clock::now();
    for (int64_t c = 0; c < count; c++) {
        LWE::LWE_ciphertext a, b;
        fill_lwe(a, c1 + c * row);
        fill_lwe(b, c2 + c * row);
        LWE::LWE_ciphertext r = RLWE::eval_bootstrap(a, b, *g_bsk, *g_ksk, (GATES)gates[c]);
        dump_lwe(r, out + c * row);
    }
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// The accelerator's gate orders (FPGA::fpga_eval_bootstrap_x1 and _x4..x16,
// myfpga.cpp:1003-1660), driven through the reference's own CPU routines:
// LWE_evaluate, acc_initialization, RLWE_mult_RGSW, LWE_keyswitch and
// LWE_modswitch.  The card's BOOTSTRAP_INIT writes the test polynomial where
// the external product's first inverse NTT would write
// (cl_fhe_acc/design/src/iNTT_module.sv:1154-1260), which equals
// acc_initialization followed by RLWE_mult_RGSW.  mode 1: step (0,0) always
// multiplies, later residue-0 steps are skipped (myfpga.cpp:1030-1071);
// mode 2: every step multiplies, by btkey[0] for residue 0
// (myfpga.cpp:1180-1225).  cap_acc may be null.  Returns -1 for another mode.
int ref_eval_bootstrap_fpga(const uint64_t *c1, const uint64_t *c2, const int32_t *gates, int64_t count, int mode,
                            uint64_t *out, uint64_t *cap_acc) {
    if (mode != 1 && mode != 2) return -1;
    size_t row = (size_t)n + 1;
    for (int64_t c = 0; c < count; c++) {
        LWE::LWE_ciphertext a, b;
        fill_lwe(a, c1 + c * row);
        fill_lwe(b, c2 + c * row);
        LWE::LWE_ciphertext t = LWE::LWE_evaluate(a, b, (GATES)gates[c]);
        ciphertext_t bound1 = mod_pow2(gate_constant[gates[c]], q);
        ciphertext_t bound2 = mod_pow2(bound1 + q2, q);
        RLWE::RLWE_ciphertext acc = RLWE::acc_initialization(2 * N / q, bound1, bound2, t.b);
        for (int i = 0; i < n; i++) {
            ciphertext_t nega = (q - t.a[i]) % q;
            for (uint64_t j = 0; j < digitR; j++) {
                ciphertext_t r = nega % Br;
                if (r > 0 || mode == 2 || (i == 0 && j == 0)) acc = RLWE::RLWE_mult_RGSW(acc, g_bsk->btkey[r][j][i]);
                nega /= Br;
            }
        }
        if (cap_acc) dump_rlwe(acc, cap_acc + c * 2 * N);
        dump_lwe(LWE::LWE_modswitch(LWE::LWE_keyswitch(acc, *g_ksk)), out + c * row);
    }
    return 0;
}

void ref_substitute(const uint64_t *in, int f, uint64_t *out) {
    RLWE::RLWE_ciphertext a;
    fill_rlwe(a, in);
    dump_rlwe(RLWE::RLWE_substitute(a, f), out);
}

void ref_rlwe_keyswitch(const uint64_t *in, const uint64_t *ks, uint64_t *out) {
    RLWE::RLWE_ciphertext a;
    fill_rlwe(a, in);
    RLWE::keyswitch_key k;
    k.RLWE_kskey.resize(digitKS_R);
    for (uint64_t l = 0; l < digitKS_R; l++) fill_rlwe(k.RLWE_kskey[l], ks + l * 2 * N);
    dump_rlwe(RLWE::RLWE_keyswitch(a, k), out);
}

// application::expand_RLWE (app.cpp:26-97); subs: [digitN][digitKS_R][2][N]
void ref_expand_rlwe(const uint64_t *in, const uint64_t *subs, uint64_t *out) {
    RLWE::RLWE_ciphertext c;
    fill_rlwe(c, in);
    application::substitute_key sk;
    sk.subs_key.resize(digitN);
    for (int i = 0; i < digitN; i++) {
        sk.subs_key[i].RLWE_kskey.resize(digitKS_R);
        for (uint64_t l = 0; l < digitKS_R; l++)
            fill_rlwe(sk.subs_key[i].RLWE_kskey[l], subs + ((size_t)i * digitKS_R + l) * 2 * N);
    }
    auto r = application::expand_RLWE(c, sk);
    for (int k = 0; k < N; k++) dump_rlwe((*r)[k], out + (size_t)k * 2 * N);
}

// RLWE::RLWE_rotate_freq (RLWE.cpp:411-430)
void ref_rotate_freq(const uint64_t *in, int log_factor, int forward, uint64_t *out) {
    RLWE::RLWE_ciphertext c;
    fill_rlwe(c, in);
    dump_rlwe(RLWE::RLWE_rotate_freq(c, log_factor, forward != 0), out);
}

// RLWE::CMUX (RLWE.cpp:788-793)
void ref_cmux(const uint64_t *c0, const uint64_t *c1, const uint64_t *rgsw, uint64_t *out) {
    RLWE::RLWE_ciphertext a, b;
    fill_rlwe(a, c0);
    fill_rlwe(b, c1);
    dump_rlwe(RLWE::CMUX(a, b, make_rgsw(rgsw)), out);
}

// RLWE::blind_rotate (RLWE.cpp:822-838): num_rotate_poly (11) RGSWs, N = 2048
int ref_blind_rotate_rlwe(const uint64_t *in, int forward, const uint64_t *rgsw, uint64_t *out) {
    if (digitN != num_rotate_poly) return -1;
    RLWE::RLWE_ciphertext c;
    fill_rlwe(c, in);
    std::vector<RLWE::RGSW_ciphertext> C;
    for (int i = 0; i < num_rotate_poly; i++) C.push_back(make_rgsw(rgsw + (size_t)i * digitG * 4 * N));
    dump_rlwe(RLWE::blind_rotate(c, forward != 0, C), out);
    return 0;
}

// RLWE::blind_rotate_full (RLWE.cpp:795-820)
void ref_blind_rotate_full(const uint64_t *in, int b, const int32_t *a, int steps, const uint64_t *rgsw,
                           uint64_t *out) {
    RLWE::RLWE_ciphertext c;
    fill_rlwe(c, in);
    std::vector<int> av(a, a + steps);
    std::vector<RLWE::RGSW_ciphertext> C;
    for (int i = 0; i < steps; i++) C.push_back(make_rgsw(rgsw + (size_t)i * digitG * 4 * N));
    dump_rlwe(RLWE::blind_rotate_full(c, b, av, C), out);
}

// application::homo_expand (app.cpp:78-97); packed [digitG][2][N], enc_sec one
// RGSW, subs [digitN][digitKS_R][2][N]; out [N][digitG][2][2][N]
void ref_homo_expand(const uint64_t *packed, const uint64_t *enc_sec, const uint64_t *subs, uint64_t *out) {
    auto pk = std::make_unique<std::vector<RLWE::RLWE_ciphertext>>(digitG);
    for (uint32_t j = 0; j < digitG; j++) fill_rlwe((*pk)[j], packed + (size_t)j * 2 * N);
    application::substitute_key sk;
    sk.subs_key.resize(digitN);
    for (int i = 0; i < digitN; i++) {
        sk.subs_key[i].RLWE_kskey.resize(digitKS_R);
        for (uint64_t l = 0; l < digitKS_R; l++)
            fill_rlwe(sk.subs_key[i].RLWE_kskey[l], subs + ((size_t)i * digitKS_R + l) * 2 * N);
    }
    auto r = application::homo_expand(pk, make_rgsw(enc_sec), sk);
    for (int i = 0; i < N; i++)
        for (uint32_t h = 0; h < digitG; h++)
            for (int g = 0; g < 2; g++)
                dump_rlwe((*r)[i].c_text[h][g], out + (((size_t)i * digitG + h) * 2 + g) * 2 * N);
}

// The tree selection's CMUX loop (test.cpp:1576-1582) over the reference's
// RLWE::CMUX: level l halves the 2^levels leaves with rgsw[l].
void ref_cmux_tree(const uint64_t *leaves, int levels, const uint64_t *rgsw, uint64_t *out) {
    uint32_t size = 1u << levels;
    std::vector<RLWE::RLWE_ciphertext> v(size);
    for (uint32_t i = 0; i < size; i++) fill_rlwe(v[i], leaves + (size_t)i * 2 * N);
    for (int l = 0; l < levels; l++) {
        RLWE::RGSW_ciphertext C = make_rgsw(rgsw + (size_t)l * digitG * 4 * N);
        for (uint32_t j = 0; j < size; j += 2) v[j / 2] = RLWE::CMUX(v[j], v[j + 1], C);
        size >>= 1;
    }
    dump_rlwe(v[0], out);
}

// RLWE::RLWE_decrypt(s, c, t) (RLWE.cpp:362-376); sk in the time domain
// (the RLWE_secretkey(vector) constructor transforms it, RLWE.cpp:281-291)
void ref_rlwe_decrypt(const uint64_t *sk_time, const uint64_t *c, uint64_t t_rlwe, uint64_t *out) {
    RLWE::RLWE_secretkey sk(std::vector<ciphertext_t>(sk_time, sk_time + N));
    RLWE::RLWE_ciphertext ct;
    fill_rlwe(ct, c);
    RLWE::RLWE_plaintext m = RLWE::RLWE_decrypt(sk, ct, t_rlwe);
    std::memcpy(out, m.m.data(), sizeof(uint64_t) * N);
}

uint64_t ref_root_of_unity(int order, uint64_t mod) { return root_of_unity(order, mod); }
uint64_t ref_first_prime(int nbit, int order) { return first_prime(nbit, order); }
uint64_t ref_previous_prime(uint64_t p, int order) { return previous_prime(p, order); }

}  // extern "C"
