// tree_demo.cpp -- the reference's homo_expansion_and_tree_selection test
// (FHE_ACC_runtime/src/test.cpp:1510-1600; the accelerator's form
// test.cpp:3984-4575) through the reference-shaped C++ layer: a substitution
// key, RGSW(-s), digitG packed bits b_i = i & 1, homo_expand, a CMUX tree over
// 2^(18 - log2 N) leaves with RGSWs log2 N..17, then blind_rotate backward with
// RGSWs 0..log2 N - 1, and RLWE_decrypt with t = 2.
//
// usage: tree_demo [N]   (N = 2048 is the reference app's size, 54-bit Q,
//                         BG = 2^9, Bks_R = 2^9; smaller N for quick runs)
// Prints the selected leaf, the rotation and "TREE DEMO OK" when the decrypted
// polynomial is leaf sel rotated by X^-rot, as the bits fix it.
#include <stdio.h>
#include <stdlib.h>

#include <vector>

#include "fhe_acc.hpp"

int main(int argc, char **argv) {
    fheacc::Params p;
    p.n = 512;
    p.q = 512;
    p.N = argc > 1 ? atoi(argv[1]) : 2048;
    p.Q = 0x3FFFFFFFFED001ULL;
    p.BGbits = 9;
    p.Br = 23;
    p.Bks = 25;
    p.Bks_R_bits = 9;
    if (p.q > 2 * (ciphertext_t)p.N) p.q = 2 * (ciphertext_t)p.N;
    fheacc::Context ctx(p);
    const int N = ctx.N(), digitN = __builtin_ctz((unsigned)N);
    const uint32_t digitG = ctx.info().digitG;
    const ciphertext_t Q = ctx.Q(), iN = ctx.info().iN;
    RLWE::RLWE_secretkey sk(ctx, 2023);

    printf("generating the substitution key\n");
    application::substitute_key subs_key(ctx, sk, 2024);
    printf("generate RGSW encryption of -s\n");
    std::vector<ciphertext_t> neg(N);
    for (int i = 0; i < N; i++) neg[i] = (Q - sk.a[i]) % Q;
    RLWE::RGSW_ciphertext RGSW_neg_s =
        RLWE::RGSW_encrypt(ctx, sk, RLWE::RGSW_plaintext(RLWE::RLWE_plaintext(neg)), 2025, 0);

    printf("build packed bits RLWE vector\n");
    auto packed_bits = std::make_unique<std::vector<RLWE::RLWE_ciphertext>>(digitG);
    unsigned __int128 pow_BG = 1;
    for (uint32_t h = 0; h < digitG; h++) {
        std::vector<ciphertext_t> m(N, 0);
        for (int j = 1; j < N; j += 2) m[j] = (ciphertext_t)(pow_BG * iN % Q);
        (*packed_bits)[h] = RLWE::RLWE_encrypt(ctx, sk, RLWE::RLWE_plaintext(m), 2026, h);
        pow_BG = pow_BG * ((ciphertext_t)1 << p.BGbits) % Q;
    }

    printf("start homo expansion\n");
    auto expanded = application::homo_expand(packed_bits, RGSW_neg_s, subs_key);
    int rc = 0;
    auto expanded_fpga = FPGA::fpga_homo_expand(packed_bits, RGSW_neg_s, subs_key, &rc);
    if (rc || !expanded_fpga) return 1;
    for (int i = 0; i < N; i++)
        for (uint32_t h = 0; h < digitG; h++)
            for (int g = 0; g < 2; g++)
                if ((*expanded)[i].c_text[h][g].a != (*expanded_fpga)[i].c_text[h][g].a ||
                    (*expanded)[i].c_text[h][g].b != (*expanded_fpga)[i].c_text[h][g].b) {
                    printf("fpga_homo_expand differs from homo_expand at RGSW %d\n", i);
                    return 1;
                }

    const int total_digits = 18;
    uint32_t size = 1u << (total_digits - digitN);
    std::vector<RLWE::RLWE_ciphertext> v(size);
    for (uint32_t i = 0; i < size; i++) {
        std::vector<ciphertext_t> poly(N, 0);
        for (uint32_t j = 0; j <= i && j < (uint32_t)N; j++) poly[j] = Q / 2;
        v[i] = RLWE::RLWE_encrypt(ctx, sk, RLWE::RLWE_plaintext(poly), 2027, i);
    }
    printf("binary tree selection with the %d LSBs from packed RGSW bits\n", total_digits);
    for (int i = digitN; i < total_digits; i++) {
        for (uint32_t j = 0; j < size; j += 2) v[j / 2] = RLWE::CMUX(ctx, v[j], v[j + 1], (*expanded)[i]);
        size >>= 1;
    }
    printf("blind rotate\n");
    std::vector<RLWE::RGSW_ciphertext> rotate_bits(expanded->begin(), expanded->begin() + digitN);
    RLWE::RLWE_ciphertext c_rotated = RLWE::blind_rotate(ctx, v[0], false, rotate_bits);
    RLWE::RLWE_plaintext p_rotated = RLWE::RLWE_decrypt(ctx, sk, c_rotated, 2);

    // bit i of the packed vector is i & 1
    int sel = 0, rot = 0;
    for (int i = digitN; i < total_digits; i++) sel |= (i & 1) << (i - digitN);
    for (int i = 0; i < digitN; i++) rot |= (i & 1) << i;
    int bad = 0;
    for (int k = 0; k < N; k++) {
        // leaf sel (Q/2 at 0..sel) times X^-rot: coefficient k is leaf[k + rot],
        // negated past N (-Q/2 rounds to 1 as well)
        const int src = k + rot;
        const int want = src < N ? (src <= sel) : (src - N <= sel);
        bad += (int)p_rotated.m[k] != want;
    }
    printf("selected leaf %d, rotation %d, %d wrong coefficients\n", sel, rot, bad);
    if (bad) return 1;
    printf("TREE DEMO OK: N=%d digitG=%u leaves=%u\n", N, digitG, 1u << (total_digits - digitN));
    return 0;
}
