// gate_demo.cpp -- the reference's gate-bootstrapping test (test::bootstrapping,
// FHE_ACC_runtime/src/test.cpp:496-625: every gate on every input pair, then
// NOT) written against the reference-shaped host API, once one gate per call
// (RLWE::eval_bootstrap), once as one batch (fheacc::eval_bootstrap_batch),
// and through the accelerator's calls in its own step order
// (FPGA::fpga_eval_bootstrap_x1..x16, myfpga.cpp:1003-1660).
//
// usage: gate_demo [trials]   (default 4 trials per gate and input pair)
#include <stdio.h>
#include <stdlib.h>

#include <chrono>

#include "fhe_acc.hpp"

static int truth(GATES g, int a, int b) {
    switch (g) {
        case OR: return a | b;
        case AND: return a & b;
        case NOR: return !(a | b);
        case NAND: return !(a & b);
        case XOR: return a ^ b;
        default: return !(a ^ b);
    }
}

int main(int argc, char **argv) {
    const int trials = argc > 1 ? atoi(argv[1]) : 4;
    fheacc::Params p;  // the benchmark set: n=500, q=1024, N=1024, 27-bit Q, BG=2^7, Br=32, Bks=2^5
    fheacc::Context ctx(p);
    printf("Q = %llu, digitG = %u, digitR = %u, digitKS = %u\n", (unsigned long long)ctx.Q(), ctx.info().digitG,
           ctx.info().digitR, ctx.info().digitKS);
    LWE::LWE_secretkey s(ctx, 11, true);
    RLWE::RLWE_secretkey z(ctx, 11);
    auto t0 = std::chrono::steady_clock::now();
    RLWE::bootstrap_key bt(ctx, s, z, 11);
    LWE::keyswitch_key ks(bt);
    auto t1 = std::chrono::steady_clock::now();
    printf("keygen on device: %.3f s\n", std::chrono::duration<double>(t1 - t0).count());

    const GATES gates[6] = {OR, AND, NOR, NAND, XOR, XNOR};
    int fails = 0;
    uint64_t idx = 0;
    // one gate per call
    for (GATES g : gates) {
        for (int m = 0; m < 4; m++) {
            int a = m >> 1, b = m & 1;
            LWE::LWE_ciphertext c1 = LWE::LWE_encrypt(ctx, s, LWE::LWE_plaintext(a), 5, idx++);
            LWE::LWE_ciphertext c2 = LWE::LWE_encrypt(ctx, s, LWE::LWE_plaintext(b), 5, idx++);
            LWE::LWE_ciphertext r = RLWE::eval_bootstrap(c1, c2, bt, ks, g);
            if ((int)LWE::LWE_decrypt(ctx, s, r).m != truth(g, a, b)) fails++;
            LWE::LWE_ciphertext nr = LWE::LWE_eval_NOT(ctx, r);
            if ((int)LWE::LWE_decrypt(ctx, s, nr).m != !truth(g, a, b)) fails++;
        }
    }
    printf("single-call gates: %d failures\n", fails);
    // one batch of every gate, input pair and trial
    std::vector<LWE::LWE_ciphertext> c1, c2;
    std::vector<GATES> gv;
    std::vector<int> expect;
    for (int t = 0; t < trials; t++)
        for (GATES g : gates)
            for (int m = 0; m < 4; m++) {
                int a = m >> 1, b = m & 1;
                c1.push_back(LWE::LWE_encrypt(ctx, s, LWE::LWE_plaintext(a), 6, idx++));
                c2.push_back(LWE::LWE_encrypt(ctx, s, LWE::LWE_plaintext(b), 6, idx++));
                gv.push_back(g);
                expect.push_back(truth(g, a, b));
            }
    int rc = 0;
    auto t2 = std::chrono::steady_clock::now();
    std::vector<LWE::LWE_ciphertext> out = fheacc::eval_bootstrap_batch(c1, c2, ks, gv, &rc);
    auto t3 = std::chrono::steady_clock::now();
    if (rc != 0) {
        printf("batch failed rc=%d\n", rc);
        return 1;
    }
    int bfails = 0;
    for (size_t i = 0; i < out.size(); i++)
        if ((int)LWE::LWE_decrypt(ctx, s, out[i]).m != expect[i]) bfails++;
    printf("batch of %zu gates: %d failures, %.3f ms\n", out.size(), bfails,
           1e3 * std::chrono::duration<double>(t3 - t2).count());
    // the accelerator's calls: its step order and its batch sizes
    int ffails = 0, frc = 0;
    for (size_t i = 0; i < 4; i++) {
        LWE::LWE_ciphertext r = FPGA::fpga_eval_bootstrap_x1(c1[i], c2[i], ks, gv[i], &frc);
        if (frc != 0 || (int)LWE::LWE_decrypt(ctx, s, r).m != expect[i]) ffails++;
    }
    const size_t sizes[4] = {4, 8, 12, 10};
    for (int f = 0; f < 4; f++) {
        const size_t m = sizes[f];
        std::vector<LWE::LWE_ciphertext> a(c1.begin(), c1.begin() + m), b(c2.begin(), c2.begin() + m);
        std::vector<GATES> g(gv.begin(), gv.begin() + m);
        std::vector<LWE::LWE_ciphertext> r = f == 0   ? FPGA::fpga_eval_bootstrap_x4(a, b, ks, g, &frc)
                                             : f == 1 ? FPGA::fpga_eval_bootstrap_x8(a, b, ks, g, &frc)
                                             : f == 2 ? FPGA::fpga_eval_bootstrap_x12(a, b, ks, g, &frc)
                                                      : FPGA::fpga_eval_bootstrap_x16(a, b, ks, g, &frc);
        if (frc != 0) ffails++;
        for (size_t i = 0; i < m && frc == 0; i++)
            if ((int)LWE::LWE_decrypt(ctx, s, r[i]).m != expect[i]) ffails++;
    }
    // a batch size the card's call does not take
    std::vector<LWE::LWE_ciphertext> a3(c1.begin(), c1.begin() + 3), b3(c2.begin(), c2.begin() + 3);
    std::vector<GATES> g3(gv.begin(), gv.begin() + 3);
    FPGA::fpga_eval_bootstrap_x4(a3, b3, ks, g3, &frc);
    const bool size_refused = frc == -1;
    // the context's own order is back: the same words as the batch above
    std::vector<LWE::LWE_ciphertext> a4(c1.begin(), c1.begin() + 4), b4(c2.begin(), c2.begin() + 4);
    std::vector<GATES> g4(gv.begin(), gv.begin() + 4);
    std::vector<LWE::LWE_ciphertext> again = fheacc::eval_bootstrap_batch(a4, b4, ks, g4, &rc);
    bool restored = rc == 0;
    for (size_t i = 0; i < 4 && restored; i++) restored = again[i].a == out[i].a && again[i].b == out[i].b;
    printf("accelerator-order calls: %d failures, size check %s, order restored %s\n", ffails,
           size_refused ? "ok" : "FAILED", restored ? "yes" : "NO");
    // identical operands are refused by the batched call
    std::vector<GATES> one = {NAND};
    fheacc::eval_bootstrap_batch({c1[0]}, {c1[0]}, ks, one, &rc);
    printf("identical operands: rc = %d\n", rc);
    const bool ok = fails == 0 && bfails == 0 && ffails == 0 && size_refused && restored && rc == -1;
    printf(ok ? "ALL GATES PASSED\n" : "GATE TEST FAILED\n");
    return ok ? 0 : 1;
}
