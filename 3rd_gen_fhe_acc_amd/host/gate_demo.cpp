// gate_demo.cpp -- the reference's gate-bootstrapping test (test::bootstrapping,
// FHE_ACC_runtime/src/test.cpp:496-625: every gate on every input pair, then
// NOT) written against the reference-shaped host API, once one gate per call
// (RLWE::eval_bootstrap) and once as one batch (fheacc::eval_bootstrap_batch).
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
    // identical operands are refused by the batched call
    std::vector<GATES> one = {NAND};
    fheacc::eval_bootstrap_batch({c1[0]}, {c1[0]}, ks, one, &rc);
    printf("identical operands: rc = %d\n", rc);
    const bool ok = fails == 0 && bfails == 0 && rc == -1;
    printf(ok ? "ALL GATES PASSED\n" : "GATE TEST FAILED\n");
    return ok ? 0 : 1;
}
