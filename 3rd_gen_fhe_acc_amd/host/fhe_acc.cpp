// fhe_acc.cpp -- the reference-shaped C++ host API over the C ABI.
#include "fhe_acc.hpp"

#include <stdio.h>
#include <stdlib.h>

#include <mutex>

namespace fheacc {

void check(int rc) {
    if (rc != 0) throw Error(rc, std::string("fheacc: ") + fheacc_last_error());
}

Context::Context(const Params &p, int device) : p_(p) {
    fheacc_params cp{p.n, p.q, p.N, p.Q, p.nbit, p.BGbits, p.Br, p.Bks, p.Bks_R_bits, p.stddev, p.E_rlwe};
    check(fheacc_ctx_create(&cp, device, &h_));
    check(fheacc_ctx_info(h_, &info_));
}

Context::~Context() { fheacc_ctx_destroy(h_); }

KeySet::KeySet(const Context &ctx) : ctx_(ctx) { check(fheacc_keys_create(ctx.handle(), &h_)); }

KeySet::~KeySet() { fheacc_keys_destroy(h_); }

}  // namespace fheacc

namespace LWE {

LWE_secretkey::LWE_secretkey(const fheacc::Context &ctx, uint64_t seed, bool binary) {
    a.assign(ctx.n(), 0);
    fheacc::check(fheacc_lwe_secret(ctx.handle(), seed, binary ? 1 : 0, a.data()));
    b = 1;
}

keyswitch_key::keyswitch_key(const RLWE::bootstrap_key &bt) : keys(bt.keys) {}

LWE_ciphertext LWE_encrypt(const fheacc::Context &ctx, const LWE_secretkey &s, const LWE_plaintext &p, uint64_t seed,
                           uint64_t index) {
    std::vector<ciphertext_t> w(ctx.n() + 1);
    uint8_t bit = (uint8_t)p.m;
    fheacc::check(fheacc_lwe_encrypt(ctx.handle(), s.a.data(), &bit, 1, seed, index, w.data()));
    LWE_ciphertext c;
    c.a.assign(w.begin(), w.begin() + ctx.n());
    c.b = w[ctx.n()];
    return c;
}

LWE_plaintext LWE_decrypt(const fheacc::Context &ctx, const LWE_secretkey &s, const LWE_ciphertext &c) {
    std::vector<ciphertext_t> w(c.a);
    w.push_back(c.b);
    uint8_t bit = 0;
    fheacc::check(fheacc_lwe_decrypt(ctx.handle(), s.a.data(), w.data(), 1, &bit));
    return LWE_plaintext(bit);
}

// LWE.cpp:203-222
LWE_ciphertext LWE_evaluate(const fheacc::Context &ctx, const LWE_ciphertext &c1, const LWE_ciphertext &c2,
                            GATES gate) {
    const ciphertext_t q = ctx.q();
    LWE_ciphertext r(ctx.n());
    const bool x = gate == XOR || gate == XNOR;
    for (int i = 0; i < ctx.n(); i++) r.a[i] = (x ? 2 * (c1.a[i] - c2.a[i]) : c1.a[i] + c2.a[i]) % q;
    r.b = (x ? 2 * (c1.b - c2.b) : c1.b + c2.b) % q;
    return r;
}

// LWE.cpp:224-231
LWE_ciphertext LWE_eval_NOT(const fheacc::Context &ctx, const LWE_ciphertext &c) {
    const ciphertext_t q = ctx.q();
    LWE_ciphertext r(ctx.n());
    for (int i = 0; i < ctx.n(); i++) r.a[i] = (q - c.a[i]) % q;
    r.b = ((q / 4) - c.b) % q;
    return r;
}

}  // namespace LWE

namespace RLWE {

RLWE_secretkey::RLWE_secretkey(const fheacc::Context &ctx, uint64_t seed) {
    a.assign(ctx.N(), 0);
    fheacc::check(fheacc_rlwe_secret(ctx.handle(), seed, a.data()));
}

bootstrap_key::bootstrap_key(const fheacc::Context &ctx, const LWE::LWE_secretkey &s, const RLWE_secretkey &z,
                             uint64_t seed)
    : keys(std::make_shared<fheacc::KeySet>(ctx)) {
    fheacc::check(fheacc_keys_generate(keys->handle(), seed, s.a.data(), z.a.data()));
}

bootstrap_key::bootstrap_key(const fheacc::Context &ctx, const std::vector<ciphertext_t> &bt,
                             const std::vector<ciphertext_t> &ks)
    : keys(std::make_shared<fheacc::KeySet>(ctx)) {
    if (bt.size() != ctx.info().bsk_words || ks.size() != ctx.info().ksk_words)
        throw fheacc::Error(-1, "key sizes do not match the parameter set");
    fheacc::check(fheacc_keys_upload_bsk(keys->handle(), bt.data()));
    fheacc::check(fheacc_keys_upload_ksk(keys->handle(), ks.data()));
}

LWE::LWE_ciphertext eval_bootstrap(const LWE::LWE_ciphertext &c1, const LWE::LWE_ciphertext &c2,
                                   const bootstrap_key &btkey, const LWE::keyswitch_key &kskey, const GATES gate) {
    if (c1 == c2) {
        // RLWE.cpp:727-730
        printf("Please only use independent ciphertexts!!!\n");
        exit(1);
    }
    if (!btkey.keys || btkey.keys != kskey.keys) throw fheacc::Error(-1, "bootstrap and key-switch keys differ");
    int rc = 0;
    std::vector<LWE::LWE_ciphertext> r = fheacc::eval_bootstrap_batch({c1}, {c2}, kskey, {gate}, &rc);
    fheacc::check(rc);
    return r[0];
}

static std::vector<ciphertext_t> flat(const fheacc::Context &ctx, const RLWE_ciphertext &c) {
    const int N = ctx.N();
    if ((int)c.a.size() != N || (int)c.b.size() != N || !c.NTT_form)
        throw fheacc::Error(-1, "RLWE ciphertext must hold N words per polynomial, NTT domain");
    std::vector<ciphertext_t> w(c.a);
    w.insert(w.end(), c.b.begin(), c.b.end());
    return w;
}

static RLWE_ciphertext unflat(int N, const ciphertext_t *w) {
    RLWE_ciphertext c;
    c.a.assign(w, w + N);
    c.b.assign(w + N, w + 2 * N);
    return c;
}

// the reference RGSW word order [h][g][a, b] (RLWE.h:313-329)
static void flat_rgsw(const fheacc::Context &ctx, const RGSW_ciphertext &C, std::vector<ciphertext_t> &w) {
    const uint32_t d = ctx.info().digitG;
    if (C.c_text.size() != d) throw fheacc::Error(-1, "RGSW ciphertext must hold digitG rows");
    for (uint32_t h = 0; h < d; h++) {
        if (C.c_text[h].size() != 2) throw fheacc::Error(-1, "RGSW rows hold two RLWE ciphertexts");
        for (int g = 0; g < 2; g++) {
            std::vector<ciphertext_t> r = flat(ctx, C.c_text[h][g]);
            w.insert(w.end(), r.begin(), r.end());
        }
    }
}

static std::vector<ciphertext_t> plain_words(const fheacc::Context &ctx, const RLWE_plaintext &p) {
    if ((int)p.m.size() != ctx.N()) throw fheacc::Error(-1, "RLWE plaintext must hold N words");
    return p.m;
}

RLWE_ciphertext RLWE_encrypt(const fheacc::Context &ctx, const RLWE_secretkey &s, const RLWE_plaintext &p,
                             uint64_t seed, uint64_t index) {
    std::vector<ciphertext_t> m = plain_words(ctx, p), out(2 * (size_t)ctx.N());
    fheacc::check(fheacc_rlwe_encrypt(ctx.handle(), s.a.data(), m.data(), 1, seed, index, out.data()));
    return unflat(ctx.N(), out.data());
}

RLWE_plaintext RLWE_decrypt(const fheacc::Context &ctx, const RLWE_secretkey &s, const RLWE_ciphertext &c,
                            ciphertext_t t) {
    std::vector<ciphertext_t> x = flat(ctx, c), out(ctx.N());
    fheacc::check(fheacc_rlwe_decrypt(ctx.handle(), s.a.data(), x.data(), 1, t, out.data()));
    return RLWE_plaintext(out);
}

static RGSW_ciphertext unflat_rgsw(const fheacc::Context &ctx, const ciphertext_t *w) {
    const uint32_t d = ctx.info().digitG;
    const int N = ctx.N();
    std::vector<std::vector<RLWE_ciphertext>> rows(d);
    for (uint32_t h = 0; h < d; h++)
        for (int g = 0; g < 2; g++) rows[h].push_back(unflat(N, w + ((size_t)h * 2 + g) * 2 * N));
    return RGSW_ciphertext(std::move(rows));
}

RGSW_ciphertext RGSW_encrypt(const fheacc::Context &ctx, const RLWE_secretkey &z, const RGSW_plaintext &m,
                             uint64_t seed, uint64_t index) {
    std::vector<ciphertext_t> w = plain_words(ctx, m.m), out((size_t)ctx.info().digitG * 4 * ctx.N());
    fheacc::check(fheacc_rgsw_encrypt(ctx.handle(), z.a.data(), w.data(), 1, seed, index, out.data()));
    return unflat_rgsw(ctx, out.data());
}

RLWE_ciphertext RLWE_mult_RGSW(const fheacc::Context &ctx, const RLWE_ciphertext &c, const RGSW_ciphertext &C) {
    std::vector<ciphertext_t> in = flat(ctx, c), g, out(2 * (size_t)ctx.N());
    flat_rgsw(ctx, C, g);
    fheacc::check(fheacc_external_product(ctx.handle(), in.data(), g.data(), 0, out.data(), 1));
    return unflat(ctx.N(), out.data());
}

RLWE_ciphertext CMUX(const fheacc::Context &ctx, const RLWE_ciphertext &c0, const RLWE_ciphertext &c1,
                     const RGSW_ciphertext &SEL) {
    std::vector<ciphertext_t> x0 = flat(ctx, c0), x1 = flat(ctx, c1), g, out(2 * (size_t)ctx.N());
    flat_rgsw(ctx, SEL, g);
    fheacc::check(fheacc_cmux(ctx.handle(), x0.data(), x1.data(), g.data(), 0, out.data(), 1));
    return unflat(ctx.N(), out.data());
}

RLWE_ciphertext blind_rotate(const fheacc::Context &ctx, const RLWE_ciphertext &c_in, const bool forward,
                             const std::vector<RGSW_ciphertext> &C) {
    std::vector<ciphertext_t> x = flat(ctx, c_in), g, out(2 * (size_t)ctx.N());
    for (const RGSW_ciphertext &r : C) flat_rgsw(ctx, r, g);
    fheacc::check(fheacc_blind_rotate_rlwe(ctx.handle(), x.data(), forward ? 1 : 0, g.data(), (int)C.size(),
                                           out.data(), 1));
    return unflat(ctx.N(), out.data());
}

RLWE_ciphertext blind_rotate_full(const fheacc::Context &ctx, const RLWE_ciphertext &c_in, const int b,
                                  const std::vector<int> &a, const std::vector<RGSW_ciphertext> &C) {
    if (a.size() != C.size()) {
        // RLWE.cpp:796-799
        printf("The sizes of rotation vector and control vector are not equal!!!\n");
        exit(1);
    }
    std::vector<ciphertext_t> x = flat(ctx, c_in), g, out(2 * (size_t)ctx.N());
    for (const RGSW_ciphertext &r : C) flat_rgsw(ctx, r, g);
    std::vector<int32_t> av(a.begin(), a.end());
    fheacc::check(fheacc_blind_rotate_full(ctx.handle(), x.data(), b, av.data(), g.empty() ? nullptr : g.data(),
                                           (int)C.size(), out.data(), 1));
    return unflat(ctx.N(), out.data());
}

}  // namespace RLWE

namespace application {

substitute_key::substitute_key(const fheacc::Context &ctx, const RLWE::RLWE_secretkey &sk, uint64_t seed)
    : keys(std::make_shared<fheacc::KeySet>(ctx)) {
    fheacc::check(fheacc_keys_generate_subs(keys->handle(), seed, sk.a.data()));
}

std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> expand_RLWE(const RLWE::RLWE_ciphertext &c1,
                                                                const substitute_key &subs_key) {
    if (!subs_key.keys) throw fheacc::Error(-1, "substitute key not generated");
    const fheacc::Context &ctx = subs_key.keys->context();
    const int N = ctx.N();
    std::vector<ciphertext_t> in = RLWE::flat(ctx, c1), out((size_t)N * 2 * N);
    fheacc::check(fheacc_expand_rlwe(ctx.handle(), subs_key.keys->handle(), in.data(), out.data(), 1));
    auto res = std::make_unique<std::vector<RLWE::RLWE_ciphertext>>(N);
    for (int k = 0; k < N; k++) (*res)[k] = RLWE::unflat(N, out.data() + (size_t)k * 2 * N);
    return res;
}

std::shared_ptr<std::vector<RLWE::RGSW_ciphertext>> homo_expand(
    const std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> &packed_bits, const RLWE::RGSW_ciphertext &enc_sec,
    const substitute_key &subs_key) {
    if (!subs_key.keys) throw fheacc::Error(-1, "substitute key not generated");
    const fheacc::Context &ctx = subs_key.keys->context();
    const int N = ctx.N();
    const uint32_t d = ctx.info().digitG;
    if (!packed_bits || packed_bits->size() != d) throw fheacc::Error(-1, "homo_expand takes digitG packed RLWEs");
    std::vector<ciphertext_t> in, g;
    for (const RLWE::RLWE_ciphertext &c : *packed_bits) {
        std::vector<ciphertext_t> w = RLWE::flat(ctx, c);
        in.insert(in.end(), w.begin(), w.end());
    }
    RLWE::flat_rgsw(ctx, enc_sec, g);
    const size_t gw = (size_t)d * 4 * N;
    std::vector<ciphertext_t> out((size_t)N * gw);
    fheacc::check(fheacc_homo_expand(ctx.handle(), subs_key.keys->handle(), in.data(), g.data(), out.data(), 1));
    auto res = std::make_shared<std::vector<RLWE::RGSW_ciphertext>>(N);
    for (int i = 0; i < N; i++) (*res)[i] = RLWE::unflat_rgsw(ctx, out.data() + (size_t)i * gw);
    return res;
}

}  // namespace application

namespace fheacc {

std::vector<LWE::LWE_ciphertext> eval_bootstrap_batch(const std::vector<LWE::LWE_ciphertext> &c1,
                                                      const std::vector<LWE::LWE_ciphertext> &c2,
                                                      const LWE::keyswitch_key &kskey, const std::vector<GATES> &gates,
                                                      int *rc) {
    *rc = 0;
    if (!kskey.keys || c1.size() != c2.size() || c1.size() != gates.size()) {
        *rc = -1;
        return std::vector<LWE::LWE_ciphertext>(c1.size());
    }
    const Context &ctx = kskey.keys->context();
    const int n = ctx.n();
    const size_t B = c1.size(), lw = (size_t)n + 1;
    std::vector<ciphertext_t> w1(B * lw), w2(B * lw), out(B * lw);
    std::vector<int32_t> g(B);
    for (size_t i = 0; i < B; i++) {
        if ((int)c1[i].a.size() != n || (int)c2[i].a.size() != n) {
            *rc = -1;
            return std::vector<LWE::LWE_ciphertext>(B);
        }
        std::copy(c1[i].a.begin(), c1[i].a.end(), w1.begin() + i * lw);
        w1[i * lw + n] = c1[i].b;
        std::copy(c2[i].a.begin(), c2[i].a.end(), w2.begin() + i * lw);
        w2[i * lw + n] = c2[i].b;
        g[i] = (int32_t)gates[i];
    }
    int r = fheacc_bootstrap_batch(ctx.handle(), kskey.keys->handle(), g.data(), w1.data(), w2.data(), out.data(),
                                   (int64_t)B, nullptr);
    std::vector<LWE::LWE_ciphertext> res(B);
    if (r != 0) {
        fprintf(stderr, "eval_bootstrap_batch: %s\n", fheacc_last_error());
        *rc = r;
        return res;
    }
    for (size_t i = 0; i < B; i++) {
        res[i].a.assign(out.begin() + i * lw, out.begin() + i * lw + n);
        res[i].b = out[i * lw + n];
    }
    return res;
}

}  // namespace fheacc

namespace FPGA {

namespace {

std::mutex order_mu;  // one accelerator-order call at a time changes a context's order

std::vector<LWE::LWE_ciphertext> in_order(const std::vector<LWE::LWE_ciphertext> &c1,
                                          const std::vector<LWE::LWE_ciphertext> &c2, const LWE::keyswitch_key &kskey,
                                          const std::vector<GATES> &gate, size_t want, int mode, const char *name,
                                          int *rc) {
    *rc = 0;
    if (c1.size() != want || c2.size() != want || gate.size() != want) {
        fprintf(stderr, "%s: input vectors must hold %zu ciphertexts\n", name, want);
        *rc = -1;
        return std::vector<LWE::LWE_ciphertext>(want);
    }
    if (!kskey.keys) {
        fprintf(stderr, "%s: no key set\n", name);
        *rc = -1;
        return std::vector<LWE::LWE_ciphertext>(want);
    }
    fheacc_ctx *h = kskey.keys->context().handle();
    std::lock_guard<std::mutex> lock(order_mu);
    int prev = FHEACC_EVAL_SOFTWARE;
    if (fheacc_ctx_get_eval_mode(h, &prev) != 0 || fheacc_ctx_set_eval_mode(h, mode) != 0) {
        fprintf(stderr, "%s: %s\n", name, fheacc_last_error());
        *rc = -1;
        return std::vector<LWE::LWE_ciphertext>(want);
    }
    std::vector<LWE::LWE_ciphertext> out = fheacc::eval_bootstrap_batch(c1, c2, kskey, gate, rc);
    fheacc_ctx_set_eval_mode(h, prev);
    return out;
}

}  // namespace

LWE::LWE_ciphertext fpga_eval_bootstrap_x1(const LWE::LWE_ciphertext &c1, const LWE::LWE_ciphertext &c2,
                                           const LWE::keyswitch_key &kskey, const GATES gate, int *rc) {
    return in_order({c1}, {c2}, kskey, {gate}, 1, FHEACC_EVAL_FPGA_X1, "fpga_eval_bootstrap_x1", rc)[0];
}

std::vector<LWE::LWE_ciphertext> fpga_eval_bootstrap_x4(const std::vector<LWE::LWE_ciphertext> &c1,
                                                        const std::vector<LWE::LWE_ciphertext> &c2,
                                                        const LWE::keyswitch_key &kskey,
                                                        const std::vector<GATES> gate, int *rc) {
    return in_order(c1, c2, kskey, gate, 4, FHEACC_EVAL_FPGA_XN, "fpga_eval_bootstrap_x4", rc);
}

std::vector<LWE::LWE_ciphertext> fpga_eval_bootstrap_x8(const std::vector<LWE::LWE_ciphertext> &c1,
                                                        const std::vector<LWE::LWE_ciphertext> &c2,
                                                        const LWE::keyswitch_key &kskey,
                                                        const std::vector<GATES> gate, int *rc) {
    return in_order(c1, c2, kskey, gate, 8, FHEACC_EVAL_FPGA_XN, "fpga_eval_bootstrap_x8", rc);
}

std::vector<LWE::LWE_ciphertext> fpga_eval_bootstrap_x12(const std::vector<LWE::LWE_ciphertext> &c1,
                                                         const std::vector<LWE::LWE_ciphertext> &c2,
                                                         const LWE::keyswitch_key &kskey,
                                                         const std::vector<GATES> gate, int *rc) {
    return in_order(c1, c2, kskey, gate, 12, FHEACC_EVAL_FPGA_XN, "fpga_eval_bootstrap_x12", rc);
}

std::vector<LWE::LWE_ciphertext> fpga_eval_bootstrap_x16(const std::vector<LWE::LWE_ciphertext> &c1,
                                                         const std::vector<LWE::LWE_ciphertext> &c2,
                                                         const LWE::keyswitch_key &kskey,
                                                         const std::vector<GATES> gate, int *rc) {
    return in_order(c1, c2, kskey, gate, 10, FHEACC_EVAL_FPGA_XN, "fpga_eval_bootstrap_x16", rc);
}

std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> fpga_expand_RLWE(const RLWE::RLWE_ciphertext &c1,
                                                                     const application::substitute_key &subs_key,
                                                                     int *rc) {
    *rc = 0;
    try {
        return application::expand_RLWE(c1, subs_key);
    } catch (const fheacc::Error &e) {
        fprintf(stderr, "fpga_expand_RLWE: %s\n", e.what());
        *rc = e.rc ? e.rc : -1;
        return nullptr;
    }
}

std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> fpga_expand_RLWE_cont_rd_wr(
    const RLWE::RLWE_ciphertext &c1, const application::substitute_key &subs_key, int *rc) {
    return fpga_expand_RLWE(c1, subs_key, rc);
}

std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> fpga_expand_RLWE_batch_rd_wr(
    const RLWE::RLWE_ciphertext &c1, const application::substitute_key &subs_key, int *rc) {
    return fpga_expand_RLWE(c1, subs_key, rc);
}

std::shared_ptr<std::vector<RLWE::RGSW_ciphertext>> fpga_homo_expand(
    const std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> &packed_bits, const RLWE::RGSW_ciphertext &enc_sec,
    const application::substitute_key &subs_key, int *rc) {
    *rc = 0;
    try {
        return application::homo_expand(packed_bits, enc_sec, subs_key);
    } catch (const fheacc::Error &e) {
        fprintf(stderr, "fpga_homo_expand: %s\n", e.what());
        *rc = e.rc ? e.rc : -1;
        return nullptr;
    }
}

}  // namespace FPGA
