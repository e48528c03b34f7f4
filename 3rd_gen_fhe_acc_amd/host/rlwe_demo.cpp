// rlwe_demo.cpp -- the reference's RLWE-side host API (application::
// expand_RLWE, app.cpp:26-97; RLWE::RLWE_mult_RGSW, CMUX, blind_rotate,
// blind_rotate_full, RLWE.cpp:642-838) driven through the reference-shaped
// C++ layer on words read from files, results written back as raw uint64.
//
// usage: rlwe_demo DIR N Q BGbits Bks_R_bits sk_seed subs_seed b
//   DIR/in.bin    2 RLWE (c0, c1), [2][2][N] words, NTT domain
//   DIR/rgsw.bin  log2 N RGSW in the reference layout [h][g][2][N]
//   DIR/a.bin     log2 N int32 rotation factors (blind_rotate_full)
// writes DIR/{expand,mult,cmux,br_fwd,br_bwd,brf}.bin
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "fhe_acc.hpp"

template <typename T>
static std::vector<T> read_all(const std::string &path, size_t count) {
    std::vector<T> v(count);
    FILE *f = fopen(path.c_str(), "rb");
    if (!f || fread(v.data(), sizeof(T), count, f) != count) {
        fprintf(stderr, "cannot read %s\n", path.c_str());
        exit(2);
    }
    fclose(f);
    return v;
}

static void write_rlwe(const std::string &path, const std::vector<RLWE::RLWE_ciphertext> &cs) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) exit(2);
    for (const auto &c : cs) {
        fwrite(c.a.data(), 8, c.a.size(), f);
        fwrite(c.b.data(), 8, c.b.size(), f);
    }
    fclose(f);
}

static RLWE::RLWE_ciphertext rlwe_at(const std::vector<ciphertext_t> &w, size_t off, int N) {
    RLWE::RLWE_ciphertext c(N);
    c.a.assign(w.begin() + off, w.begin() + off + N);
    c.b.assign(w.begin() + off + N, w.begin() + off + 2 * N);
    return c;
}

int main(int argc, char **argv) {
    if (argc != 9) {
        fprintf(stderr, "usage: rlwe_demo DIR N Q BGbits Bks_R_bits sk_seed subs_seed b\n");
        return 2;
    }
    const std::string dir = argv[1];
    fheacc::Params p;
    p.n = 8;
    p.N = atoi(argv[2]);
    p.q = 2 * (ciphertext_t)p.N < 1024 ? 2 * (ciphertext_t)p.N : 1024;
    p.Q = strtoull(argv[3], nullptr, 0);
    p.BGbits = (uint32_t)atoi(argv[4]);
    p.Bks_R_bits = (uint32_t)atoi(argv[5]);
    fheacc::Context ctx(p);
    const int N = ctx.N(), logN = __builtin_ctz((unsigned)N);
    const uint32_t d = ctx.info().digitG;
    RLWE::RLWE_secretkey sk(ctx, strtoull(argv[6], nullptr, 0));
    application::substitute_key subs(ctx, sk, strtoull(argv[7], nullptr, 0));
    const int b = atoi(argv[8]);

    const std::vector<ciphertext_t> in = read_all<ciphertext_t>(dir + "/in.bin", 4 * (size_t)N);
    const std::vector<ciphertext_t> gw = read_all<ciphertext_t>(dir + "/rgsw.bin", (size_t)logN * d * 4 * N);
    const std::vector<int32_t> a32 = read_all<int32_t>(dir + "/a.bin", logN);
    RLWE::RLWE_ciphertext c0 = rlwe_at(in, 0, N), c1 = rlwe_at(in, 2 * N, N);
    std::vector<RLWE::RGSW_ciphertext> C(logN);
    for (int i = 0; i < logN; i++) {
        std::vector<std::vector<RLWE::RLWE_ciphertext>> rows(d);
        for (uint32_t h = 0; h < d; h++)
            for (int g = 0; g < 2; g++) rows[h].push_back(rlwe_at(gw, (((size_t)i * d + h) * 2 + g) * 2 * N, N));
        C[i] = RLWE::RGSW_ciphertext(rows);
    }
    std::vector<int> a(a32.begin(), a32.end());

    auto expanded = application::expand_RLWE(c0, subs);
    write_rlwe(dir + "/expand.bin", *expanded);
    write_rlwe(dir + "/mult.bin", {RLWE::RLWE_mult_RGSW(ctx, c0, C[0])});
    write_rlwe(dir + "/cmux.bin", {RLWE::CMUX(ctx, c0, c1, C[0])});
    write_rlwe(dir + "/br_fwd.bin", {RLWE::blind_rotate(ctx, c0, true, C)});
    write_rlwe(dir + "/br_bwd.bin", {RLWE::blind_rotate(ctx, c0, false, C)});
    write_rlwe(dir + "/brf.bin", {RLWE::blind_rotate_full(ctx, c0, b, a, C)});
    printf("RLWE DEMO DONE: N=%d d=%u expansions=%zu\n", N, d, expanded->size());
    return 0;
}
