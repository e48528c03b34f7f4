// fhe_acc.hpp -- C++ host API over libfheacc.so, shaped like the reference's
// host API so that its callers (FHE_ACC_runtime/apps/*.cpp, src/test.cpp)
// port by swapping includes and passing a Context where the reference reads
// its global parameter block.
//
//   reference                                   here
//   params.h globals n, q, N, Q, BG, Br, Bks... fheacc::Context(Params)
//   LWE::LWE_secretkey(RNG_uniform&)            LWE::LWE_secretkey(ctx, seed, binary)
//   RLWE::RLWE_secretkey(RNG_uniform&)          RLWE::RLWE_secretkey(ctx, seed)
//   RLWE::bootstrap_key(s, z, uni, norm)        RLWE::bootstrap_key(ctx, s, z, seed)
//   LWE::keyswitch_key(z, s, uni, norm)         LWE::keyswitch_key(btkey) (same device set)
//   LWE::LWE_encrypt(s, p, uni, norm)           LWE::LWE_encrypt(ctx, s, p, seed, index)
//   LWE::LWE_decrypt(s, c)                      LWE::LWE_decrypt(ctx, s, c)
//   LWE::LWE_evaluate / LWE_eval_NOT            same names (host, LWE.cpp:203-231)
//   RLWE::eval_bootstrap(c1, c2, bt, ks, gate)  same name and arguments
//   FPGA::fpga_eval_bootstrap_x{4,8,12,16}      fheacc::eval_bootstrap_batch (any B, int* rc)
//   FPGA::fpga_eval_bootstrap_x1..x16 (bar, fd) FPGA::fpga_eval_bootstrap_x1..x16 (no bar/fd),
//                                               the accelerator's step order and batch sizes
//   RLWE::RLWE_ciphertext / RGSW_ciphertext     same members (a, b, NTT_form; c_text[h][g])
//   RLWE::RLWE_mult_RGSW(c, C)                  RLWE::RLWE_mult_RGSW(ctx, c, C)
//   RLWE::CMUX / blind_rotate / blind_rotate_full  same names, ctx first
//   application::substitute_key(sk, uni, norm)  application::substitute_key(ctx, sk, seed)
//   application::expand_RLWE(c, subs_key)       same name and arguments
//   application::homo_expand(bits, enc_sec, sk) same name and arguments
//   RLWE::RLWE_encrypt(s, p, uni, norm)         RLWE::RLWE_encrypt(ctx, s, p, seed, index)
//   RLWE::RLWE_decrypt(s, c, t)                 RLWE::RLWE_decrypt(ctx, s, c, t)
//   RGSW_ciphertext(m, z, uni, norm)            RLWE::RGSW_encrypt(ctx, z, m, seed, index)
//   FPGA::fpga_expand_RLWE{,_cont_rd_wr,_batch_rd_wr}, fpga_homo_expand
//     (bar, rc, fds; keys DMA'd to the card)    same names (subs key / RGSW(-s) passed, rc)
//
// Keys live on the device; the classes hold shared handles.  eval_bootstrap
// on identical operands prints the reference's message and exits the process
// as RLWE.cpp:727-730 does; the batched call reports it through *rc = -1
// like the accelerator driver (myfpga.cpp:1011-1022).
#pragma once
#include <stdint.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fheacc.h"

typedef uint64_t message_t;
typedef uint64_t ciphertext_t;

// params.h:107
enum GATES { OR, AND, NOR, NAND, XOR, XNOR };

namespace fheacc {

struct Params {
    int n = 500;
    ciphertext_t q = 1024;
    int N = 1024;
    ciphertext_t Q = 0;  // 0: previous_prime(first_prime(nbit, N), N)
    int nbit = 27;
    uint32_t BGbits = 7;
    ciphertext_t Br = 32;
    ciphertext_t Bks = 32;
    uint32_t Bks_R_bits = 9;
    double stddev = 3.19;
    ciphertext_t E_rlwe = 18;
};

class Error : public std::runtime_error {
   public:
    Error(int rc, const std::string &what) : std::runtime_error(what), rc(rc) {}
    int rc;
};

class Context {
   public:
    explicit Context(const Params &p, int device = 0);
    ~Context();
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    fheacc_ctx *handle() const { return h_; }
    const fheacc_info &info() const { return info_; }
    const Params &params() const { return p_; }
    int n() const { return p_.n; }
    ciphertext_t q() const { return p_.q; }
    int N() const { return p_.N; }
    ciphertext_t Q() const { return info_.Q; }

   private:
    fheacc_ctx *h_ = nullptr;
    Params p_;
    fheacc_info info_{};
};

// A client's device-resident bootstrap + key-switch keys.
class KeySet {
   public:
    explicit KeySet(const Context &ctx);
    ~KeySet();
    KeySet(const KeySet &) = delete;
    KeySet &operator=(const KeySet &) = delete;
    fheacc_keys *handle() const { return h_; }
    const Context &context() const { return ctx_; }

   private:
    const Context &ctx_;
    fheacc_keys *h_ = nullptr;
};

void check(int rc);

}  // namespace fheacc

namespace RLWE {
class RLWE_secretkey;
class bootstrap_key;
}  // namespace RLWE

namespace LWE {

class LWE_plaintext {
   public:
    LWE_plaintext() : m(0) {}
    LWE_plaintext(message_t m_in) : m(m_in) {}
    message_t m;
};

class LWE_ciphertext {
   public:
    LWE_ciphertext() = default;
    explicit LWE_ciphertext(int n) : a(n, 0), b(0) {}
    bool operator==(const LWE_ciphertext &o) const { return b == o.b && a == o.a; }
    std::vector<ciphertext_t> a;
    ciphertext_t b = 0;
};

class LWE_secretkey : public LWE_ciphertext {
   public:
    LWE_secretkey() = default;
    // seeded (the reference draws from std::random_device, RNG.cpp:5-6)
    LWE_secretkey(const fheacc::Context &ctx, uint64_t seed, bool binary = true);
};

class keyswitch_key {
   public:
    keyswitch_key() = default;
    // the key-switch key is generated with, and lives beside, the bootstrap key
    explicit keyswitch_key(const RLWE::bootstrap_key &bt);
    std::shared_ptr<fheacc::KeySet> keys;
};

LWE_ciphertext LWE_encrypt(const fheacc::Context &ctx, const LWE_secretkey &s, const LWE_plaintext &p, uint64_t seed,
                           uint64_t index);
LWE_plaintext LWE_decrypt(const fheacc::Context &ctx, const LWE_secretkey &s, const LWE_ciphertext &c);
LWE_ciphertext LWE_evaluate(const fheacc::Context &ctx, const LWE_ciphertext &c1, const LWE_ciphertext &c2,
                            GATES gate);
LWE_ciphertext LWE_eval_NOT(const fheacc::Context &ctx, const LWE_ciphertext &c);

}  // namespace LWE

namespace RLWE {

class RLWE_secretkey {
   public:
    RLWE_secretkey() = default;
    RLWE_secretkey(const fheacc::Context &ctx, uint64_t seed);
    std::vector<ciphertext_t> a;  // time domain, mod Q
};

class bootstrap_key {
   public:
    bootstrap_key() = default;
    // Generates btkey (RLWE.cpp:579-606) and the LWE key-switch key
    // (LWE.cpp:55-106) on the device from a seed.
    bootstrap_key(const fheacc::Context &ctx, const LWE::LWE_secretkey &s, const RLWE_secretkey &z, uint64_t seed);
    // Uploads keys in the reference layouts (flattened btkey / kskey).
    bootstrap_key(const fheacc::Context &ctx, const std::vector<ciphertext_t> &btkey_flat,
                  const std::vector<ciphertext_t> &kskey_flat);
    std::shared_ptr<fheacc::KeySet> keys;
};

LWE::LWE_ciphertext eval_bootstrap(const LWE::LWE_ciphertext &c1, const LWE::LWE_ciphertext &c2,
                                   const bootstrap_key &btkey, const LWE::keyswitch_key &kskey, const GATES gate);

// RLWE.h:127-159: the two polynomials, NTT domain (the GPU calls take and
// return the NTT form)
class RLWE_ciphertext {
   public:
    RLWE_ciphertext() = default;
    explicit RLWE_ciphertext(int N) : a(N, 0), b(N, 0) {}
    std::vector<ciphertext_t> a, b;
    bool NTT_form = true;
};

// RLWE.h:313-329: c_text[h][g], h < digitG, g in {0, 1}
class RGSW_ciphertext {
   public:
    RGSW_ciphertext() = default;
    explicit RGSW_ciphertext(std::vector<std::vector<RLWE_ciphertext>> rows) : c_text(std::move(rows)) {}
    std::vector<std::vector<RLWE_ciphertext>> c_text;
};

// RLWE.h RLWE_plaintext: the message polynomial, here kept in the time domain
// (the reference's RLWE_plaintext(m, false) constructor transforms it and
// RLWE_decrypt's denoise transforms it back, RLWE.cpp:71-84,138-155)
class RLWE_plaintext {
   public:
    RLWE_plaintext() = default;
    explicit RLWE_plaintext(std::vector<ciphertext_t> m_time) : m(std::move(m_time)) {}
    std::vector<ciphertext_t> m;  // time domain, mod Q
};

// RLWE.h RGSW_plaintext
class RGSW_plaintext {
   public:
    RGSW_plaintext() = default;
    explicit RGSW_plaintext(RLWE_plaintext m_in) : m(std::move(m_in)) {}
    RLWE_plaintext m;
};

// RLWE.cpp:324-344; seeded (stream `index` of the library's sampler)
RLWE_ciphertext RLWE_encrypt(const fheacc::Context &ctx, const RLWE_secretkey &s, const RLWE_plaintext &p,
                             uint64_t seed, uint64_t index);
// RLWE.cpp:362-376: RLWE_decrypt(s, c, t), the message rounded to [0, t]
RLWE_plaintext RLWE_decrypt(const fheacc::Context &ctx, const RLWE_secretkey &s, const RLWE_ciphertext &c,
                            ciphertext_t t);
// RLWE.cpp:561-577: RGSW_ciphertext(m, z, uni, norm), seeded
RGSW_ciphertext RGSW_encrypt(const fheacc::Context &ctx, const RLWE_secretkey &z, const RGSW_plaintext &m,
                             uint64_t seed, uint64_t index);

// RLWE.cpp:642-704
RLWE_ciphertext RLWE_mult_RGSW(const fheacc::Context &ctx, const RLWE_ciphertext &c, const RGSW_ciphertext &C);
// RLWE.cpp:788-793
RLWE_ciphertext CMUX(const fheacc::Context &ctx, const RLWE_ciphertext &c0, const RLWE_ciphertext &c1,
                     const RGSW_ciphertext &SEL);
// RLWE.cpp:822-838 (any number of RGSWs up to log2 N; the reference fixes 11)
RLWE_ciphertext blind_rotate(const fheacc::Context &ctx, const RLWE_ciphertext &c_in, const bool forward,
                             const std::vector<RGSW_ciphertext> &C);
// RLWE.cpp:795-820
RLWE_ciphertext blind_rotate_full(const fheacc::Context &ctx, const RLWE_ciphertext &c_in, const int b,
                                  const std::vector<int> &a, const std::vector<RGSW_ciphertext> &C);

}  // namespace RLWE

namespace application {

// app.h substitute_key: the RLWE key switch keys from sk(X^(N/2^i + 1)) to
// sk(X), i < log2 N, generated on the device from a seed and kept resident
class substitute_key {
   public:
    substitute_key() = default;
    substitute_key(const fheacc::Context &ctx, const RLWE::RLWE_secretkey &sk, uint64_t seed);
    std::shared_ptr<fheacc::KeySet> keys;
};

// app.cpp:26-97
std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> expand_RLWE(const RLWE::RLWE_ciphertext &c1,
                                                                const substitute_key &subs_key);
// app.cpp:78-97: digitG packed RLWEs -> N RGSWs, RGSW i row h =
// (u_h[i] (x) enc_sec, u_h[i]) with u_h = expand_RLWE(packed_bits[h])
std::shared_ptr<std::vector<RLWE::RGSW_ciphertext>> homo_expand(
    const std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> &packed_bits, const RLWE::RGSW_ciphertext &enc_sec,
    const substitute_key &subs_key);

}  // namespace application

namespace fheacc {

// FPGA::fpga_eval_bootstrap_x{4,8,12,16} with any batch size; *rc = 0, -1, -ENOMEM.
// Evaluates in the context's order (fheacc_ctx_set_eval_mode; by default the
// software order of RLWE::eval_bootstrap).
std::vector<LWE::LWE_ciphertext> eval_bootstrap_batch(const std::vector<LWE::LWE_ciphertext> &c1,
                                                      const std::vector<LWE::LWE_ciphertext> &c2,
                                                      const LWE::keyswitch_key &kskey, const std::vector<GATES> &gates,
                                                      int *rc);

}  // namespace fheacc

namespace FPGA {

// The accelerator's gate calls (myfpga.h:141-401) with the card's BAR handle
// and DMA descriptor dropped: the GPU context is the one kskey's device key
// set belongs to.  Each evaluates in the card's step order, which differs
// from RLWE::eval_bootstrap in noise and therefore in words:
//   x1  the fused BOOTSTRAP_INIT at step (0, 0) multiplies by btkey[r][0][0]
//       even for r = 0, later residue-0 steps are skipped (myfpga.cpp:1003-1127);
//   xN  every step multiplies, by btkey[0] = RGSW(1) for residue 0
//       (myfpga.cpp:1130-1660).
// Batch sizes are the reference's: 4, 8, 12, and 10 for x16 (myfpga.cpp:
// 1140,1282,1408,1541); other sizes give *rc = -1.  The order is set on the
// shared context for the call and restored afterwards: other gate
// bootstraps on the same context must not run concurrently with these.
LWE::LWE_ciphertext fpga_eval_bootstrap_x1(const LWE::LWE_ciphertext &c1, const LWE::LWE_ciphertext &c2,
                                           const LWE::keyswitch_key &kskey, const GATES gate, int *rc);
std::vector<LWE::LWE_ciphertext> fpga_eval_bootstrap_x4(const std::vector<LWE::LWE_ciphertext> &c1,
                                                        const std::vector<LWE::LWE_ciphertext> &c2,
                                                        const LWE::keyswitch_key &kskey,
                                                        const std::vector<GATES> gate, int *rc);
std::vector<LWE::LWE_ciphertext> fpga_eval_bootstrap_x8(const std::vector<LWE::LWE_ciphertext> &c1,
                                                        const std::vector<LWE::LWE_ciphertext> &c2,
                                                        const LWE::keyswitch_key &kskey,
                                                        const std::vector<GATES> gate, int *rc);
std::vector<LWE::LWE_ciphertext> fpga_eval_bootstrap_x12(const std::vector<LWE::LWE_ciphertext> &c1,
                                                         const std::vector<LWE::LWE_ciphertext> &c2,
                                                         const LWE::keyswitch_key &kskey,
                                                         const std::vector<GATES> gate, int *rc);
std::vector<LWE::LWE_ciphertext> fpga_eval_bootstrap_x16(const std::vector<LWE::LWE_ciphertext> &c1,
                                                         const std::vector<LWE::LWE_ciphertext> &c2,
                                                         const LWE::keyswitch_key &kskey,
                                                         const std::vector<GATES> gate, int *rc);

// The accelerator's expansion calls (myfpga.h:342-359).  The card reads the
// subs key and RGSW(-s) it was sent beforehand (dma_write_subs_key,
// dma_write_RGSW_enc_sk); here they are the resident substitute_key and an
// RGSW passed in.  The card runs RLWESUBS and the host the additions and
// rotations (myfpga.cpp:1662-2276), the CPU form's words: the three
// expansion variants differ only in how they stream RLWEs over PCIe, so all
// three are the one device expansion.  *rc = 0 or -1 (the failure printed);
// a failed call returns a null pointer, as the card's calls do.
std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> fpga_expand_RLWE(const RLWE::RLWE_ciphertext &c1,
                                                                     const application::substitute_key &subs_key,
                                                                     int *rc);
std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> fpga_expand_RLWE_cont_rd_wr(
    const RLWE::RLWE_ciphertext &c1, const application::substitute_key &subs_key, int *rc);
std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> fpga_expand_RLWE_batch_rd_wr(
    const RLWE::RLWE_ciphertext &c1, const application::substitute_key &subs_key, int *rc);
std::shared_ptr<std::vector<RLWE::RGSW_ciphertext>> fpga_homo_expand(
    const std::unique_ptr<std::vector<RLWE::RLWE_ciphertext>> &packed_bits, const RLWE::RGSW_ciphertext &enc_sec,
    const application::substitute_key &subs_key, int *rc);

}  // namespace FPGA
