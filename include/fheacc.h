/*
 * fheacc.h -- C ABI of the MI355X gate-bootstrapping library (libfheacc.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, int return codes
 * (0 = ok, -1 = failure, -ENOMEM = allocation failure, the convention of the
 * reference's accelerator driver, FHE_ACC_runtime/src/myfpga.cpp:1011-1022).
 * On failure fheacc_last_error() describes the cause (thread-local).
 *
 * Word layouts (uint64 words, canonical in [0, modulus), the reference's
 * std::vector<uint64_t> contents):
 *   LWE ciphertext   [n+1]             a[0..n-1] then b         (LWE.h:36-64)
 *   RLWE ciphertext  [2][N]            a then b, NTT domain     (RLWE.h:127-159)
 *   RGSW ciphertext  [digitG][2][2][N] c_text[h][g].{a,b}       (RLWE.h:313-329)
 *   bootstrap key    [Br][digitR][n] RGSW = btkey[r][j][k]      (RLWE.cpp:579-606)
 *   LWE ks key       [Bks][digitKS][N] LWE(n) mod Q = kskey[d][j][i] (LWE.cpp:55-106)
 *   RLWE ks key      [digitKS_R] RLWE = RLWE_kskey[l]           (RLWE.cpp:293-322)
 *   subs key         [log2 N][digitKS_R] RLWE = subs_key[i]     (app.cpp:4-24)
 * Gates follow enum GATES {OR, AND, NOR, NAND, XOR, XNOR} (params.h:107).
 *
 * Functions named *_dev take device pointers and are asynchronous on the
 * context's stream; the others take host pointers and return after the
 * results are in host memory.
 */
#ifndef FHEACC_H
#define FHEACC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FHEACC_OK 0
#define FHEACC_ERR (-1)
#define FHEACC_ENOMEM (-12)

typedef struct fheacc_ctx fheacc_ctx;
typedef struct fheacc_keys fheacc_keys;

/* Parameter set.  Replaces the global block every app assigns
 * (FHE_ACC_runtime/include/params.h:13-134, apps/fhe_wo_fpga.cpp:35-162). */
typedef struct fheacc_params {
    int32_t  n;           /* LWE dimension */
    uint64_t q;           /* LWE modulus: power of two dividing 2N */
    int32_t  N;           /* ring degree: power of two, 16..2048 */
    uint64_t Q;           /* RLWE prime, Q = 1 mod 2N; 0 = derive from nbit */
    int32_t  nbit;        /* Q = previous_prime(first_prime(nbit, N), N) when Q == 0 */
    uint32_t BGbits;      /* gadget base 2^BGbits */
    uint64_t Br;          /* blind-rotation radix */
    uint64_t Bks;         /* RLWE->LWE key-switch base, 2..256 */
    uint32_t Bks_R_bits;  /* RLWE->RLWE key-switch base 2^Bks_R_bits */
    double   stddev;      /* key-generation Gaussian sigma (3.19) */
    uint64_t E_rlwe;      /* key-generation noise bound (18) */
} fheacc_params;

/* Derived values (params.h globals computed as the apps compute them). */
typedef struct fheacc_info {
    uint64_t Q, iN, ROU;
    uint32_t digitG, digitR, digitKS, digitKS_R;
    uint32_t narrow;        /* 1: 32-bit device words (Q < 2^30) */
    uint32_t harvey;        /* 1: 32-bit words with Harvey's forward butterflies (28- to 30-bit Q) */
    uint64_t gate_constant[6];
    uint64_t bsk_words;     /* words of the reference-layout bootstrap key */
    uint64_t ksk_words;     /* words of the reference-layout LWE ks key */
    uint64_t subs_words;    /* words of the reference-layout subs key */
    uint64_t bsk_device_bytes, ksk_device_bytes, subs_device_bytes;
    uint32_t lazy_inverse_stages; /* inverse NTT passes of up to this many stages leave their
                                     sums unreduced: 2^(stages+1) Q fits the device word */
    uint32_t reserved;
} fheacc_info;

/* Capture buffers for intermediate words (each may be NULL).  Host or device
 * memory to match the call. */
typedef struct fheacc_capture {
    uint64_t *acc;   /* [B][2][N] accumulator after blind rotation (NTT domain) */
    uint64_t *ks;    /* [B][n+1] LWE mod Q after the key switch */
} fheacc_capture;

/* ---- context ---------------------------------------------------------
 * Replaces the parameter derivation of the apps (Q, iN, ROU, tables:
 * fhe_wo_fpga.cpp:89-113, nbt.cpp:209-346, params.cpp:149-175) and the
 * accelerator configuration FPGA::OCL_config_wr_rd + BAR1_ROU_table_*_wr
 * (myfpga.h:141-180, myfpga.cpp:82-594). */
int fheacc_ctx_create(const fheacc_params *params, int device, fheacc_ctx **out);
/* Releases the caller's handle.  Key sets, graphs and servers made from the
 * context hold references of their own: its device memory is freed when the
 * last of them is destroyed, so they may be destroyed in any order. */
int fheacc_ctx_destroy(fheacc_ctx *ctx);
int fheacc_ctx_info(const fheacc_ctx *ctx, fheacc_info *out);
/* Launch on a caller-owned hipStream_t from now on (NULL: the null stream).
 * A new context launches on a stream of its own. */
int fheacc_ctx_set_stream(fheacc_ctx *ctx, void *hip_stream);
int fheacc_sync(fheacc_ctx *ctx);
const char *fheacc_last_error(void);
/* Device-side bounds checks.  The checked library, lib/libfheacc_checked.so
 * (every source built with -DFHEACC_DEVICE_CHECKS=1), checks each LDS access
 * of the transforms and external products against the workgroup's LDS
 * allocation, and each gate workgroup's ciphertext against its launch's
 * range.  A failed check sets a bit and the kernel goes on: it never traps.
 * No counterpart in the reference, whose only checks are host compiler
 * warnings (FHE_ACC_runtime/CMakeLists.txt:24,42) and the DEBUG define's rejection bounds
 * (LWE.cpp:40-51).
 * fheacc_device_checks_enabled: 1 in the checked library, 0 otherwise.
 * fheacc_device_check_read: synchronizes the device, then stores in *bits the
 * FHEACC_DEVCHECK_* bits set since the last read and clears them (always 0
 * in the default library). */
#define FHEACC_DEVCHECK_LDS 1
#define FHEACC_DEVCHECK_CIPHERTEXT 2
int fheacc_device_checks_enabled(void);
int fheacc_device_check_read(int device, uint32_t *bits);
/* Per-kernel timing with HIP events on the launch stream (the role of the
 * reference's chrono prints around each phase, myfpga.cpp:1431-1525).
 * enable != 0 starts recording and clears the totals; fheacc_ctx_kernel_ms
 * synchronizes and returns total milliseconds and launch counts per kernel
 * class: [0] blind rotation, [1] LWE key switch, [2] external product,
 * [3] NTT. */
int fheacc_ctx_profile(fheacc_ctx *ctx, int enable);
int fheacc_ctx_kernel_ms(fheacc_ctx *ctx, double *ms4, int64_t *launches4);

/* Launch shapes of the blind-rotation kernel: threads per workgroup (128,
 * 256, 512, or 1024 when N >= 1024) and digit polys per LDS batch (0: as
 * many as fit).  A batch of 2d polys runs the product in place; d polys run
 * it in place in two batches (the digits of a, then of b) in half the LDS;
 * any other size uses a separate digit buffer beside the accumulator.  Every
 * shape gives the same words.  Once set, the CMUX chains use the same shape; until then
 * they run 1024 threads from N = 1024 up.  No
 * counterpart in the reference (the FPGA's pipeline depth is fixed,
 * myfpga.cpp:1533). */
int fheacc_ctx_set_tuning(fheacc_ctx *ctx, int br_threads, int br_polys_per_batch);
int fheacc_ctx_get_tuning(const fheacc_ctx *ctx, int *br_threads, int *br_polys_per_batch);
/* Schedule of the blind rotation for dense batches (the serving loop's
 * gathered launches always run one workgroup per ciphertext):
 *   persist = 0   one workgroup per ciphertext, the hardware dispatcher
 *                 refilling CUs as gates finish (the default);
 *   persist > 0, sync_window = 0  tiles: launches of min(persist,
 *                 occupancy) workgroups per CU each, back to back, so the
 *                 resident workgroups of a tile start together;
 *   persist > 0, sync_window > 0  one persistent launch of that many
 *                 workgroups per CU, each taking every G-th ciphertext;
 *                 workgroups of a group (workgroup index mod 8) keep within
 *                 sync_window steps of each other, counted across their
 *                 gates, on device progress words, so each step's keys are
 *                 fetched into L2 once for the group.  A wait gives up after
 *                 a bounded spin: pacing never changes words and never blocks.
 *                 At the benchmark's batch the tightest pacing cuts the
 *                 blind rotation's L2 misses 3.7x, but costs more time than
 *                 the misses do (the kernel is VALU-bound), so the
 *                 autotuner keeps persist = 0 there.
 * No counterpart in the reference: the accelerator streams one key per step
 * for its 12-14 ciphertexts in flight (myfpga.cpp:1446-1492, 1533).  The
 * autotuner picks it. */
int fheacc_ctx_set_schedule(fheacc_ctx *ctx, int persist, int sync_window);
int fheacc_ctx_get_schedule(const fheacc_ctx *ctx, int *persist, int *sync_window);
/* Autotuner: times the batched gate bootstrap on B synthetic ciphertexts
 * with `keys` for every candidate shape, keeps the fastest in the context,
 * and writes a JSON report of all candidates to report (may be NULL).  A
 * schedule other than persist = 0 is kept only when it is at least 3 %
 * faster than persist = 0 at the chosen shape (timings of one shape repeat
 * within 1-2 %). */
int fheacc_autotune(fheacc_ctx *ctx, fheacc_keys *keys, int64_t B, char *report, size_t report_len);

/* Evaluation order of the blind rotation for every later gate bootstrap on
 * the context (batches, graphs captured afterwards, the serving loop):
 *   FHEACC_EVAL_SOFTWARE  residue-0 steps skipped: RLWE::eval_bootstrap
 *                         (RLWE.cpp:754-767).  The default.
 *   FHEACC_EVAL_FPGA_X1   the accelerator's one-ciphertext form: step (0,0)
 *                         is the fused BOOTSTRAP_INIT and multiplies by
 *                         btkey[r][0][0] even for r = 0; later residue-0
 *                         steps are skipped (fpga_eval_bootstrap_x1,
 *                         myfpga.cpp:1030-1071).
 *   FHEACC_EVAL_FPGA_XN   the batched forms: every step multiplies, by
 *                         btkey[0] = RGSW(1) when the residue is 0
 *                         (fpga_eval_bootstrap_x4..x16, myfpga.cpp:1180-1225).
 * The FPGA orders add noise terms, so their words differ from the software
 * order's; each is bit-exact with its own order. */
#define FHEACC_EVAL_SOFTWARE 0
#define FHEACC_EVAL_FPGA_X1 1
#define FHEACC_EVAL_FPGA_XN 2
int fheacc_ctx_set_eval_mode(fheacc_ctx *ctx, int mode);
int fheacc_ctx_get_eval_mode(const fheacc_ctx *ctx, int *mode);

/* ---- keys (device resident, one set per client) -----------------------
 * Replaces RLWE::bootstrap_key / LWE::keyswitch_key /
 * application::substitute_key construction (RLWE.cpp:579-606,
 * LWE.cpp:55-106, app.cpp:4-24) and the key DMA to the card
 * (FPGA::dma_write_bootstrap_key, dma_write_subs_key, myfpga.cpp:596-712). */
int fheacc_keys_create(fheacc_ctx *ctx, fheacc_keys **out);
/* Releases the caller's handle; graphs and servers that use the key set keep
 * it resident until they are destroyed.  Every call that takes a context and
 * a key set fails with -1 when the keys were created under another context. */
int fheacc_keys_destroy(fheacc_keys *keys);
/* Host arrays in the reference layouts above.  Every residue is stored,
 * btkey[0] (RGSW(1)) included: the FPGA evaluation orders read it. */
int fheacc_keys_upload_bsk(fheacc_keys *keys, const uint64_t *bsk);
int fheacc_keys_upload_ksk(fheacc_keys *keys, const uint64_t *ksk);
int fheacc_keys_upload_subs(fheacc_keys *keys, const uint64_t *subs);
/* Read back in the reference layouts. */
int fheacc_keys_download_bsk(fheacc_keys *keys, uint64_t *bsk);
int fheacc_keys_download_ksk(fheacc_keys *keys, uint64_t *ksk);
int fheacc_keys_download_subs(fheacc_keys *keys, uint64_t *subs);
/* The accelerator's DDR key images, word for word what
 * FPGA::dma_write_bootstrap_key / dma_write_subs_key DMA to the card from
 * DDR_ADDR (myfpga.cpp:596-712, myfpga.h:122-125).  Bootstrap key: for
 * i < n, j < digitR, r < Br, g < 2, h < digitG the RLWE
 * btkey[r][j][i].c_text[h][g] (bsk_words in all).  Subs key: for
 * i < log2 N, l < digitKS_R the RLWE subs_key[i].RLWE_kskey[l] (subs_words).
 * Each RLWE is in the card's buffer order, 4 words of a then 4 of b, N/4
 * times (convert_RLWE_to_buffer, myfpga.h:201-223).  Export writes an image;
 * import makes the key resident from one (a key read back from a card with
 * dma_read_compare_bootstrap_key's reads, for instance). */
int fheacc_keys_export_fpga_bsk(fheacc_keys *keys, uint64_t *image);
int fheacc_keys_import_fpga_bsk(fheacc_keys *keys, const uint64_t *image);
int fheacc_keys_export_fpga_subs(fheacc_keys *keys, uint64_t *image);
int fheacc_keys_import_fpga_subs(fheacc_keys *keys, const uint64_t *image);
/* Seeded generation on the device, straight into the device layouts.
 * lwe_sk: [n] mod q; rlwe_sk: [N] mod Q, time domain (host arrays). */
int fheacc_keys_generate(fheacc_keys *keys, uint64_t seed, const uint64_t *lwe_sk, const uint64_t *rlwe_sk);
int fheacc_keys_generate_subs(fheacc_keys *keys, uint64_t seed, const uint64_t *rlwe_sk);
/* 64-bit digest of the resident keys (read-back integrity check, the role of
 * FPGA::dma_read_compare_bootstrap_key, myfpga.cpp:772-850). */
int fheacc_keys_digest(fheacc_keys *keys, uint64_t *bsk_digest, uint64_t *ksk_digest);

/* Checkpoint / restore of a resident key set: a header naming the parameter
 * set and a digest of each key, then the keys in the device layout.  No
 * counterpart in the reference, whose keys live only in process memory (the
 * accelerator's key DMA, myfpga.cpp:596-712, is the nearest).  Load checks
 * the parameter set and every digest and leaves the key set unchanged on any
 * failure. */
int fheacc_keys_save(fheacc_keys *keys, const char *path);
int fheacc_keys_load(fheacc_keys *keys, const char *path);

/* ---- host-side helpers of the reference host API ------------------------
 * LWE::LWE_secretkey, RLWE::RLWE_secretkey, LWE::LWE_encrypt,
 * LWE::LWE_decrypt (LWE.cpp:22-53,108-190; RLWE.cpp:243-279), with this
 * library's seeded sampler in place of std::random_device. */
int fheacc_lwe_secret(const fheacc_ctx *ctx, uint64_t seed, int binary, uint64_t *sk);
int fheacc_rlwe_secret(const fheacc_ctx *ctx, uint64_t seed, uint64_t *sk);
int fheacc_lwe_encrypt(const fheacc_ctx *ctx, const uint64_t *sk, const uint8_t *bits, int64_t count,
                       uint64_t seed, uint64_t first_index, uint64_t *out);
int fheacc_lwe_decrypt(const fheacc_ctx *ctx, const uint64_t *sk, const uint64_t *c, int64_t count,
                       uint8_t *bits);
/* RLWE::RLWE_encrypt of message polynomials (RLWE.cpp:324-344, with
 * RLWE_plaintext(m, false) transforming m, RLWE.cpp:71-84): sk [N] time
 * domain; msg [count][N] time domain mod Q; out [count][2][N] NTT domain.
 * Ciphertext i draws from the seeded sampler's stream first_index + i.  Runs
 * on the device; host buffers. */
int fheacc_rlwe_encrypt(fheacc_ctx *ctx, const uint64_t *sk, const uint64_t *msg, int64_t count, uint64_t seed,
                        uint64_t first_index, uint64_t *out);
/* RGSW encryption of message polynomials, RGSW_ciphertext(m, z, ...)
 * (RLWE.cpp:561-577): out [count] RGSWs in the reference layout
 * [digitG][2][2][N], row h adding NTT(m) BG^h to a (column 0) or b
 * (column 1).  An RGSW of -s feeds fheacc_homo_expand. */
int fheacc_rgsw_encrypt(fheacc_ctx *ctx, const uint64_t *sk, const uint64_t *msg, int64_t count, uint64_t seed,
                        uint64_t first_index, uint64_t *out);
/* RLWE::RLWE_decrypt(s, c, t) (RLWE.cpp:362-376): b - a s, to the time
 * domain, rounded to multiples of Q / t (denoise, RLWE.cpp:138-155);
 * out [count][N] in [0, t]. */
int fheacc_rlwe_decrypt(fheacc_ctx *ctx, const uint64_t *sk, const uint64_t *ct, int64_t count, uint64_t t,
                        uint64_t *out);

/* ---- the gate path ------------------------------------------------------
 * RLWE::eval_bootstrap (RLWE.cpp:726-786) for B independent pairs, and the
 * accelerator's batched form FPGA::fpga_eval_bootstrap_x{1,4,8,12,16}
 * (myfpga.h:141-401, myfpga.cpp:1003-1660) with any B up to 2^31 - 1 (larger
 * grids go out as several launches).  gates[i] per pair, an enum GATES value
 * in [0, 5].  Identical operands (the reference exits the process,
 * RLWE.cpp:727-730) and gates outside [0, 5] fail the call with -1: the host
 * form checks them before launching; the device form and graph replays flag
 * them on the device (those ciphertexts skip their blind rotation, and
 * fheacc_sync, which the host form calls, reports the error).  Output:
 * [B][n+1] LWE mod q. */
int fheacc_bootstrap_batch(fheacc_ctx *ctx, fheacc_keys *keys, const int32_t *gates, const uint64_t *c1,
                           const uint64_t *c2, uint64_t *out, int64_t B, const fheacc_capture *cap);
int fheacc_bootstrap_batch_dev(fheacc_ctx *ctx, fheacc_keys *keys, const int32_t *gates, const uint64_t *c1,
                               const uint64_t *c2, uint64_t *out, int64_t B, const fheacc_capture *cap);

/* The gate path of fheacc_bootstrap_batch_dev with no device allocation.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory: code under HIP graph capture, serving loops and harnesses
 * that forbid hipMalloc.  Its scratch (the key-switch digits, B * N * digitKS
 * bytes, the b words and the gather key switch's partial sums) lives in a
 * static device array of FHEACC_NOALLOC_SCRATCH_BYTES on each GPU, held
 * until the process ends; a batch that needs more fails with -1 (at the
 * benchmark's N = 1024, n = 500, digitKS = 6 that is above about 5,100
 * ciphertexts).  Every call on one GPU shares that GPU's array and leaves its
 * words behind: calls that overlap on one GPU (two streams, threads or
 * contexts) race; calls on one stream are ordered.  Per-kernel profiling is
 * skipped.  Words, checks and status reporting are those of
 * fheacc_bootstrap_batch_dev. */
#define FHEACC_NOALLOC_SCRATCH_BYTES (32u << 20)
int fheacc_bootstrap_batch_dev_noalloc(fheacc_ctx *ctx, fheacc_keys *keys, const int32_t *gates,
                                       const uint64_t *c1, const uint64_t *c2, uint64_t *out, int64_t B,
                                       const fheacc_capture *cap);

/* The gate path for fixed device buffers captured once into a HIP graph and
 * replayed: each launch re-reads gates/c1/c2 and rewrites out (the FPGA form
 * issues one MMIO instruction per external product instead,
 * myfpga.cpp:1446-1492).  The buffers must stay valid while the graph
 * exists.  Identical operands are reported by fheacc_sync. */
typedef struct fheacc_graph fheacc_graph;
int fheacc_graph_capture_bootstrap(fheacc_ctx *ctx, fheacc_keys *keys, const int32_t *gates, const uint64_t *c1,
                                   const uint64_t *c2, uint64_t *out, int64_t B, fheacc_graph **graph);
/* Mixed parameter sets in one launch: set i is a dense batch of B[i] pairs
 * under its own context and keys (each context one parameter set, all on one
 * device).  One graph holds every set's kernels, set after set, and launches
 * (on ctxs[0]'s stream) as one unit.  Each set runs at its context's tuning,
 * with its own kernel instantiation.  Each set's words are those of
 * fheacc_bootstrap_batch_dev on its own. */
int fheacc_graph_capture_mixed(int nsets, fheacc_ctx *const *ctxs, fheacc_keys *const *keys,
                               const int32_t *const *gates, const uint64_t *const *c1, const uint64_t *const *c2,
                               uint64_t *const *out, const int64_t *B, fheacc_graph **graph);
/* A mixed batch as it arrives: total ciphertexts of several parameter sets
 * interleaved in one buffer, row r (row_words words apart, at least the
 * largest n + 1) belonging to set set_of[r] (a host array read at capture).
 * Each set's kernels gather their rows (operands, gate, output) by index on
 * the device, so nothing is sorted or copied per launch; out has the same
 * rows.  Each row's words are those of fheacc_bootstrap_batch_dev on its
 * set's context and keys.  Launches like fheacc_graph_capture_mixed. */
int fheacc_graph_capture_interleaved(int nsets, fheacc_ctx *const *ctxs, fheacc_keys *const *keys,
                                     const int32_t *set_of, const int32_t *gates, const uint64_t *c1,
                                     const uint64_t *c2, uint64_t *out, int64_t total, int64_t row_words,
                                     fheacc_graph **graph);
int fheacc_graph_launch(fheacc_graph *graph);
int fheacc_graph_destroy(fheacc_graph *graph);

/* ---- kernels of the path, batched (device pointers) ----------------------
 * NTT / iNTT in place over count polys of N words (nbt.cpp:423-491). */
int fheacc_ntt_dev(fheacc_ctx *ctx, uint64_t *polys, int64_t count, int inverse);
/* RLWE_mult_RGSW (RLWE.cpp:642-704): out[i] = acc[i] (x) rgsw[i]; rgsw is in
 * the reference RGSW layout, rgsw_stride words apart (0: one shared RGSW). */
int fheacc_external_product_dev(fheacc_ctx *ctx, const uint64_t *acc, const uint64_t *rgsw,
                                int64_t rgsw_stride, uint64_t *out, int64_t count);
/* RLWE_substitute then RLWE_keyswitch with subs key[sub_factor]
 * (RLWE.cpp:444-510; the accelerator's RLWESUBS, myfpga.cpp:1662-1795). */
int fheacc_subs_keyswitch_dev(fheacc_ctx *ctx, fheacc_keys *keys, const uint64_t *in, int sub_factor,
                              uint64_t *out, int64_t count);

/* ---- RLWE expansion and leveled operations (device pointers) -------------
 * application::expand_RLWE (app.cpp:26-97; the accelerator's batched
 * expansion, myfpga.cpp:1662-2452): in [count][2][N] (NTT domain), out
 * [count][N][2][N], slot k of input c at out + (c N + k) 2N.  Needs the subs
 * key.  in and out must not overlap. */
int fheacc_expand_rlwe_dev(fheacc_ctx *ctx, fheacc_keys *keys, const uint64_t *in, uint64_t *out, int64_t count);
/* application::homo_expand (app.cpp:78-97) and the accelerator's
 * fpga_homo_expand (myfpga.h:359, myfpga.cpp:2278-2452) for count queries:
 * packed [count][digitG][2][N] (packed[h] encrypts sum_i b_i iN BG^h X^i),
 * enc_sec one RGSW of -s in the reference layout; out [count][N] RGSWs in
 * the reference layout, RGSW i row h = (u_h[i] (x) enc_sec, u_h[i]) with
 * u_h = expand_RLWE(packed[h]).  Needs the subs key. */
int fheacc_homo_expand_dev(fheacc_ctx *ctx, fheacc_keys *keys, const uint64_t *packed, const uint64_t *enc_sec,
                           uint64_t *out, int64_t count);
/* The CMUX tree of the reference's tree selection (test.cpp:1576-1582 and the
 * accelerator's RLWE_MULT_RGSW form, test.cpp:4185-4330): count trees of
 * leaves [count][2^levels][2][N]; level l (< levels) halves every tree,
 * v[j/2] = CMUX(v[j], v[j+1], C[l]), with C[l] the reference-layout RGSW at
 * rgsw + l rgsw_stride (0: RGSWs back to back); out [count][2][N].  The
 * reference then blind-rotates the result (fheacc_blind_rotate_rlwe_dev). */
int fheacc_cmux_tree_dev(fheacc_ctx *ctx, const uint64_t *leaves, int levels, const uint64_t *rgsw,
                         int64_t rgsw_stride, uint64_t *out, int64_t count);
/* RLWE::CMUX (RLWE.cpp:788-793): out[i] = (c1[i] - c0[i]) (x) rgsw[i] + c0[i];
 * rgsw in the reference RGSW layout, rgsw_stride words apart (0: one shared
 * RGSW). */
int fheacc_cmux_dev(fheacc_ctx *ctx, const uint64_t *c0, const uint64_t *c1, const uint64_t *rgsw,
                    int64_t rgsw_stride, uint64_t *out, int64_t count);
/* RLWE::blind_rotate (RLWE.cpp:822-838): acc = CMUX(acc, acc X^(+-2^i), C[i])
 * for i < steps (<= log2 N; the reference fixes 11 at N = 2048); rgsw:
 * [steps] reference-layout RGSWs shared by all count inputs. */
int fheacc_blind_rotate_rlwe_dev(fheacc_ctx *ctx, const uint64_t *in, int forward, const uint64_t *rgsw, int steps,
                                 uint64_t *out, int64_t count);
/* RLWE::blind_rotate_full (RLWE.cpp:795-820): rotation by X^b, then
 * acc = CMUX(acc, acc X^a[i], C[i]); a is a host array, rotation factors are
 * reduced mod N as poly_rotate_time does (RLWE.cpp:49-68). */
int fheacc_blind_rotate_full_dev(fheacc_ctx *ctx, const uint64_t *in, int b, const int32_t *a, const uint64_t *rgsw,
                                 int steps, uint64_t *out, int64_t count);

/* Host-buffer forms of the same calls (the reference-shaped C++ layer uses
 * them): copies in, the device call, copies out; they return when out holds
 * the result. */
int fheacc_expand_rlwe(fheacc_ctx *ctx, fheacc_keys *keys, const uint64_t *in, uint64_t *out, int64_t count);
int fheacc_subs_keyswitch(fheacc_ctx *ctx, fheacc_keys *keys, const uint64_t *in, int sub_factor, uint64_t *out,
                          int64_t count);
int fheacc_homo_expand(fheacc_ctx *ctx, fheacc_keys *keys, const uint64_t *packed, const uint64_t *enc_sec,
                       uint64_t *out, int64_t count);
int fheacc_cmux_tree(fheacc_ctx *ctx, const uint64_t *leaves, int levels, const uint64_t *rgsw, int64_t rgsw_stride,
                     uint64_t *out, int64_t count);
int fheacc_external_product(fheacc_ctx *ctx, const uint64_t *acc, const uint64_t *rgsw, int64_t rgsw_stride,
                            uint64_t *out, int64_t count);
int fheacc_cmux(fheacc_ctx *ctx, const uint64_t *c0, const uint64_t *c1, const uint64_t *rgsw, int64_t rgsw_stride,
                uint64_t *out, int64_t count);
int fheacc_blind_rotate_rlwe(fheacc_ctx *ctx, const uint64_t *in, int forward, const uint64_t *rgsw, int steps,
                             uint64_t *out, int64_t count);
int fheacc_blind_rotate_full(fheacc_ctx *ctx, const uint64_t *in, int b, const int32_t *a, const uint64_t *rgsw,
                             int steps, uint64_t *out, int64_t count);

/* ---- serving loop -------------------------------------------------------
 * No counterpart in the reference: its apps evaluate one gate
 * (RLWE::eval_bootstrap, RLWE.cpp:726-786) or one fixed batch
 * (fpga_eval_bootstrap_xN, myfpga.cpp:1003-1660) per call.  A request loop
 * over one context for callers that serve many clients:
 *  - clients register their resident keys once (fheacc_server_add_client);
 *  - a job is a gate circuit: n_in input LWE ciphertexts ([n_in][n+1] words
 *    mod q) and n_gates gates, ops[g] = {gate, src1, src2}, whose operands
 *    are earlier nodes (node i < n_in is input i, node n_in + g is gate g's
 *    output; src1 != src2).  outputs[] names the nodes copied to out
 *    ([n_out][n+1] words), which must stay valid until fheacc_server_wait;
 *  - a job is admitted, in arrival order, when its n_in + n_gates
 *    ciphertexts fit in what is left of hbm_budget;
 *  - a worker thread coalesces the ready gates of all admitted jobs, of all
 *    clients, into gate-path launches of at most max_batch; up to
 *    max_inflight launches (one stream each) run at once, so requests that
 *    arrive while others are on the GPU join the next launch.
 * Every gate is the same evaluation as fheacc_bootstrap_batch, word for word.
 * Jobs must be collected with fheacc_server_wait before
 * fheacc_server_destroy, which fails jobs that have not finished. */
typedef struct fheacc_server fheacc_server;
typedef struct fheacc_server_config {
    uint64_t hbm_budget;   /* bytes of device memory for job ciphertexts */
    int32_t max_batch;     /* most gate bootstraps per launch */
    int32_t max_inflight;  /* launches in flight at once, 1..4 */
    int32_t min_fill;      /* with a launch in flight, wait for this many ready gates */
    int32_t fair_share;    /* 1: every job with ready gates first gets max_batch / jobs of a launch,
                              the rest goes oldest job first; 0: oldest job first throughout */
} fheacc_server_config;
typedef struct fheacc_server_stats {
    int64_t jobs_submitted, jobs_done, jobs_failed, jobs_queued, jobs_active;
    int64_t bootstraps, batches, max_batch_seen;
    double wall_s;                 /* since creation or the last reset */
    double gpu_busy_s;             /* wall time with a launch in flight */
    double gate_ms_p50, gate_ms_p99;  /* gate ready -> its launch complete */
    double job_ms_p50, job_ms_p99;    /* submit -> outputs in host memory */
    uint64_t pool_bytes, peak_bytes;  /* pool size; peak admitted footprint */
} fheacc_server_stats;
int fheacc_server_create(fheacc_ctx *ctx, const fheacc_server_config *cfg, fheacc_server **server);
int fheacc_server_destroy(fheacc_server *server);
int fheacc_server_add_client(fheacc_server *server, fheacc_keys *keys, int32_t *client);
int fheacc_server_submit(fheacc_server *server, int32_t client, const uint64_t *inputs, int64_t n_in,
                         const int32_t *ops, int64_t n_gates, const int64_t *outputs, int64_t n_out,
                         uint64_t *out, int64_t *job);
/* Blocks until the job is done; returns its status (identical operands:
 * -1, RLWE.cpp:727-730).  latency_ms: submit -> outputs in host memory. */
int fheacc_server_wait(fheacc_server *server, int64_t job, double *latency_ms);
int fheacc_server_get_stats(fheacc_server *server, fheacc_server_stats *stats);
int fheacc_server_reset_stats(fheacc_server *server);

#ifdef __cplusplus
}
#endif
#endif
