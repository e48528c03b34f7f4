// api.cpp -- the C ABI (include/fheacc.h) over the HIP kernels.
#include <errno.h>
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "fheacc_internal.hpp"
#include "kernels.hpp"

using namespace fheacc;

namespace {
thread_local std::string g_err;
}

int fheacc::fail(const std::string &m, int rc) {
    g_err = m;
    // A failed runtime call (an allocation that does not fit, say) stays in
    // the thread's last-error slot until read; the launchers check launches
    // with hipGetLastError, so a later, good launch would report it.  The
    // failure is reported here, so clear it.
    (void)hipGetLastError();
    return rc;
}

namespace {

typedef unsigned __int128 u128;

uint64_t shoup64h(uint64_t w, uint64_t Q) { return (uint64_t)(((u128)w << 64) / Q); }
uint32_t shoup32h(uint32_t w, uint32_t Q) { return (uint32_t)(((uint64_t)w << 32) / Q); }
uint64_t mulmodh(uint64_t a, uint64_t b, uint64_t Q) { return (uint64_t)((u128)a * b % Q); }

ShoupGen make_sg(uint64_t Q) {
    ShoupGen s;
    s.Q = Q;
    s.a = (uint64_t)(((u128)1 << 64) / Q);
    s.b = (uint64_t)(((u128)1 << 64) % Q);
    s.bp = shoup64h(s.b, Q);
    return s;
}

// The largest inverse NTT pass whose sums may stay unreduced: 2^(R+1) Q <=
// 2^bits (passes have at most 8 stages).
static int lazy_inverse_stages(uint64_t Q, int bits) {
    const u128 top = (u128)1 << bits;
    int r = 0;
    while (r < 8 && ((u128)Q << (r + 2)) <= top) r++;
    return r;
}

template <typename W>
Arith<W> make_arith(const HostParams &p) {
    Arith<W> A;
    A.Q = (W)p.Q;
    A.Q2 = (W)(2 * p.Q);
    A.mu = (uint64_t)(((u128)1 << 64) / p.Q);
    if constexpr (sizeof(W) == 4) A.harvey = p.harvey;
    else {
        A.Q5 = 5 * p.Q;
        A.NQ = 0 - p.Q;
        A.invQ = 1.0 / (double)p.Q;
    }
    A.lzr = lazy_inverse_stages(p.Q, 8 * sizeof(W));
    return A;
}

template <typename W>
Twiddles<W> make_tw(const fheacc_ctx *ctx) {
    Twiddles<W> t;
    t.fw = (const W *)ctx->t.fw;
    t.fwp = (const W *)ctx->t.fwp;
    t.iw = (const W *)ctx->t.iw;
    t.iwp = (const W *)ctx->t.iwp;
    t.iN = (W)ctx->t.iN;
    t.iNp = (W)ctx->t.iNp;
    t.iNw = (W)ctx->t.iNw;
    t.iNwp = (W)ctx->t.iNwp;
    t.onep = sizeof(W) == 4 ? (W)shoup32h(1, (uint32_t)ctx->p.Q) : (W)shoup64h(1, ctx->p.Q);
    return t;
}

size_t wsize(const fheacc_ctx *c) { return c->p.narrow ? 4 : 8; }

// W elements of one device RGSW: narrow d*N*4, wide d*N*8
size_t rgsw_elems(const HostParams &p) { return (size_t)p.digitG * p.N * (p.narrow ? 4 : 8); }
size_t rgsw_bytes(const HostParams &p) { return rgsw_elems(p) * (p.narrow ? 4 : 8); }
int n_pad(const HostParams &p) { return (p.n + 1 + 7) / 8 * 8; }
size_t bsk_dev_bytes(const HostParams &p) { return (size_t)p.n * p.digitR * p.Br * rgsw_bytes(p); }
size_t ksk_dev_bytes(const HostParams &p) {
    return (size_t)p.N * p.digitKS * p.Bks * n_pad(p) * (p.narrow ? 4 : 8);
}
size_t subs_elems(const HostParams &p) { return (size_t)p.logN * p.digitKS_R * p.N * (p.narrow ? 2 : 4); }
size_t subs_dev_bytes(const HostParams &p) { return subs_elems(p) * (p.narrow ? 4 : 8); }

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

int ensure_scratch(fheacc_ctx *ctx, size_t bytes) {
    if (ctx->scratch_bytes >= bytes) return 0;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->scratch) HIP_TRY(hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
    size_t b = std::max(bytes, (size_t)1 << 20);
    HIP_TRY(hipMalloc(&ctx->scratch, b));
    ctx->scratch_bytes = b;
    return 0;
}

// Digit polys per NTT batch: up to `total`, as many as fit in `budget` bytes
// of LDS next to `fixed_bytes`.
int pick_pb(const HostParams &p, int total, size_t fixed_bytes, size_t budget, int requested) {
    if (requested > 0) return std::min(requested, total);
    const size_t w = p.narrow ? 4 : 8;
    int pb = total;
    while (pb > 1 && fixed_bytes + (size_t)padded(pb * p.N) * w > budget) pb--;
    return pb;
}

int check_ctx(const fheacc_ctx *c) { return c ? 0 : fail("null context"); }

// Keys index the device with their own context's strides: a key set used
// under another context would read past its allocations.
int check_keys(const fheacc_ctx *c, const fheacc_keys *k) {
    if (!k) return fail("null key set");
    return k->ctx == c ? 0 : fail("keys belong to another context");
}

// LDS of the blind-rotation kernel with pb digit polys per batch: pb >= 2d
// runs the product in place in 2d polys, pb == d in place in d polys (two
// batches), otherwise acc (2 polys) plus a pb-poly digit buffer; then the n
// evaluated a words.
size_t br_lds_bytes(const HostParams &p, int pb) {
    const size_t w = p.narrow ? 4 : 8;
    const int d = (int)p.digitG;
    const int form = product_form(pb, d);
    const size_t polys = form == kInPlace ? (size_t)padded(std::max(2, 2 * d) * p.N)
                         : form == kSplit ? (size_t)padded(std::max(2, d) * p.N)
                                          : (size_t)padded(2 * p.N) + (size_t)padded(pb * p.N);
    return polys * w + (size_t)p.n * 2 + 16;
}

// Records a (start, stop) event pair around one kernel launch when profiling.
struct Prof {
    fheacc_ctx *c;
    int cls;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    Prof(fheacc_ctx *c_, int cls_) : Prof(c_, cls_, c_->stream, c_->profile) {}
    Prof(fheacc_ctx *c_, int cls_, hipStream_t st_, bool on) : c(c_), cls(cls_), st(st_) {
        if (!on) return;
        if (hipEventCreate(&a) != hipSuccess) {
            a = nullptr;
            return;
        }
        if (hipEventCreate(&b) != hipSuccess || hipEventRecord(a, st) != hipSuccess) release();
    }
    void release() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
        a = b = nullptr;
    }
    ~Prof() {
        if (!a) return;
        if (hipEventRecord(b, st) == hipSuccess) c->events.push_back({cls, a, b});
        else release();
    }
};

void drop_events(fheacc_ctx *c) {
    for (auto &e : c->events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    c->events.clear();
}

}  // namespace

int fheacc_ctx_profile(fheacc_ctx *c, int enable) {
    if (check_ctx(c)) return -1;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    drop_events(c);
    c->profile = enable != 0;
    return 0;
}

int fheacc_ctx_kernel_ms(fheacc_ctx *c, double *ms, int64_t *n) {
    if (check_ctx(c) || !ms || !n) return fail("null argument");
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 4; i++) {
        ms[i] = 0;
        n[i] = 0;
    }
    for (auto &e : c->events) {
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, e.a, e.b));
        ms[e.cls] += t;
        n[e.cls] += 1;
    }
    drop_events(c);
    return 0;
}

const char *fheacc_last_error(void) { return g_err.c_str(); }

namespace {
// readers of the checked translation units' violation words (none in the
// default build); registered while the library loads
std::vector<unsigned (*)()> &check_readers() {
    static std::vector<unsigned (*)()> r;
    return r;
}
}  // namespace

static_assert(kCheckLds == FHEACC_DEVCHECK_LDS && kCheckCiphertext == FHEACC_DEVCHECK_CIPHERTEXT,
              "device check bits");

void fheacc::register_check_reader(unsigned (*reader)()) { check_readers().push_back(reader); }

int fheacc_device_checks_enabled(void) { return check_readers().empty() ? 0 : 1; }

int fheacc_device_check_read(int device, uint32_t *bits) {
    if (!bits) return fail("null output");
    *bits = 0;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return fail("no such device");
    int prev = 0;
    HIP_TRY(hipGetDevice(&prev));
    HIP_TRY(hipSetDevice(device));
    const hipError_t e = hipDeviceSynchronize();
    unsigned v = 0;
    for (auto r : check_readers()) v |= r();
    HIP_TRY(hipSetDevice(prev));
    if (e != hipSuccess) return fail(std::string("device synchronize: ") + hipGetErrorString(e));
    if (v == ~0u) return fail("reading the device check words failed");
    *bits = v;
    return 0;
}

int fheacc_ctx_create(const fheacc_params *params, int device, fheacc_ctx **out) {
    if (!out) return fail("null output handle");
    *out = nullptr;
    fheacc_ctx *c = new (std::nothrow) fheacc_ctx();
    if (!c) return fail("out of host memory", FHEACC_ENOMEM);
    std::string err;
    if (derive_params(params, &c->p, &err)) {
        delete c;
        return fail(err);
    }
    c->device = device;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
        delete c;
        return fail("no HIP device " + std::to_string(device));
    }
    auto cleanup = [&](int rc) {
        fheacc_ctx_destroy(c);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return cleanup(fail("hipSetDevice failed"));
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
        lds_max > 0)
        c->lds_budget = std::min<size_t>((size_t)lds_max, 160 * 1024);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0) {
        c->tune.small_max = cus;
        c->cus = cus;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(fail("stream"));
    c->own_stream = true;
    const HostParams &p = c->p;
    const int N = p.N;
    const size_t ws = p.narrow ? 4 : 8;
    std::vector<uint64_t> fw(N), fwp(N), iw(N), iwp(N);
    for (int i = 0; i < N; i++) {
        fw[i] = p.rou[i];
        iw[i] = p.irou[i];
        fwp[i] = p.narrow ? shoup32h((uint32_t)fw[i], (uint32_t)p.Q) : shoup64h(fw[i], p.Q);
        iwp[i] = p.narrow ? shoup32h((uint32_t)iw[i], (uint32_t)p.Q) : shoup64h(iw[i], p.Q);
    }
    void **dst[4] = {&c->t.fw, &c->t.fwp, &c->t.iw, &c->t.iwp};
    std::vector<uint64_t> *srcs[4] = {&fw, &fwp, &iw, &iwp};
    for (int t = 0; t < 4; t++) {
        if (hipMalloc(dst[t], N * ws) != hipSuccess) return cleanup(fail("table alloc", FHEACC_ENOMEM));
        std::vector<uint32_t> narrow;
        const void *h = srcs[t]->data();
        if (p.narrow) {
            narrow.assign(srcs[t]->begin(), srcs[t]->end());
            h = narrow.data();
        }
        if (hipMemcpy(*dst[t], h, N * ws, hipMemcpyHostToDevice) != hipSuccess) return cleanup(fail("table copy"));
    }
    c->t.iN = p.iN;
    c->t.iNp = p.narrow ? shoup32h((uint32_t)p.iN, (uint32_t)p.Q) : shoup64h(p.iN, p.Q);
    c->t.iNw = (uint64_t)((u128)p.iN * p.irou[1] % p.Q);
    c->t.iNwp = p.narrow ? shoup32h((uint32_t)c->t.iNw, (uint32_t)p.Q) : shoup64h(c->t.iNw, p.Q);
    c->t.cdt_len = (int)p.cdt.size();
    if (hipMalloc(&c->t.cdt, p.cdt.size() * 8) != hipSuccess) return cleanup(fail("cdt alloc", FHEACC_ENOMEM));
    if (hipMemcpy(c->t.cdt, p.cdt.data(), p.cdt.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(fail("cdt copy"));
    if (hipMalloc(&c->status, sizeof(int)) != hipSuccess) return cleanup(fail("status alloc", FHEACC_ENOMEM));
    if (hipMemset(c->status, 0, sizeof(int)) != hipSuccess) return cleanup(fail("status clear"));
    *out = c;
    return 0;
}

static void ctx_free(fheacc_ctx *c) {
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : {c->t.fw, c->t.fwp, c->t.iw, c->t.iwp, (void *)c->t.cdt, c->scratch, (void *)c->status})
        if (p) (void)hipFree(p);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void fheacc::ctx_retain(fheacc_ctx *c) { c->refs.fetch_add(1); }

void fheacc::ctx_release(fheacc_ctx *c) {
    if (c && c->refs.fetch_sub(1) == 1) ctx_free(c);
}

int fheacc_ctx_destroy(fheacc_ctx *c) {
    ctx_release(c);
    return 0;
}

int fheacc_ctx_info(const fheacc_ctx *c, fheacc_info *o) {
    if (check_ctx(c) || !o) return fail("null argument");
    const HostParams &p = c->p;
    o->Q = p.Q;
    o->iN = p.iN;
    o->ROU = p.ROU;
    o->digitG = p.digitG;
    o->digitR = p.digitR;
    o->digitKS = p.digitKS;
    o->digitKS_R = p.digitKS_R;
    o->narrow = p.narrow;
    o->harvey = p.harvey;
    o->lazy_inverse_stages = (uint32_t)lazy_inverse_stages(p.Q, p.narrow ? 32 : 64);
    o->reserved = 0;
    for (int g = 0; g < 6; g++) o->gate_constant[g] = p.gate_constant[g];
    o->bsk_words = (uint64_t)p.Br * p.digitR * p.n * p.digitG * 4 * p.N;
    o->ksk_words = (uint64_t)p.Bks * p.digitKS * p.N * (p.n + 1);
    o->subs_words = (uint64_t)p.logN * p.digitKS_R * 2 * p.N;
    o->bsk_device_bytes = bsk_dev_bytes(p);
    o->ksk_device_bytes = ksk_dev_bytes(p);
    o->subs_device_bytes = subs_dev_bytes(p);
    return 0;
}

int fheacc_ctx_set_stream(fheacc_ctx *c, void *s) {
    if (check_ctx(c)) return -1;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (c->own_stream) HIP_TRY(hipStreamDestroy(c->stream));
    c->stream = (hipStream_t)s;  // NULL: the device's null stream
    c->own_stream = false;
    return 0;
}

int fheacc_sync(fheacc_ctx *c) {
    if (check_ctx(c)) return -1;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    int st = 0;
    HIP_TRY(hipMemcpy(&st, c->status, sizeof(int), hipMemcpyDeviceToHost));
    if (st) {
        HIP_TRY(hipMemset(c->status, 0, sizeof(int)));
        if (st & FHEACC_STATUS_GATE) return fail("gate out of range: gates are enum GATES values 0..5 (params.h:107)");
        return fail("identical operands passed to the gate bootstrap (RLWE.cpp:727-730)");
    }
    return 0;
}

// ---------------------------------------------------------------- keys

int fheacc_keys_create(fheacc_ctx *c, fheacc_keys **out) {
    if (check_ctx(c) || !out) return fail("null argument");
    fheacc_keys *k = new (std::nothrow) fheacc_keys();
    if (!k) return fail("out of host memory", FHEACC_ENOMEM);
    k->ctx = c;
    ctx_retain(c);
    *out = k;
    return 0;
}

void fheacc::keys_retain(fheacc_keys *k) { k->refs.fetch_add(1); }

void fheacc::keys_release(fheacc_keys *k) {
    if (!k || k->refs.fetch_sub(1) != 1) return;
    fheacc_ctx *c = k->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (void *p : {k->bsk, k->ksk, k->subs})
        if (p) (void)hipFree(p);
    delete k;
    ctx_release(c);
}

int fheacc_keys_destroy(fheacc_keys *k) {
    keys_release(k);
    return 0;
}

static int alloc_key(void **p, size_t bytes) {
    if (*p) return 0;
    HIP_TRY(hipMalloc(p, bytes));
    return 0;
}

// map[device RGSW e] = reference RGSW index: device [n][digitR][Br],
// reference btkey[Br][digitR][n] (RLWE.cpp:579-606)
static std::vector<int64_t> bsk_map(const HostParams &p) {
    std::vector<int64_t> m((size_t)p.n * p.digitR * p.Br);
    size_t e = 0;
    for (int k = 0; k < p.n; k++)
        for (uint32_t j = 0; j < p.digitR; j++)
            for (uint64_t r = 0; r < p.Br; r++) m[e++] = ((int64_t)r * p.digitR + j) * p.n + k;
    return m;
}

// Bootstrap-key words in (host -> device layout) and out.  ddr = false: the
// reference's btkey[Br][digitR][n] order; ddr = true: the accelerator's DDR
// image, whose RGSW order [n][digitR][Br] is the device order.
static int bsk_in(fheacc_keys *k, const uint64_t *words, bool ddr) {
    fheacc_ctx *c = k->ctx;
    const HostParams &p = c->p;
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = alloc_key(&k->bsk, bsk_dev_bytes(p))) return rc;
    k->bsk_bytes = bsk_dev_bytes(p);
    const size_t rw = (size_t)p.digitG * 4 * p.N;
    const int64_t count = (int64_t)p.Br * p.digitR * p.n;
    DevBuf stage, dmap;
    HIP_TRY(hipMalloc(&stage.p, count * rw * 8));
    HIP_TRY(hipMemcpy(stage.p, words, count * rw * 8, hipMemcpyHostToDevice));
    const int64_t *map = nullptr;
    if (!ddr) {
        std::vector<int64_t> m = bsk_map(p);
        HIP_TRY(hipMalloc(&dmap.p, m.size() * 8));
        HIP_TRY(hipMemcpy(dmap.p, m.data(), m.size() * 8, hipMemcpyHostToDevice));
        map = (const int64_t *)dmap.p;
    }
    hipError_t e = p.narrow
        ? launch_rgsw_to_device<uint32_t>((const uint64_t *)stage.p, rw, (uint32_t *)k->bsk, rgsw_elems(p),
                                          p.digitG, p.N, make_sg(p.Q), map, count, c->stream, ddr)
        : launch_rgsw_to_device<uint64_t>((const uint64_t *)stage.p, rw, (uint64_t *)k->bsk, rgsw_elems(p),
                                          p.digitG, p.N, make_sg(p.Q), map, count, c->stream, ddr);
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

static int bsk_out(fheacc_keys *k, uint64_t *words, bool ddr) {
    fheacc_ctx *c = k->ctx;
    const HostParams &p = c->p;
    HIP_TRY(hipSetDevice(c->device));
    const size_t rw = (size_t)p.digitG * 4 * p.N;
    const int64_t count = (int64_t)p.Br * p.digitR * p.n;
    DevBuf stage, dmap;
    HIP_TRY(hipMalloc(&stage.p, count * rw * 8));
    const int64_t *map = nullptr;
    if (!ddr) {
        std::vector<int64_t> m = bsk_map(p);
        HIP_TRY(hipMalloc(&dmap.p, m.size() * 8));
        HIP_TRY(hipMemcpy(dmap.p, m.data(), m.size() * 8, hipMemcpyHostToDevice));
        map = (const int64_t *)dmap.p;
    }
    hipError_t e = p.narrow
        ? launch_rgsw_to_ref<uint32_t>((const uint32_t *)k->bsk, rgsw_elems(p), (uint64_t *)stage.p, rw, p.digitG,
                                       p.N, map, count, c->stream, ddr)
        : launch_rgsw_to_ref<uint64_t>((const uint64_t *)k->bsk, rgsw_elems(p), (uint64_t *)stage.p, rw, p.digitG,
                                       p.N, map, count, c->stream, ddr);
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(words, stage.p, count * rw * 8, hipMemcpyDeviceToHost));
    return 0;
}

int fheacc_keys_upload_bsk(fheacc_keys *k, const uint64_t *bsk) {
    if (!k || !bsk) return fail("null argument");
    return bsk_in(k, bsk, false);
}

int fheacc_keys_download_bsk(fheacc_keys *k, uint64_t *bsk) {
    if (!k || !bsk || !k->bsk) return fail("no bootstrap key resident");
    return bsk_out(k, bsk, false);
}

int fheacc_keys_import_fpga_bsk(fheacc_keys *k, const uint64_t *image) {
    if (!k || !image) return fail("null argument");
    return bsk_in(k, image, true);
}

int fheacc_keys_export_fpga_bsk(fheacc_keys *k, uint64_t *image) {
    if (!k || !image || !k->bsk) return fail("no bootstrap key resident");
    return bsk_out(k, image, true);
}

int fheacc_keys_upload_ksk(fheacc_keys *k, const uint64_t *ksk) {
    if (!k || !ksk) return fail("null argument");
    fheacc_ctx *c = k->ctx;
    const HostParams &p = c->p;
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = alloc_key(&k->ksk, ksk_dev_bytes(p))) return rc;
    k->ksk_bytes = ksk_dev_bytes(p);
    const size_t words = (size_t)p.Bks * p.digitKS * p.N * (p.n + 1);
    DevBuf stage;
    HIP_TRY(hipMalloc(&stage.p, words * 8));
    HIP_TRY(hipMemcpy(stage.p, ksk, words * 8, hipMemcpyHostToDevice));
    hipError_t e = p.narrow
        ? launch_ksk_convert<uint32_t>((const uint64_t *)stage.p, (uint32_t *)k->ksk, nullptr, (int)p.Bks, p.digitKS,
                                       p.N, p.n + 1, n_pad(p), c->stream)
        : launch_ksk_convert<uint64_t>((const uint64_t *)stage.p, (uint64_t *)k->ksk, nullptr, (int)p.Bks, p.digitKS,
                                       p.N, p.n + 1, n_pad(p), c->stream);
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int fheacc_keys_download_ksk(fheacc_keys *k, uint64_t *ksk) {
    if (!k || !ksk || !k->ksk) return fail("no key-switch key resident");
    fheacc_ctx *c = k->ctx;
    const HostParams &p = c->p;
    HIP_TRY(hipSetDevice(c->device));
    const size_t words = (size_t)p.Bks * p.digitKS * p.N * (p.n + 1);
    DevBuf stage;
    HIP_TRY(hipMalloc(&stage.p, words * 8));
    hipError_t e = p.narrow
        ? launch_ksk_convert<uint32_t>(nullptr, (uint32_t *)k->ksk, (uint64_t *)stage.p, (int)p.Bks, p.digitKS, p.N,
                                       p.n + 1, n_pad(p), c->stream)
        : launch_ksk_convert<uint64_t>(nullptr, (uint64_t *)k->ksk, (uint64_t *)stage.p, (int)p.Bks, p.digitKS, p.N,
                                       p.n + 1, n_pad(p), c->stream);
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(ksk, stage.p, words * 8, hipMemcpyDeviceToHost));
    return 0;
}

int fheacc_keys_upload_subs(fheacc_keys *k, const uint64_t *subs) {
    if (!k || !subs) return fail("null argument");
    fheacc_ctx *c = k->ctx;
    const HostParams &p = c->p;
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = alloc_key(&k->subs, subs_dev_bytes(p))) return rc;
    k->subs_bytes = subs_dev_bytes(p);
    const int64_t total = (int64_t)p.logN * p.digitKS_R * p.N;
    DevBuf stage;
    HIP_TRY(hipMalloc(&stage.p, total * 2 * 8));
    HIP_TRY(hipMemcpy(stage.p, subs, total * 2 * 8, hipMemcpyHostToDevice));
    hipError_t e = p.narrow
        ? launch_subs_convert<uint32_t>(stage.p, k->subs, p.N, total, false, make_sg(p.Q), c->stream)
        : launch_subs_convert<uint64_t>(stage.p, k->subs, p.N, total, false, make_sg(p.Q), c->stream);
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

static int subs_download(fheacc_keys *k, uint64_t *subs) {
    fheacc_ctx *c = k->ctx;
    const HostParams &p = c->p;
    const int64_t total = (int64_t)p.logN * p.digitKS_R * p.N;
    DevBuf stage;
    HIP_TRY(hipMalloc(&stage.p, total * 2 * 8));
    hipError_t e = p.narrow
        ? launch_subs_convert<uint32_t>(k->subs, stage.p, p.N, total, true, make_sg(p.Q), c->stream)
        : launch_subs_convert<uint64_t>(k->subs, stage.p, p.N, total, true, make_sg(p.Q), c->stream);
    HIP_TRY(e);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(subs, stage.p, total * 2 * 8, hipMemcpyDeviceToHost));
    return 0;
}

int fheacc_keys_download_subs(fheacc_keys *k, uint64_t *subs) {
    if (!k || !subs || !k->subs) return fail("no subs key resident");
    HIP_TRY(hipSetDevice(k->ctx->device));
    return subs_download(k, subs);
}

// One RLWE (a then b, N words each) <-> the accelerator's buffer order: 4 words
// of a, then 4 of b, N/4 times (convert_RLWE_to_buffer / convert_buffer_to_RLWE,
// myfpga.h:201-243).
static void rlwe_ddr_order(const uint64_t *ab, uint64_t *img, int N) {
    for (int k = 0; k < N; k++) {
        img[(k >> 2) * 8 + (k & 3)] = ab[k];
        img[(k >> 2) * 8 + 4 + (k & 3)] = ab[N + k];
    }
}

static void rlwe_from_ddr_order(const uint64_t *img, uint64_t *ab, int N) {
    for (int k = 0; k < N; k++) {
        ab[k] = img[(k >> 2) * 8 + (k & 3)];
        ab[N + k] = img[(k >> 2) * 8 + 4 + (k & 3)];
    }
}

// The subs key's DDR image (FPGA::dma_write_subs_key, myfpga.cpp:650-712):
// subs_key[i].RLWE_kskey[l] for i < log2 N, l < digitKS_R, in that order.
int fheacc_keys_export_fpga_subs(fheacc_keys *k, uint64_t *image) {
    if (!k || !image || !k->subs) return fail("no subs key resident");
    HIP_TRY(hipSetDevice(k->ctx->device));
    const HostParams &p = k->ctx->p;
    const size_t cnt = (size_t)p.logN * p.digitKS_R, rl = 2 * (size_t)p.N;
    std::vector<uint64_t> ref;
    try {
        ref.resize(cnt * rl);
    } catch (const std::bad_alloc &) {
        return fail("out of host memory", FHEACC_ENOMEM);
    }
    if (int rc = subs_download(k, ref.data())) return rc;
    for (size_t r = 0; r < cnt; r++) rlwe_ddr_order(ref.data() + r * rl, image + r * rl, p.N);
    return 0;
}

int fheacc_keys_import_fpga_subs(fheacc_keys *k, const uint64_t *image) {
    if (!k || !image) return fail("null argument");
    const HostParams &p = k->ctx->p;
    const size_t cnt = (size_t)p.logN * p.digitKS_R, rl = 2 * (size_t)p.N;
    std::vector<uint64_t> ref;
    try {
        ref.resize(cnt * rl);
    } catch (const std::bad_alloc &) {
        return fail("out of host memory", FHEACC_ENOMEM);
    }
    for (size_t r = 0; r < cnt; r++) rlwe_from_ddr_order(image + r * rl, ref.data() + r * rl, p.N);
    return fheacc_keys_upload_subs(k, ref.data());
}

template <typename W>
static int keygen_impl(fheacc_keys *k, uint64_t seed, const uint64_t *lwe_sk, const uint64_t *rlwe_sk, bool subs) {
    fheacc_ctx *c = k->ctx;
    const HostParams &p = c->p;
    const uint64_t Q = p.Q;
    const size_t ws = sizeof(W);
    DevBuf dz, dzn, dznp, dsk, dpow, dpowp, ds, dsp, dcij, dcijp;
    HIP_TRY(hipMalloc(&dz.p, p.N * 8));
    HIP_TRY(hipMemcpy(dz.p, rlwe_sk, p.N * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&dzn.p, p.N * ws));
    HIP_TRY(hipMalloc(&dznp.p, p.N * ws));
    KeygenArgs<W> a{};
    a.A = make_arith<W>(p);
    a.tw = make_tw<W>(c);
    a.seed = seed;
    a.cdt = c->t.cdt;
    a.cdt_len = c->t.cdt_len;
    a.E_rlwe = p.E_rlwe;
    a.zn = (W *)dzn.p;
    a.znp = (W *)dznp.p;
    a.z = (const uint64_t *)dz.p;
    a.sg = make_sg(Q);
    a.q = (uint32_t)p.q;
    a.Br = (uint32_t)p.Br;
    a.digitR = p.digitR;
    KskGenArgs<W> ka{};
    const int d = subs ? p.digitKS_R : p.digitG;
    std::vector<uint64_t> pw(d), pwp(d);
    uint64_t base = subs ? (1ULL << p.Bks_R_bits) % Q : (1ULL << p.BGbits) % Q, cur = 1 % Q;
    for (int h = 0; h < d; h++) {
        pw[h] = cur;
        pwp[h] = p.narrow ? shoup32h((uint32_t)cur, (uint32_t)Q) : shoup64h(cur, Q);
        cur = mulmodh(cur, base, Q);
    }
    HIP_TRY(hipMalloc(&dpow.p, d * 8));
    HIP_TRY(hipMalloc(&dpowp.p, d * 8));
    HIP_TRY(hipMemcpy(dpow.p, pw.data(), d * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dpowp.p, pwp.data(), d * 8, hipMemcpyHostToDevice));
    a.bg_pow = (const uint64_t *)dpow.p;
    a.bg_powp = (const uint64_t *)dpowp.p;
    a.d = d;
    if (subs) {
        if (int rc = alloc_key(&k->subs, subs_dev_bytes(p))) return rc;
        k->subs_bytes = subs_dev_bytes(p);
        a.bsk = (W *)k->subs;
        a.rgsw_elems = (size_t)p.N * (p.narrow ? 2 : 4);
        a.n = p.logN;
    } else {
        if (int rc = alloc_key(&k->bsk, bsk_dev_bytes(p))) return rc;
        if (int rc = alloc_key(&k->ksk, ksk_dev_bytes(p))) return rc;
        k->bsk_bytes = bsk_dev_bytes(p);
        k->ksk_bytes = ksk_dev_bytes(p);
        HIP_TRY(hipMalloc(&dsk.p, p.n * 8));
        HIP_TRY(hipMemcpy(dsk.p, lwe_sk, p.n * 8, hipMemcpyHostToDevice));
        a.lwe_sk = (const uint64_t *)dsk.p;
        a.bsk = (W *)k->bsk;
        a.rgsw_elems = rgsw_elems(p);
        a.n = p.n;
        // LWE key-switch key inputs: s mapped into Z_Q (LWE.cpp:62-75)
        std::vector<W> s(p.n), sp(p.n);
        for (int h = 0; h < p.n; h++) {
            uint64_t v = lwe_sk[h] >= p.q / 2 ? lwe_sk[h] + (Q - p.q) : lwe_sk[h];
            s[h] = (W)v;
            sp[h] = (W)(p.narrow ? shoup32h((uint32_t)v, (uint32_t)Q) : shoup64h(v, Q));
        }
        std::vector<uint64_t> cij((size_t)p.Bks * p.digitKS), cijp(cij.size());
        for (uint64_t i = 0; i < p.Bks; i++) {
            uint64_t pk = 1;
            for (uint32_t j = 0; j < p.digitKS; j++) {
                uint64_t v = mulmodh(i % Q, pk, Q);
                cij[i * p.digitKS + j] = v;
                cijp[i * p.digitKS + j] = p.narrow ? shoup32h((uint32_t)v, (uint32_t)Q) : shoup64h(v, Q);
                pk = mulmodh(pk, p.Bks, Q);
            }
        }
        HIP_TRY(hipMalloc(&ds.p, p.n * ws));
        HIP_TRY(hipMalloc(&dsp.p, p.n * ws));
        HIP_TRY(hipMalloc(&dcij.p, cij.size() * 8));
        HIP_TRY(hipMalloc(&dcijp.p, cij.size() * 8));
        HIP_TRY(hipMemcpy(ds.p, s.data(), p.n * ws, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dsp.p, sp.data(), p.n * ws, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dcij.p, cij.data(), cij.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(dcijp.p, cijp.data(), cij.size() * 8, hipMemcpyHostToDevice));
        ka.A = a.A;
        ka.seed = seed;
        ka.cdt = a.cdt;
        ka.cdt_len = a.cdt_len;
        ka.E_rlwe = p.E_rlwe;
        ka.ksk = (W *)k->ksk;
        ka.rows = (int64_t)p.N * p.digitKS * p.Bks;
        ka.n = p.n;
        ka.n_pad = n_pad(p);
        ka.N = p.N;
        ka.dks = p.digitKS;
        ka.Bks = (int)p.Bks;
        ka.s = (const W *)ds.p;
        ka.sp = (const W *)dsp.p;
        ka.z = (const uint64_t *)dz.p;
        ka.cij = (const uint64_t *)dcij.p;
        ka.cijp = (const uint64_t *)dcijp.p;
    }
    HIP_TRY(launch_keygen<W>(a, ka, p.logN, (const uint64_t *)dz.p, c->stream, subs));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int fheacc_keys_generate(fheacc_keys *k, uint64_t seed, const uint64_t *lwe_sk, const uint64_t *rlwe_sk) {
    if (!k || !lwe_sk || !rlwe_sk) return fail("null argument");
    HIP_TRY(hipSetDevice(k->ctx->device));
    return k->ctx->p.narrow ? keygen_impl<uint32_t>(k, seed, lwe_sk, rlwe_sk, false)
                            : keygen_impl<uint64_t>(k, seed, lwe_sk, rlwe_sk, false);
}

int fheacc_keys_generate_subs(fheacc_keys *k, uint64_t seed, const uint64_t *rlwe_sk) {
    if (!k || !rlwe_sk) return fail("null argument");
    HIP_TRY(hipSetDevice(k->ctx->device));
    return k->ctx->p.narrow ? keygen_impl<uint32_t>(k, seed, nullptr, rlwe_sk, true)
                            : keygen_impl<uint64_t>(k, seed, nullptr, rlwe_sk, true);
}

static int digest_of(fheacc_ctx *c, const void *buf, size_t bytes, uint64_t *out) {
    if (!buf) {
        *out = 0;
        return 0;
    }
    const int blocks = 1024;
    DevBuf part;
    HIP_TRY(hipMalloc(&part.p, blocks * 8));
    HIP_TRY(launch_digest((const uint64_t *)buf, (int64_t)(bytes / 8), (uint64_t *)part.p, blocks, c->stream));
    std::vector<uint64_t> h(blocks);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(h.data(), part.p, blocks * 8, hipMemcpyDeviceToHost));
    uint64_t s = 0;
    for (uint64_t v : h) s += v;
    *out = s;
    return 0;
}

int fheacc_keys_digest(fheacc_keys *k, uint64_t *bd, uint64_t *kd) {
    if (!k) return fail("null argument");
    HIP_TRY(hipSetDevice(k->ctx->device));
    uint64_t b = 0, s = 0;
    if (int rc = digest_of(k->ctx, k->bsk, k->bsk_bytes, &b)) return rc;
    if (int rc = digest_of(k->ctx, k->ksk, k->ksk_bytes, &s)) return rc;
    if (bd) *bd = b;
    if (kd) *kd = s;
    return 0;
}

// ------------------------------------------------------- key checkpoints
//
// File: 17 little-endian uint64 header fields, then the resident keys in the
// device layout, bsk, ksk, subs (absent ones have 0 bytes):
//   magic "FHEACCK1", version, device word bytes, n, N, q, Q, BGbits, Br,
//   Bks, Bks_R_bits, bsk_bytes, ksk_bytes, subs_bytes, and a digest of each.
// Keys stream through a pinned 64 MiB buffer; a multi-GB key never needs a
// host copy of its own.

namespace {
constexpr uint64_t kKeyMagic = 0x314B434341454846ULL;  // "FHEACCK1"
constexpr uint64_t kKeyVersion = 2;  // 2: the bootstrap key holds every residue, 0 included
constexpr size_t kKeyChunk = 64 << 20;

struct PinnedBuf {
    void *p = nullptr;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
};

struct File {
    FILE *f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};
}  // namespace

int fheacc_keys_save(fheacc_keys *k, const char *path) {
    if (!k || !path) return fail("null argument");
    fheacc_ctx *c = k->ctx;
    const HostParams &p = c->p;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    uint64_t dg[3] = {0, 0, 0};
    const void *sec[3] = {k->bsk, k->ksk, k->subs};
    const size_t len[3] = {k->bsk ? k->bsk_bytes : 0, k->ksk ? k->ksk_bytes : 0, k->subs ? k->subs_bytes : 0};
    for (int i = 0; i < 3; i++)
        if (int rc = digest_of(c, sec[i], len[i], &dg[i])) return rc;
    const uint64_t hdr[17] = {kKeyMagic, kKeyVersion, (uint64_t)wsize(c), (uint64_t)p.n, (uint64_t)p.N, p.q, p.Q,
                              p.BGbits, p.Br, p.Bks, p.Bks_R_bits, len[0], len[1], len[2], dg[0], dg[1], dg[2]};
    File f;
    f.f = fopen(path, "wb");
    if (!f.f) return fail(std::string("cannot open ") + path + " for writing");
    if (fwrite(hdr, 8, 17, f.f) != 17) return fail("write failed");
    PinnedBuf pb;
    HIP_TRY(hipHostMalloc(&pb.p, kKeyChunk, hipHostMallocDefault));
    for (int i = 0; i < 3; i++) {
        for (size_t off = 0; off < len[i]; off += kKeyChunk) {
            const size_t m = std::min(kKeyChunk, len[i] - off);
            HIP_TRY(hipMemcpyAsync(pb.p, (const char *)sec[i] + off, m, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            if (fwrite(pb.p, 1, m, f.f) != m) return fail("write failed");
        }
    }
    if (fclose(f.f) != 0) {
        f.f = nullptr;
        return fail("write failed");
    }
    f.f = nullptr;
    return 0;
}

int fheacc_keys_load(fheacc_keys *k, const char *path) {
    if (!k || !path) return fail("null argument");
    fheacc_ctx *c = k->ctx;
    const HostParams &p = c->p;
    File f;
    f.f = fopen(path, "rb");
    if (!f.f) return fail(std::string("cannot open ") + path);
    uint64_t hdr[17];
    if (fread(hdr, 8, 17, f.f) != 17 || hdr[0] != kKeyMagic) return fail("not a key file");
    if (hdr[1] != kKeyVersion) return fail("unsupported key file version " + std::to_string(hdr[1]));
    if (hdr[2] != wsize(c) || hdr[3] != (uint64_t)p.n || hdr[4] != (uint64_t)p.N || hdr[5] != p.q || hdr[6] != p.Q ||
        hdr[7] != p.BGbits || hdr[8] != p.Br || hdr[9] != p.Bks || hdr[10] != p.Bks_R_bits)
        return fail("key file belongs to another parameter set");
    const size_t want[3] = {bsk_dev_bytes(p), ksk_dev_bytes(p), subs_dev_bytes(p)};
    for (int i = 0; i < 3; i++)
        if (hdr[11 + i] != 0 && hdr[11 + i] != want[i]) return fail("key file section size mismatch");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // load into fresh buffers; the key set changes only when all of it checks out
    void *buf[3] = {nullptr, nullptr, nullptr};
    auto release = [&](int rc) {
        for (void *b : buf)
            if (b) (void)hipFree(b);
        return rc;
    };
    PinnedBuf pb;
    HIP_TRY(hipHostMalloc(&pb.p, kKeyChunk, hipHostMallocDefault));
    for (int i = 0; i < 3; i++) {
        const size_t n = hdr[11 + i];
        if (!n) continue;
        hipError_t e = hipMalloc(&buf[i], n);
        if (e != hipSuccess) return release(fail("key allocation", FHEACC_ENOMEM));
        for (size_t off = 0; off < n; off += kKeyChunk) {
            const size_t m = std::min(kKeyChunk, n - off);
            if (fread(pb.p, 1, m, f.f) != m) return release(fail("key file truncated"));
            e = hipMemcpyAsync((char *)buf[i] + off, pb.p, m, hipMemcpyHostToDevice, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) return release(fail(std::string("key upload: ") + hipGetErrorString(e)));
        }
        uint64_t d = 0;
        if (int rc = digest_of(c, buf[i], n, &d)) return release(rc);
        if (d != hdr[14 + i]) return release(fail("key file digest mismatch (corrupted file)"));
    }
    void **dst[3] = {&k->bsk, &k->ksk, &k->subs};
    size_t *dlen[3] = {&k->bsk_bytes, &k->ksk_bytes, &k->subs_bytes};
    for (int i = 0; i < 3; i++) {
        if (!buf[i]) continue;
        if (*dst[i]) (void)hipFree(*dst[i]);
        *dst[i] = buf[i];
        *dlen[i] = hdr[11 + i];
        buf[i] = nullptr;
    }
    k->bsk_digest = hdr[14];
    k->ksk_digest = hdr[15];
    return 0;
}

// ---------------------------------------------------------------- host API

int fheacc_lwe_secret(const fheacc_ctx *c, uint64_t seed, int binary, uint64_t *sk) {
    if (check_ctx(c) || !sk) return fail("null argument");
    const HostParams &p = c->p;
    for (int i = 0; i < p.n; i++) {
        uint64_t w = rng_word(seed, stream_id(TAG_LWE_SK, 0), (uint64_t)i);
        sk[i] = binary ? (w >> 63) : ((uint64_t)(((u128)w * 3) >> 64) + p.q - 1) % p.q;
    }
    return 0;
}

int fheacc_rlwe_secret(const fheacc_ctx *c, uint64_t seed, uint64_t *sk) {
    if (check_ctx(c) || !sk) return fail("null argument");
    const HostParams &p = c->p;
    for (int i = 0; i < p.N; i++) {
        uint64_t w = rng_word(seed, stream_id(TAG_RLWE_SK, 0), (uint64_t)i);
        sk[i] = ((uint64_t)(((u128)w * 3) >> 64) + p.Q - 1) % p.Q;
    }
    return 0;
}

// LWE.cpp:108-146 (t = 4): b = e + m q/4 + <a, s> mod q
int fheacc_lwe_encrypt(const fheacc_ctx *c, const uint64_t *sk, const uint8_t *bits, int64_t count, uint64_t seed,
                       uint64_t first_index, uint64_t *out) {
    if (check_ctx(c) || !sk || !bits || !out) return fail("null argument");
    const HostParams &p = c->p;
    const uint64_t q = p.q;
    for (int64_t i = 0; i < count; i++) {
        const uint64_t stream = stream_id(TAG_LWE_ENC, first_index + i);
        uint64_t *o = out + i * (p.n + 1);
        int64_t e = gauss_sample(p, rng_word(seed, stream, p.n)) % (int64_t)p.E;
        uint64_t b = (uint64_t)(e + (int64_t)((uint64_t)bits[i] * (q / 4))) % q;
        for (int h = 0; h < p.n; h++) {
            o[h] = (uint64_t)(((u128)rng_word(seed, stream, h) * q) >> 64);
            b = (b + (o[h] * sk[h]) % q) % q;
        }
        o[p.n] = b;
    }
    return 0;
}

// LWE.cpp:148-190
int fheacc_lwe_decrypt(const fheacc_ctx *c, const uint64_t *sk, const uint64_t *ct, int64_t count, uint8_t *bits) {
    if (check_ctx(c) || !sk || !ct || !bits) return fail("null argument");
    const HostParams &p = c->p;
    for (int64_t i = 0; i < count; i++) {
        const uint64_t *x = ct + i * (p.n + 1);
        uint64_t m = x[p.n];
        for (int h = 0; h < p.n; h++) m = (m - x[h] * sk[h]) % p.q;
        m = (m + p.E * 2) % p.q;
        bits[i] = (uint8_t)(m * 4 / p.q);
    }
    return 0;
}

// ---------------------------------------------------------------- gate path

// Step ranges of the gather key switch for a batch of B: about 1024
// workgroups in all (4 per CU), at least 64 steps per range, at most 256
// ranges.  Large batches fill the chip with one range per ciphertext.
static int ks_gather_split(const HostParams &p, int64_t B) {
    const int64_t cb = (p.n + 1 + 255) / 256;
    const int64_t S = (int64_t)p.N * p.digitKS;
    int64_t split = 1024 / std::max<int64_t>(1, cb * B);
    split = std::min<int64_t>(split, std::min<int64_t>(256, S / 64));
    return (int)std::max<int64_t>(split, 1);
}

// Ciphertexts per staged key-switch workgroup: each workgroup streams its
// column block of the whole key once, so more ciphertexts per workgroup means
// fewer passes over the key.  On 64-bit words (keys of several GB: 3.6 GB at
// N = 2048, n = 630) the largest of 1024, 512 that still gives every CU a
// workgroup (the grid is column blocks x ciphertext groups): 16.2 -> 14.7 ms
// for 2458 ciphertexts at 512.  32-bit keys stay at 256, where 512 measured
// slower (1.37 -> 1.52 ms at the benchmark batch).
static int ks_staged_threads(const HostParams &p, int64_t B, int cus) {
    if (p.narrow) return 256;
    const int64_t col_blocks = (p.n + 1 + 7) / 8;
    for (int nt : {1024, 512})
        if (col_blocks * ((B + nt - 1) / nt) >= std::max(cus, 1)) return nt;
    return 256;
}

// partial-sum rows the gather key switch may need for any launch of at most
// B ciphertexts (a serving-loop segment can be smaller than its launch)
static int64_t ks_part_rows(const HostParams &p, int64_t B) {
    int64_t rows = 0;
    for (int64_t b = 1; b <= std::min<int64_t>(B, 1024); b++) {
        const int split = ks_gather_split(p, b);
        if (split > 1) rows = std::max(rows, b * split);
    }
    return rows;
}

// Progress words of a paced tiled blind rotation of B ciphertexts (none unless
// the context's schedule paces): one per ciphertext.
static size_t pacing_words(const fheacc_ctx *c, int64_t B) {
    const Tuning &t = c->tune;
    if (t.br_persist <= 0 || t.br_sync_window <= 0 || B <= 0) return 0;
    return (size_t)B;
}

// Scratch of one gate-path launch: the key-switch digits u8[B][N][digitKS],
// the key-switch b words u64[B], the gather key switch's partial sums, and
// the lock-step counters of a paced launch.
size_t fheacc::gate_scratch_bytes(const fheacc_ctx *c, int64_t B) {
    const HostParams &p = c->p;
    const size_t ksd_bytes = (size_t)B * p.N * p.digitKS;
    const size_t ksb_end = (ksd_bytes + 255) / 256 * 256 + (size_t)B * 8;
    const size_t part_end = (ksb_end + 255) / 256 * 256 + (size_t)ks_part_rows(p, B) * (p.n + 1) * 8;
    return (part_end + 255) / 256 * 256 + pacing_words(c, B) * 4;
}

template <typename W>
static int gate_path_impl(fheacc_ctx *c, const fheacc_keys *k, const GateLaunch &g) {
    const HostParams &p = c->p;
    const int64_t B = g.B;
    const size_t ksd_bytes = (size_t)B * p.N * p.digitKS;
    const size_t ksd_al = (ksd_bytes + 255) / 256 * 256;
    uint8_t *ksd = (uint8_t *)g.scratch;
    uint64_t *ksb = (uint64_t *)((char *)g.scratch + ksd_al);
    const size_t part_off = (ksd_al + (size_t)B * 8 + 255) / 256 * 256;
    uint64_t *part = (uint64_t *)((char *)g.scratch + part_off);
    const size_t pacing_off = (part_off + (size_t)ks_part_rows(p, B) * (p.n + 1) * 8 + 255) / 256 * 256;
    BRArgs<W> a{};
    a.A = make_arith<W>(p);
    a.tw = make_tw<W>(c);
    a.c1 = g.c1;
    a.c2 = g.c2;
    a.gates = g.gates;
    a.bsk = k ? (const W *)k->bsk : nullptr;
    a.rgsw_elems = rgsw_elems(p);
    a.n = p.n;
    a.q = (uint32_t)p.q;
    a.digitR = p.digitR;
    a.Br = (uint32_t)p.Br;
    a.d = p.digitG;
    a.bits = p.BGbits;
    a.pb = c->tune.br_pb > 0 ? std::min(c->tune.br_pb, 2 * (int)p.digitG) : 2 * (int)p.digitG;
    if (c->tune.br_pb <= 0)
        while (a.pb > 1 && br_lds_bytes(p, a.pb) > c->lds_budget) a.pb--;
    for (int i = 0; i < 6; i++) a.gate_constant[i] = p.gate_constant[i];
    a.cap_acc = g.cap ? g.cap->acc : nullptr;
    a.ksd = ksd;
    a.ksb = ksb;
    a.dks = p.digitKS;
    a.Bks = (uint32_t)p.Bks;
    a.status = c->status;
    a.idx1 = g.idx1;
    a.idx2 = g.idx2;
    a.gidx = g.gidx;
    a.row_words = g.row_words ? g.row_words : p.n + 1;
    a.key_of = g.key_of;
    a.bsk_tab = (const W *const *)g.bsk_tab;
    a.ct_status = g.ct_status;
    a.mode = c->eval_mode.load();
    const size_t lds = br_lds_bytes(p, a.pb);
    if (lds > c->lds_budget) return fail("parameter set does not fit the LDS budget");
    const bool small = B <= c->tune.small_max && c->tune.br_threads_small;
    const int nt = small ? c->tune.br_threads_small : c->tune.br_threads;
    // Tiles serve dense throughput batches; gathered (serving) launches may
    // share the GPU with another launch in flight, where a workgroup of a
    // paced tile would wait (up to its spin limit) for one not yet resident.
    const int persist = small ? 0 : c->tune.br_persist;
    a.sync = nullptr;
    a.sync_groups = kPacingGroups;
    a.sync_window = c->tune.br_sync_window;
    a.spin_limit = c->tune.br_spin_limit;
    if (persist > 0 && !g.key_of && pacing_words(c, B) > 0) a.sync = (uint32_t *)((char *)g.scratch + pacing_off);
    {
        Prof pr(c, 0, g.stream, g.profile);
        HIP_TRY(launch_blind_rotate<W>(a, p.logN, nt, B, lds, g.stream, g.key_of ? 0 : persist, c->cus, nullptr));
    }
    KSArgs<W> s{};
    s.A = a.A;
    s.n = p.n;
    s.n_pad = n_pad(p);
    s.S = p.N * p.digitKS;
    s.Bks = (uint32_t)p.Bks;
    s.Q = p.Q;
    int lq = 0;
    while ((1ULL << lq) < p.q) lq++;
    s.lq1 = lq + 1;
    // one key-switch launch per key segment (a dense batch is one segment)
    const int nseg = g.nseg > 0 ? g.nseg : 1;
    for (int sg = 0; sg < nseg; sg++) {
        const int64_t o0 = g.nseg > 0 ? g.seg_off[sg] : 0, o1 = g.nseg > 0 ? g.seg_off[sg + 1] : B;
        if (o1 <= o0) continue;
        s.ksk = g.nseg > 0 ? (const W *)g.seg_ksk[sg] : (const W *)k->ksk;
        s.ksd = ksd + (size_t)o0 * s.S;
        s.ksb = ksb + o0;
        s.oidx = g.oidx ? g.oidx + o0 : nullptr;
        s.out_row_words = g.row_words ? g.row_words : p.n + 1;
        s.out = g.oidx ? g.out : g.out + o0 * (p.n + 1);
        s.cap_ks = g.cap && g.cap->ks ? g.cap->ks + o0 * (p.n + 1) : nullptr;
        s.B = o1 - o0;
        s.nt = c->tune.ks_threads > 0 ? c->tune.ks_threads : ks_staged_threads(p, s.B, c->cus);
        if (ks_staged_lds<W>(s.Bks, s.nt) > c->lds_budget) s.nt = 256;
        const bool gather = s.B <= c->tune.ks_gather_max || ks_staged_lds<W>(s.Bks, s.nt) > c->lds_budget;
        s.split = gather ? ks_gather_split(p, s.B) : 1;
        s.part = part;  // segments run in order on one stream and reuse it
        Prof pr(c, 1, g.stream, g.profile);
        HIP_TRY(launch_lwe_keyswitch<W>(s, gather, g.stream));
    }
    return 0;
}

int fheacc::launch_gate_path(fheacc_ctx *c, const fheacc_keys *k, const GateLaunch &g) {
    return c->p.narrow ? gate_path_impl<uint32_t>(c, k, g) : gate_path_impl<uint64_t>(c, k, g);
}

static int bootstrap_dense(fheacc_ctx *c, fheacc_keys *k, const int32_t *gates, const uint64_t *c1,
                           const uint64_t *c2, uint64_t *out, int64_t B, const fheacc_capture *cap) {
    if (int rc = ensure_scratch(c, gate_scratch_bytes(c, B))) return rc;
    GateLaunch g;
    g.gates = gates;
    g.c1 = c1;
    g.c2 = c2;
    g.out = out;
    g.B = B;
    g.cap = cap;
    g.scratch = c->scratch;
    g.stream = c->stream;
    g.profile = c->profile;
    return launch_gate_path(c, k, g);
}

int fheacc_bootstrap_batch_dev(fheacc_ctx *c, fheacc_keys *k, const int32_t *gates, const uint64_t *c1,
                               const uint64_t *c2, uint64_t *out, int64_t B, const fheacc_capture *cap) {
    if (check_ctx(c) || !k || !gates || !c1 || !c2 || !out) return fail("null argument");
    if (check_keys(c, k)) return -1;
    if (!k->bsk || !k->ksk) return fail("bootstrap and key-switch keys must be resident");
    if (B <= 0) return B == 0 ? 0 : fail("negative batch");
    if (B > 0x7fffffff) return fail("batch too large");
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return bootstrap_dense(c, k, gates, c1, c2, out, B, cap);
}

int fheacc_bootstrap_batch(fheacc_ctx *c, fheacc_keys *k, const int32_t *gates, const uint64_t *c1,
                           const uint64_t *c2, uint64_t *out, int64_t B, const fheacc_capture *cap) {
    if (check_ctx(c) || !k || !gates || !c1 || !c2 || !out) return fail("null argument");
    if (check_keys(c, k)) return -1;
    if (B <= 0) return B == 0 ? 0 : fail("negative batch");
    const HostParams &p = c->p;
    const size_t lw = (size_t)p.n + 1;
    for (int64_t i = 0; i < B; i++) {
        if (gates[i] < 0 || gates[i] > 5) return fail("gate out of range");
        if (!memcmp(c1 + i * lw, c2 + i * lw, lw * 8))
            return fail("identical operands passed to the gate bootstrap (RLWE.cpp:727-730)");
    }
    HIP_TRY(hipSetDevice(c->device));
    const size_t in_b = B * lw * 8, acc_b = cap && cap->acc ? B * 2 * (size_t)p.N * 8 : 0,
                 ks_b = cap && cap->ks ? in_b : 0;
    DevBuf buf;
    HIP_TRY(hipMalloc(&buf.p, 3 * in_b + acc_b + ks_b + B * 4 + 64));
    char *base = (char *)buf.p;
    uint64_t *d1 = (uint64_t *)base, *d2 = (uint64_t *)(base + in_b), *dout = (uint64_t *)(base + 2 * in_b);
    uint64_t *dacc = acc_b ? (uint64_t *)(base + 3 * in_b) : nullptr;
    uint64_t *dks = ks_b ? (uint64_t *)(base + 3 * in_b + acc_b) : nullptr;
    int32_t *dg = (int32_t *)(base + 3 * in_b + acc_b + ks_b);
    HIP_TRY(hipMemcpyAsync(d1, c1, in_b, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(d2, c2, in_b, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(dg, gates, B * 4, hipMemcpyHostToDevice, c->stream));
    fheacc_capture dcap{dacc, dks};
    if (int rc = fheacc_bootstrap_batch_dev(c, k, dg, d1, d2, dout, B, &dcap)) return rc;
    HIP_TRY(hipMemcpyAsync(out, dout, in_b, hipMemcpyDeviceToHost, c->stream));
    if (dacc) HIP_TRY(hipMemcpyAsync(cap->acc, dacc, acc_b, hipMemcpyDeviceToHost, c->stream));
    if (dks) HIP_TRY(hipMemcpyAsync(cap->ks, dks, ks_b, hipMemcpyDeviceToHost, c->stream));
    return fheacc_sync(c);
}

// ---------------------------------------------------------------- kernels

int fheacc_ntt_dev(fheacc_ctx *c, uint64_t *polys, int64_t count, int inverse) {
    if (check_ctx(c) || !polys) return fail("null argument");
    if (count <= 0) return 0;
    HIP_TRY(hipSetDevice(c->device));
    const HostParams &p = c->p;
    const int per_wg = p.N >= 1024 ? 2 : std::max(1, 2048 / p.N);
    if (p.narrow) {
        NTTArgs<uint32_t> a{make_arith<uint32_t>(p), make_tw<uint32_t>(c), polys, count, per_wg, inverse};
        Prof pr(c, 3);
        HIP_TRY(launch_ntt<uint32_t>(a, p.logN, c->stream));
    } else {
        NTTArgs<uint64_t> a{make_arith<uint64_t>(p), make_tw<uint64_t>(c), polys, count, per_wg, inverse};
        Prof pr(c, 3);
        HIP_TRY(launch_ntt<uint64_t>(a, p.logN, c->stream));
    }
    return 0;
}

// RLWE_mult_RGSW over count inputs acc + i in_stride -> out + i out_stride
// (words; 2N for dense batches).
template <typename W>
static int ep_impl(fheacc_ctx *c, const uint64_t *acc, const uint64_t *rgsw, int64_t stride, uint64_t *out,
                   int64_t count, int64_t in_stride = 0, int64_t out_stride = 0) {
    const HostParams &p = c->p;
    const int64_t nk = stride ? count : 1;
    const size_t ke = rgsw_elems(p);
    if (int rc = ensure_scratch(c, nk * ke * sizeof(W))) return rc;
    const size_t rw = (size_t)p.digitG * 4 * p.N;
    HIP_TRY(launch_rgsw_to_device<W>(rgsw, stride ? stride : (int64_t)rw, (W *)c->scratch, ke, p.digitG, p.N,
                                     make_sg(p.Q), nullptr, nk, c->stream));
    EPArgs<W> a{};
    a.A = make_arith<W>(p);
    a.tw = make_tw<W>(c);
    a.acc = acc;
    a.out = out;
    a.in_stride = in_stride ? in_stride : 2 * (int64_t)p.N;
    a.out_stride = out_stride ? out_stride : 2 * (int64_t)p.N;
    a.key = (const W *)c->scratch;
    a.key_stride = stride ? (int64_t)ke : 0;
    a.d = p.digitG;
    a.bits = p.BGbits;
    a.pb = pick_pb(p, 2 * p.digitG, (size_t)padded(2 * p.N) * sizeof(W), c->lds_budget, c->tune.br_pb);
    Prof pr(c, 2);
    HIP_TRY(launch_external_product<W>(a, p.logN, count, c->stream));
    return 0;
}

int fheacc_external_product_dev(fheacc_ctx *c, const uint64_t *acc, const uint64_t *rgsw, int64_t stride,
                                uint64_t *out, int64_t count) {
    if (check_ctx(c) || !acc || !rgsw || !out) return fail("null argument");
    if (count <= 0) return 0;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return c->p.narrow ? ep_impl<uint32_t>(c, acc, rgsw, stride, out, count)
                       : ep_impl<uint64_t>(c, acc, rgsw, stride, out, count);
}

template <typename W>
static int subs_impl(fheacc_ctx *c, fheacc_keys *k, const uint64_t *in, int f, uint64_t *out, int64_t count) {
    const HostParams &p = c->p;
    SubsArgs<W> a{};
    a.A = make_arith<W>(p);
    a.tw = make_tw<W>(c);
    a.in = in;
    a.out = out;
    a.key = (const W *)k->subs;
    a.key_stride = (int64_t)p.digitKS_R * p.N * (p.narrow ? 2 : 4);
    a.sub_factor = f;
    a.d = p.digitKS_R;
    a.bits = p.Bks_R_bits;
    a.dch = pick_pb(p, p.digitKS_R, 2 * (size_t)padded(2 * p.N) * sizeof(W), c->lds_budget, 0);
    if ((2 * (size_t)padded(2 * p.N) + padded(a.dch * p.N)) * sizeof(W) > c->lds_budget)
        return fail("subs key switch exceeds the LDS budget");
    HIP_TRY(launch_subs_keyswitch<W>(a, p.logN, count, c->stream));
    return 0;
}

int fheacc_subs_keyswitch_dev(fheacc_ctx *c, fheacc_keys *k, const uint64_t *in, int f, uint64_t *out,
                              int64_t count) {
    if (check_ctx(c) || !k || !in || !out) return fail("null argument");
    if (check_keys(c, k)) return -1;
    if (!k->subs) return fail("no subs key resident");
    if (f < 0 || f >= c->p.logN) return fail("sub_factor out of range");
    if (count <= 0) return 0;
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return c->p.narrow ? subs_impl<uint32_t>(c, k, in, f, out, count)
                       : subs_impl<uint64_t>(c, k, in, f, out, count);
}


// ------------------------------------------- RLWE expansion and CMUX chains

// expand_RLWE of B inputs (in[b], 2N words apart); slot k of expansion b is
// written at out + b set_stride + k slot_stride (words; 0: dense [B][N][2][N]).
template <typename W>
static int expand_impl(fheacc_ctx *c, fheacc_keys *k, const uint64_t *in, uint64_t *out, int64_t B,
                       int64_t set_stride = 0, int64_t slot_stride = 0) {
    const HostParams &p = c->p;
    const size_t rw = 2 * (size_t)p.N;
    if (!slot_stride) slot_stride = (int64_t)rw;
    if (!set_stride) set_stride = (int64_t)p.N * slot_stride;
    // slot 0 of every expansion starts as its input
    HIP_TRY(hipMemcpy2DAsync(out, (size_t)set_stride * 8, in, rw * 8, rw * 8, (size_t)B, hipMemcpyDeviceToDevice,
                             c->stream));
    SubsArgs<W> a{};
    a.set_stride = set_stride;
    a.slot_stride = slot_stride;
    a.A = make_arith<W>(p);
    a.tw = make_tw<W>(c);
    a.in = out;
    a.out = out;
    a.key = (const W *)k->subs;
    a.key_stride = (int64_t)p.digitKS_R * p.N * (p.narrow ? 2 : 4);
    a.d = p.digitKS_R;
    a.bits = p.Bks_R_bits;
    a.dch = pick_pb(p, p.digitKS_R, 2 * (size_t)padded(2 * p.N) * sizeof(W), c->lds_budget, 0);
    if ((2 * (size_t)padded(2 * p.N) + padded(a.dch * p.N)) * sizeof(W) > c->lds_budget)
        return fail("subs key switch exceeds the LDS budget");
    for (int i = 0; i < p.logN; i++) {
        a.sub_factor = i;
        Prof pr(c, 2);
        HIP_TRY(launch_expand_level<W>(a, p.logN, B, c->stream));
    }
    return 0;
}

int fheacc_expand_rlwe_dev(fheacc_ctx *c, fheacc_keys *k, const uint64_t *in, uint64_t *out, int64_t count) {
    if (check_ctx(c) || !k || !in || !out) return fail("null argument");
    if (check_keys(c, k)) return -1;
    if (!k->subs) return fail("no subs key resident");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    if ((count << c->p.logN) > 0x7fffffff) return fail("batch too large");
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return c->p.narrow ? expand_impl<uint32_t>(c, k, in, out, count) : expand_impl<uint64_t>(c, k, in, out, count);
}

// Rotation factor of poly_rotate_time (RLWE.cpp:49-68): reduced mod N with
// the sign of C++ '%', then X^rf as the exponent in [0, 2N).
static int rotate_exponent(int64_t rf, int N) {
    rf %= N;
    return (int)(rf >= 0 ? rf : 2 * N + rf);
}

template <typename W>
static int cmux_impl(fheacc_ctx *c, const uint64_t *c0, const uint64_t *c1, const uint64_t *rgsw, int64_t nk,
                     int64_t rgsw_stride, bool per_ct, const std::vector<int32_t> &rot, int pre_rot, uint64_t *out,
                     int64_t count) {
    const HostParams &p = c->p;
    const size_t ke = rgsw_elems(p);
    const size_t kbytes = (nk * ke * sizeof(W) + 255) / 256 * 256;
    if (int rc = ensure_scratch(c, kbytes + rot.size() * 4 + 4)) return rc;
    const size_t rw = (size_t)p.digitG * 4 * p.N;
    HIP_TRY(launch_rgsw_to_device<W>(rgsw, rgsw_stride ? rgsw_stride : (int64_t)rw, (W *)c->scratch, ke, p.digitG,
                                     p.N, make_sg(p.Q), nullptr, nk, c->stream));
    int32_t *drot = (int32_t *)((char *)c->scratch + kbytes);
    if (!rot.empty())
        HIP_TRY(hipMemcpyAsync(drot, rot.data(), rot.size() * 4, hipMemcpyHostToDevice, c->stream));
    CmuxArgs<W> a{};
    a.A = make_arith<W>(p);
    a.tw = make_tw<W>(c);
    a.c0 = c0;
    a.c1 = c1;
    a.in_stride = 2 * (int64_t)p.N;
    a.out = out;
    a.key = (const W *)c->scratch;
    a.key_stride = per_ct ? 0 : (int64_t)ke;
    a.key_ct_stride = per_ct ? (int64_t)ke : 0;
    a.rot = drot;
    a.steps = c1 ? 1 : (int)rot.size();
    a.pre_rot = pre_rot;
    a.d = p.digitG;
    a.bits = p.BGbits;
    a.pb = c->tune.br_pb > 0 ? std::min(c->tune.br_pb, 2 * (int)p.digitG) : 2 * (int)p.digitG;
    while (a.pb > 1 && br_lds_bytes(p, a.pb) > c->lds_budget) a.pb--;
    const size_t lds = br_lds_bytes(p, a.pb);
    if (lds > c->lds_budget) return fail("parameter set does not fit the LDS budget");
    Prof pr(c, 2);
    const int nt = c->tune.shape_set || p.N < 1024 ? c->tune.br_threads : 1024;
    HIP_TRY(launch_cmux_chain<W>(a, p.logN, nt, count, lds, c->stream));
    return 0;
}

static int cmux_dispatch(fheacc_ctx *c, const uint64_t *c0, const uint64_t *c1, const uint64_t *rgsw, int64_t nk,
                         int64_t stride, bool per_ct, const std::vector<int32_t> &rot, int pre_rot, uint64_t *out,
                         int64_t count) {
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return c->p.narrow ? cmux_impl<uint32_t>(c, c0, c1, rgsw, nk, stride, per_ct, rot, pre_rot, out, count)
                       : cmux_impl<uint64_t>(c, c0, c1, rgsw, nk, stride, per_ct, rot, pre_rot, out, count);
}

int fheacc_cmux_dev(fheacc_ctx *c, const uint64_t *c0, const uint64_t *c1, const uint64_t *rgsw, int64_t stride,
                    uint64_t *out, int64_t count) {
    if (check_ctx(c) || !c0 || !c1 || !rgsw || !out) return fail("null argument");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    if (count > 0x7fffffff) return fail("batch too large");
    return cmux_dispatch(c, c0, c1, rgsw, stride ? count : 1, stride, stride != 0, {}, 0, out, count);
}

int fheacc_blind_rotate_rlwe_dev(fheacc_ctx *c, const uint64_t *in, int forward, const uint64_t *rgsw, int steps,
                                 uint64_t *out, int64_t count) {
    if (check_ctx(c) || !in || !rgsw || !out) return fail("null argument");
    if (steps < 1 || steps > c->p.logN) return fail("steps must be in [1, log2 N]");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    if (count > 0x7fffffff) return fail("batch too large");
    // rotate_poly_forward[i] = NTT(X^(2^i)), rotate_poly_backward[i] =
    // NTT((Q - 1) X^(N - 2^i)) = NTT(X^(2N - 2^i)) (params.cpp:177-210)
    std::vector<int32_t> rot(steps);
    for (int i = 0; i < steps; i++) rot[i] = forward ? (1 << i) : 2 * c->p.N - (1 << i);
    return cmux_dispatch(c, in, nullptr, rgsw, steps, 0, false, rot, 0, out, count);
}

int fheacc_blind_rotate_full_dev(fheacc_ctx *c, const uint64_t *in, int b, const int32_t *a, const uint64_t *rgsw,
                                 int steps, uint64_t *out, int64_t count) {
    if (check_ctx(c) || !in || !out || (steps > 0 && (!a || !rgsw))) return fail("null argument");
    if (steps < 0 || steps > 4096) return fail("steps must be in [0, 4096]");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    if (count > 0x7fffffff) return fail("batch too large");
    std::vector<int32_t> rot(steps);
    for (int i = 0; i < steps; i++) rot[i] = rotate_exponent(a[i], c->p.N);
    // steps == 0 leaves only the public rotation (a copy when b reduces to 0)
    return cmux_dispatch(c, in, nullptr, steps ? rgsw : in, steps, 0, false, rot, rotate_exponent(b, c->p.N), out,
                         count);
}

// ------------------------------------------ homo_expand and the CMUX tree

// application::homo_expand (app.cpp:78-97) for count queries: the digitG
// expansions of each query are written straight into column 1 of the output
// RGSWs' rows (RGSW i, row j, column 1 = slot i of expansion j), then every
// row's column 0 is its column 1 times RGSW(-s), one batched external product.
template <typename W>
static int homo_expand_impl(fheacc_ctx *c, fheacc_keys *k, const uint64_t *packed, const uint64_t *enc_sec,
                            uint64_t *out, int64_t count) {
    const HostParams &p = c->p;
    const int64_t N = p.N, G = (int64_t)p.digitG * 4 * N;  // words per RGSW
    for (int64_t q = 0; q < count; q++)
        if (int rc = expand_impl<W>(c, k, packed + q * (int64_t)p.digitG * 2 * N, out + q * N * G + 2 * N,
                                    p.digitG, 4 * N, G))
            return rc;
    return ep_impl<W>(c, out + 2 * N, enc_sec, 0, out, count * N * p.digitG, 4 * N, 4 * N);
}

int fheacc_homo_expand_dev(fheacc_ctx *c, fheacc_keys *k, const uint64_t *packed, const uint64_t *enc_sec,
                           uint64_t *out, int64_t count) {
    if (check_ctx(c) || !k || !packed || !enc_sec || !out) return fail("null argument");
    if (check_keys(c, k)) return -1;
    if (!k->subs) return fail("no subs key resident");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    if ((count << c->p.logN) * (int64_t)c->p.digitG > 0x7fffffff) return fail("batch too large");
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return c->p.narrow ? homo_expand_impl<uint32_t>(c, k, packed, enc_sec, out, count)
                       : homo_expand_impl<uint64_t>(c, k, packed, enc_sec, out, count);
}

// The tree selection's CMUX tree (test.cpp:1576-1582) over count trees of
// 2^levels leaves: level l halves every tree with RGSW l (v[j/2] =
// CMUX(v[j], v[j+1], C[l])), ping-ponging between two scratch buffers; the
// last level writes out.
template <typename W>
static int cmux_tree_impl(fheacc_ctx *c, const uint64_t *leaves, int levels, const uint64_t *rgsw,
                          int64_t rgsw_stride, uint64_t *out, int64_t count) {
    const HostParams &p = c->p;
    const int64_t rw = 2 * (int64_t)p.N;
    if (levels == 0) {
        HIP_TRY(hipMemcpyAsync(out, leaves, (size_t)count * rw * 8, hipMemcpyDeviceToDevice, c->stream));
        return 0;
    }
    const size_t ke = rgsw_elems(p);
    const size_t kbytes = ((size_t)levels * ke * sizeof(W) + 255) / 256 * 256;
    const size_t pbytes = ((size_t)(count << (levels - 1)) * rw * 8 + 255) / 256 * 256;
    if (int rc = ensure_scratch(c, kbytes + 2 * pbytes)) return rc;
    const int64_t ref_words = (int64_t)p.digitG * 4 * p.N;
    HIP_TRY(launch_rgsw_to_device<W>(rgsw, rgsw_stride ? rgsw_stride : ref_words, (W *)c->scratch, ke, p.digitG,
                                     p.N, make_sg(p.Q), nullptr, levels, c->stream));
    uint64_t *buf[2] = {(uint64_t *)((char *)c->scratch + kbytes), (uint64_t *)((char *)c->scratch + kbytes + pbytes)};
    CmuxArgs<W> a{};
    a.A = make_arith<W>(p);
    a.tw = make_tw<W>(c);
    a.in_stride = 2 * rw;  // the leaf pairs (2p, 2p + 1) of every tree
    a.steps = 1;
    a.d = p.digitG;
    a.bits = p.BGbits;
    a.pb = c->tune.br_pb > 0 ? std::min(c->tune.br_pb, 2 * (int)p.digitG) : 2 * (int)p.digitG;
    while (a.pb > 1 && br_lds_bytes(p, a.pb) > c->lds_budget) a.pb--;
    const size_t lds = br_lds_bytes(p, a.pb);
    if (lds > c->lds_budget) return fail("parameter set does not fit the LDS budget");
    const int nt = c->tune.shape_set || p.N < 1024 ? c->tune.br_threads : 1024;
    const uint64_t *src = leaves;
    for (int l = 0; l < levels; l++) {
        uint64_t *dst = l == levels - 1 ? out : buf[l & 1];
        a.c0 = src;
        a.c1 = src + rw;
        a.out = dst;
        a.key = (const W *)c->scratch + l * ke;
        Prof pr(c, 2);
        HIP_TRY(launch_cmux_chain<W>(a, p.logN, nt, count << (levels - 1 - l), lds, c->stream));
        src = dst;
    }
    return 0;
}

int fheacc_cmux_tree_dev(fheacc_ctx *c, const uint64_t *leaves, int levels, const uint64_t *rgsw,
                         int64_t rgsw_stride, uint64_t *out, int64_t count) {
    if (check_ctx(c) || !leaves || !out || (levels > 0 && !rgsw)) return fail("null argument");
    if (levels < 0 || levels > 24) return fail("levels must be in [0, 24]");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    if ((count << levels) > 0x7fffffff) return fail("batch too large");
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return c->p.narrow ? cmux_tree_impl<uint32_t>(c, leaves, levels, rgsw, rgsw_stride, out, count)
                       : cmux_tree_impl<uint64_t>(c, leaves, levels, rgsw, rgsw_stride, out, count);
}

// ------------------------------------------ RLWE / RGSW encryption, decryption

// which: 0 RLWE_encrypt, 1 RGSW encryption, 2 RLWE_decrypt (host buffers; the
// work runs on the device with the library's seeded sampler)
template <typename W>
static int encrypt_impl(fheacc_ctx *c, int which, const uint64_t *sk, const uint64_t *in, int64_t count,
                        uint64_t seed, uint64_t first, uint64_t t, uint64_t *out) {
    const HostParams &p = c->p;
    const int64_t N = p.N;
    const int64_t in_words = which == 2 ? 2 * N : N;
    const int64_t out_words = which == 0 ? 2 * N : which == 1 ? (int64_t)p.digitG * 4 * N : N;
    DevBuf dz, dzn, dznp, din, dout, dpow, dpowp;
    HIP_TRY(hipMalloc(&dz.p, N * 8));
    HIP_TRY(hipMemcpy(dz.p, sk, N * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&dzn.p, N * sizeof(W)));
    HIP_TRY(hipMalloc(&dznp.p, N * sizeof(W)));
    HIP_TRY(hipMalloc(&din.p, (size_t)count * in_words * 8));
    HIP_TRY(hipMemcpy(din.p, in, (size_t)count * in_words * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&dout.p, (size_t)count * out_words * 8));
    std::vector<uint64_t> pw(p.digitG), pwp(p.digitG);
    for (uint32_t h = 0; h < p.digitG; h++) {
        pw[h] = h ? mulmodh(pw[h - 1], (1ULL << p.BGbits) % p.Q, p.Q) : 1 % p.Q;
        pwp[h] = p.narrow ? shoup32h((uint32_t)pw[h], (uint32_t)p.Q) : shoup64h(pw[h], p.Q);
    }
    HIP_TRY(hipMalloc(&dpow.p, p.digitG * 8));
    HIP_TRY(hipMalloc(&dpowp.p, p.digitG * 8));
    HIP_TRY(hipMemcpy(dpow.p, pw.data(), p.digitG * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dpowp.p, pwp.data(), p.digitG * 8, hipMemcpyHostToDevice));
    EncArgs<W> e{};
    KeygenArgs<W> &kg = e.k;
    kg.A = make_arith<W>(p);
    kg.tw = make_tw<W>(c);
    kg.seed = seed;
    kg.cdt = c->t.cdt;
    kg.cdt_len = c->t.cdt_len;
    kg.E_rlwe = p.E_rlwe;
    kg.zn = (W *)dzn.p;
    kg.znp = (W *)dznp.p;
    kg.sg = make_sg(p.Q);
    kg.d = p.digitG;
    kg.bg_pow = (const uint64_t *)dpow.p;
    kg.bg_powp = (const uint64_t *)dpowp.p;
    e.msg = (const uint64_t *)din.p;
    e.ct = (const uint64_t *)din.p;
    e.out = (uint64_t *)dout.p;
    e.first = first;
    e.scale = t ? p.Q / t : 1;
    HIP_TRY(launch_secret_ntt<W>(kg, p.logN, (const uint64_t *)dz.p, c->stream));
    HIP_TRY(launch_encrypt<W>(e, which, p.logN, count, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(out, dout.p, (size_t)count * out_words * 8, hipMemcpyDeviceToHost));
    return 0;
}

static int encrypt_call(fheacc_ctx *c, int which, const uint64_t *sk, const uint64_t *in, int64_t count,
                        uint64_t seed, uint64_t first, uint64_t t, uint64_t *out) {
    if (check_ctx(c) || !sk || !in || !out) return fail("null argument");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    if (count * (int64_t)c->p.digitG > 0x7fffffff) return fail("batch too large");
    std::lock_guard<std::mutex> lk(c->mu);
    HIP_TRY(hipSetDevice(c->device));
    return c->p.narrow ? encrypt_impl<uint32_t>(c, which, sk, in, count, seed, first, t, out)
                       : encrypt_impl<uint64_t>(c, which, sk, in, count, seed, first, t, out);
}

int fheacc_rlwe_encrypt(fheacc_ctx *c, const uint64_t *sk, const uint64_t *msg, int64_t count, uint64_t seed,
                        uint64_t first_index, uint64_t *out) {
    return encrypt_call(c, 0, sk, msg, count, seed, first_index, 0, out);
}

int fheacc_rgsw_encrypt(fheacc_ctx *c, const uint64_t *sk, const uint64_t *msg, int64_t count, uint64_t seed,
                        uint64_t first_index, uint64_t *out) {
    return encrypt_call(c, 1, sk, msg, count, seed, first_index, 0, out);
}

int fheacc_rlwe_decrypt(fheacc_ctx *c, const uint64_t *sk, const uint64_t *ct, int64_t count, uint64_t t,
                        uint64_t *out) {
    if (c && (t < 2 || t > c->p.Q)) return fail("plaintext modulus t must be in [2, Q]");
    return encrypt_call(c, 2, sk, ct, count, 0, 0, t, out);
}

// ------------------------------------- host-buffer forms of the RLWE calls

namespace {
// Device copies of host inputs, device outputs, and the copy back, around
// one *_dev call on the context stream.
struct HostCall {
    fheacc_ctx *c;
    std::vector<void *> bufs;
    explicit HostCall(fheacc_ctx *c_) : c(c_) {}
    ~HostCall() {
        for (void *p : bufs) (void)hipFree(p);
    }
    int in(const void *h, size_t bytes, uint64_t **d) {
        *d = nullptr;
        if (!bytes) return 0;
        HIP_TRY(hipMalloc((void **)d, bytes));
        bufs.push_back(*d);
        HIP_TRY(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
        return 0;
    }
    int out(size_t bytes, uint64_t **d) {
        HIP_TRY(hipMalloc((void **)d, bytes));
        bufs.push_back(*d);
        return 0;
    }
    int back(void *h, const uint64_t *d, size_t bytes) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        HIP_TRY(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
        return 0;
    }
};
}  // namespace

int fheacc_expand_rlwe(fheacc_ctx *c, fheacc_keys *k, const uint64_t *in, uint64_t *out, int64_t count) {
    if (check_ctx(c) || !k || !in || !out) return fail("null argument");
    if (check_keys(c, k)) return -1;
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    const size_t rb = 2 * (size_t)c->p.N * 8;
    HIP_TRY(hipSetDevice(c->device));
    HostCall h(c);
    uint64_t *di, *dout;
    if (int rc = h.in(in, count * rb, &di)) return rc;
    if (int rc = h.out(count * rb * c->p.N, &dout)) return rc;
    if (int rc = fheacc_expand_rlwe_dev(c, k, di, dout, count)) return rc;
    return h.back(out, dout, count * rb * c->p.N);
}

int fheacc_subs_keyswitch(fheacc_ctx *c, fheacc_keys *k, const uint64_t *in, int sub_factor, uint64_t *out,
                          int64_t count) {
    if (check_ctx(c) || !k || !in || !out) return fail("null argument");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    const size_t rb = 2 * (size_t)c->p.N * 8;
    HIP_TRY(hipSetDevice(c->device));
    HostCall h(c);
    uint64_t *di, *dout;
    if (int rc = h.in(in, count * rb, &di)) return rc;
    if (int rc = h.out(count * rb, &dout)) return rc;
    if (int rc = fheacc_subs_keyswitch_dev(c, k, di, sub_factor, dout, count)) return rc;
    return h.back(out, dout, count * rb);
}

int fheacc_homo_expand(fheacc_ctx *c, fheacc_keys *k, const uint64_t *packed, const uint64_t *enc_sec,
                       uint64_t *out, int64_t count) {
    if (check_ctx(c) || !k || !packed || !enc_sec || !out) return fail("null argument");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    const HostParams &p = c->p;
    const size_t gb = (size_t)p.digitG * 4 * p.N * 8;
    HIP_TRY(hipSetDevice(c->device));
    HostCall h(c);
    uint64_t *dp, *de, *dout;
    if (int rc = h.in(packed, count * p.digitG * 2 * (size_t)p.N * 8, &dp)) return rc;
    if (int rc = h.in(enc_sec, gb, &de)) return rc;
    if (int rc = h.out(count * (size_t)p.N * gb, &dout)) return rc;
    if (int rc = fheacc_homo_expand_dev(c, k, dp, de, dout, count)) return rc;
    return h.back(out, dout, count * (size_t)p.N * gb);
}

int fheacc_cmux_tree(fheacc_ctx *c, const uint64_t *leaves, int levels, const uint64_t *rgsw, int64_t rgsw_stride,
                     uint64_t *out, int64_t count) {
    if (check_ctx(c) || !leaves || !out || (levels > 0 && !rgsw)) return fail("null argument");
    if (levels < 0 || levels > 24) return fail("levels must be in [0, 24]");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    const HostParams &p = c->p;
    const size_t rb = 2 * (size_t)p.N * 8;
    const int64_t stride = rgsw_stride ? rgsw_stride : (int64_t)p.digitG * 4 * p.N;
    HIP_TRY(hipSetDevice(c->device));
    HostCall h(c);
    uint64_t *dl, *dg = nullptr, *dout;
    if (int rc = h.in(leaves, (size_t)(count << levels) * rb, &dl)) return rc;
    if (levels > 0)
        if (int rc = h.in(rgsw, ((size_t)(levels - 1) * stride + (size_t)p.digitG * 4 * p.N) * 8, &dg)) return rc;
    if (int rc = h.out(count * rb, &dout)) return rc;
    if (int rc = fheacc_cmux_tree_dev(c, dl, levels, dg, stride, dout, count)) return rc;
    return h.back(out, dout, count * rb);
}

int fheacc_external_product(fheacc_ctx *c, const uint64_t *acc, const uint64_t *rgsw, int64_t stride, uint64_t *out,
                            int64_t count) {
    if (check_ctx(c) || !acc || !rgsw || !out) return fail("null argument");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    const size_t rb = 2 * (size_t)c->p.N * 8, gb = (size_t)c->p.digitG * 4 * c->p.N * 8;
    HIP_TRY(hipSetDevice(c->device));
    HostCall h(c);
    uint64_t *da, *dg, *dout;
    if (int rc = h.in(acc, count * rb, &da)) return rc;
    if (int rc = h.in(rgsw, stride ? count * stride * 8 : gb, &dg)) return rc;
    if (int rc = h.out(count * rb, &dout)) return rc;
    if (int rc = fheacc_external_product_dev(c, da, dg, stride, dout, count)) return rc;
    return h.back(out, dout, count * rb);
}

int fheacc_cmux(fheacc_ctx *c, const uint64_t *c0, const uint64_t *c1, const uint64_t *rgsw, int64_t stride,
                uint64_t *out, int64_t count) {
    if (check_ctx(c) || !c0 || !c1 || !rgsw || !out) return fail("null argument");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    const size_t rb = 2 * (size_t)c->p.N * 8, gb = (size_t)c->p.digitG * 4 * c->p.N * 8;
    HIP_TRY(hipSetDevice(c->device));
    HostCall h(c);
    uint64_t *d0, *d1, *dg, *dout;
    if (int rc = h.in(c0, count * rb, &d0)) return rc;
    if (int rc = h.in(c1, count * rb, &d1)) return rc;
    if (int rc = h.in(rgsw, stride ? count * stride * 8 : gb, &dg)) return rc;
    if (int rc = h.out(count * rb, &dout)) return rc;
    if (int rc = fheacc_cmux_dev(c, d0, d1, dg, stride, dout, count)) return rc;
    return h.back(out, dout, count * rb);
}

int fheacc_blind_rotate_rlwe(fheacc_ctx *c, const uint64_t *in, int forward, const uint64_t *rgsw, int steps,
                             uint64_t *out, int64_t count) {
    if (check_ctx(c) || !in || !rgsw || !out) return fail("null argument");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    if (steps < 1 || steps > c->p.logN) return fail("steps must be in [1, log2 N]");
    const size_t rb = 2 * (size_t)c->p.N * 8, gb = (size_t)c->p.digitG * 4 * c->p.N * 8;
    HIP_TRY(hipSetDevice(c->device));
    HostCall h(c);
    uint64_t *di, *dg, *dout;
    if (int rc = h.in(in, count * rb, &di)) return rc;
    if (int rc = h.in(rgsw, steps * gb, &dg)) return rc;
    if (int rc = h.out(count * rb, &dout)) return rc;
    if (int rc = fheacc_blind_rotate_rlwe_dev(c, di, forward, dg, steps, dout, count)) return rc;
    return h.back(out, dout, count * rb);
}

int fheacc_blind_rotate_full(fheacc_ctx *c, const uint64_t *in, int b, const int32_t *a, const uint64_t *rgsw,
                             int steps, uint64_t *out, int64_t count) {
    if (check_ctx(c) || !in || !out || (steps > 0 && (!a || !rgsw))) return fail("null argument");
    if (count <= 0) return count == 0 ? 0 : fail("negative count");
    if (steps < 0 || steps > 4096) return fail("steps must be in [0, 4096]");
    const size_t rb = 2 * (size_t)c->p.N * 8, gb = (size_t)c->p.digitG * 4 * c->p.N * 8;
    HIP_TRY(hipSetDevice(c->device));
    HostCall h(c);
    uint64_t *di, *dg = nullptr, *dout;
    if (int rc = h.in(in, count * rb, &di)) return rc;
    if (int rc = h.in(rgsw, steps * gb, &dg)) return rc;
    if (int rc = h.out(count * rb, &dout)) return rc;
    if (int rc = fheacc_blind_rotate_full_dev(c, di, b, a, dg, steps, dout, count)) return rc;
    return h.back(out, dout, count * rb);
}

// ---------------------------------------------------------------- tuning

int fheacc_ctx_set_tuning(fheacc_ctx *c, int threads, int pb) {
    if (check_ctx(c)) return -1;
    if (threads != 128 && threads != 256 && threads != 512 && threads != 1024)
        return fail("threads must be 128, 256, 512 or 1024");
    if (threads == 1024 && c->p.N < 1024) return fail("1024 threads need N >= 1024");
    if (pb < 0 || pb > 2 * (int)c->p.digitG) return fail("polys per batch out of range");
    std::lock_guard<std::mutex> lk(c->mu);
    c->tune.br_threads = threads;
    c->tune.br_pb = pb;
    c->tune.br_threads_small = 0;
    c->tune.shape_set = true;
    return 0;
}

int fheacc_ctx_set_schedule(fheacc_ctx *c, int persist, int sync_window) {
    if (check_ctx(c)) return -1;
    if (persist < 0 || persist > 64) return fail("persistent workgroups per CU must be in [0, 64]");
    if (sync_window < 0 || (sync_window > 0 && persist == 0)) return fail("a sync window needs a persistent grid");
    std::lock_guard<std::mutex> lk(c->mu);
    c->tune.br_persist = persist;
    c->tune.br_sync_window = sync_window;
    return 0;
}

int fheacc_ctx_get_schedule(const fheacc_ctx *c, int *persist, int *sync_window) {
    if (check_ctx(c) || !persist || !sync_window) return fail("null argument");
    *persist = c->tune.br_persist;
    *sync_window = c->tune.br_sync_window;
    return 0;
}

int fheacc_ctx_get_tuning(const fheacc_ctx *c, int *threads, int *pb) {
    if (check_ctx(c) || !threads || !pb) return fail("null argument");
    *threads = c->tune.br_threads;
    *pb = c->tune.br_pb;
    return 0;
}

int fheacc_ctx_set_eval_mode(fheacc_ctx *c, int mode) {
    if (check_ctx(c)) return fail("null context");
    if (mode != FHEACC_EVAL_SOFTWARE && mode != FHEACC_EVAL_FPGA_X1 && mode != FHEACC_EVAL_FPGA_XN)
        return fail("unknown evaluation mode " + std::to_string(mode));
    c->eval_mode.store(mode);
    return 0;
}

int fheacc_ctx_get_eval_mode(const fheacc_ctx *c, int *mode) {
    if (check_ctx(c) || !mode) return fail("null argument");
    *mode = c->eval_mode.load();
    return 0;
}

// Synthetic operands: uniform words mod q (c1 != c2 almost surely; equal
// pairs only raise the status flag, which is cleared afterwards).
__global__ void k_fill_lwe(uint64_t *c1, uint64_t *c2, int32_t *g, int64_t words, uint64_t qm, int64_t B) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < words; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ULL + 0x2545F4914F6CDD1DULL;
        h ^= h >> 29;
        h *= 0xBF58476D1CE4E5B9ULL;
        h ^= h >> 32;
        c1[i] = h & qm;
        c2[i] = (h >> 20) & qm;
        if (i < B) g[i] = 3;
    }
}

int fheacc_autotune(fheacc_ctx *c, fheacc_keys *k, int64_t B, char *report, size_t report_len) {
    if (check_ctx(c) || !k) return fail("null argument");
    if (check_keys(c, k)) return -1;
    if (!k->bsk || !k->ksk) return fail("bootstrap and key-switch keys must be resident");
    if (B <= 0 || B > (1 << 20)) return fail("tuning batch out of range");
    HIP_TRY(hipSetDevice(c->device));
    const HostParams &p = c->p;
    const int64_t words = B * (p.n + 1);
    DevBuf buf;
    HIP_TRY(hipMalloc(&buf.p, 3 * words * 8 + B * 4 + 64));
    uint64_t *c1 = (uint64_t *)buf.p, *c2 = c1 + words, *out = c2 + words;
    int32_t *g = (int32_t *)(out + words);
    hipLaunchKernelGGL(k_fill_lwe, dim3(256), dim3(256), 0, c->stream, c1, c2, g, words, p.q - 1, B);
    HIP_TRY(hipGetLastError());
    const Tuning saved = c->tune;
    const bool saved_prof = c->profile;
    // blind-rotation time of one shape on the first Bt synthetic ciphertexts
    auto time_shape = [&](int threads, int pb, int64_t Bt, double *br) -> int {
        c->tune.br_threads = threads;
        c->tune.br_pb = pb;
        c->tune.br_threads_small = 0;
        {
            std::lock_guard<std::mutex> lk(c->mu);
            c->profile = false;
        }
        int rc = fheacc_bootstrap_batch_dev(c, k, g, c1, c2, out, Bt, nullptr);  // warm
        if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail("synchronize");
        {
            std::lock_guard<std::mutex> lk(c->mu);
            drop_events(c);
            c->profile = true;
        }
        for (int r = 0; r < 2 && rc == 0; r++) rc = fheacc_bootstrap_batch_dev(c, k, g, c1, c2, out, Bt, nullptr);
        double ms[4];
        int64_t cnt[4];
        if (rc == 0) rc = fheacc_ctx_kernel_ms(c, ms, cnt);
        if (rc == 0) *br = ms[0] / std::max<int64_t>(cnt[0], 1);
        return rc;
    };
    auto restore = [&](int rc) {
        c->tune = saved;
        std::lock_guard<std::mutex> lk(c->mu);
        c->profile = saved_prof;
        drop_events(c);
        return rc;
    };
    std::string rep = "{\"batch\": " + std::to_string(B) + ", \"candidates\": [";
    double best = 1e30;
    int best_t = saved.br_threads, best_pb = saved.br_pb;
    bool first = true;
    // digit batches: even sizes up to 2d (2d: the whole batch in place) and d
    // (in place in two batches, half the LDS)
    std::vector<int> pbs;
    for (int pb = 2; pb <= 2 * (int)p.digitG; pb += 2) pbs.push_back(pb);
    if (p.digitG & 1) pbs.push_back((int)p.digitG);
    std::sort(pbs.begin(), pbs.end());
    for (int threads : {128, 256, 512, 1024}) {
        if (threads > p.N || (threads == 1024 && p.N < 1024)) continue;
        for (int pb : pbs) {
            const size_t lds = br_lds_bytes(p, pb);
            if (lds > c->lds_budget) continue;
            double br = 0;
            if (int rc = time_shape(threads, pb, B, &br)) return restore(rc);
            if (!first) rep += ", ";
            first = false;
            rep += "{\"threads\": " + std::to_string(threads) + ", \"polys_per_batch\": " + std::to_string(pb) +
                   ", \"lds_bytes\": " + std::to_string(lds) + ", \"blind_rotate_ms\": " + std::to_string(br) + "}";
            if (br < best) {
                best = br;
                best_t = threads;
                best_pb = pb;
            }
        }
    }
    // the schedule at the chosen shape: one workgroup per ciphertext, tiles of
    // resident workgroups, or the persistent grid paced within a window of
    // steps (as many workgroups per CU as fit).  Another schedule replaces
    // one workgroup per ciphertext only when it is at least 3 % faster:
    // repeated timings of one shape differ by 1-2 %.
    std::string rep_sch;
    int best_persist = saved.br_persist, best_window = saved.br_sync_window;
    {
        double best_sch = 1e30, plain_ms = 0;
        const int cand[][2] = {{0, 0}, {64, 0}, {64, 8}, {64, 16}, {64, 32}};
        for (const auto &cd : cand) {
            c->tune.br_persist = cd[0];
            c->tune.br_sync_window = cd[1];
            double br = 0;
            if (int rc = time_shape(best_t, best_pb, B, &br)) return restore(rc);
            rep_sch += std::string(rep_sch.empty() ? "" : ", ") + "{\"persist\": " + std::to_string(cd[0]) +
                       ", \"sync_window\": " + std::to_string(cd[1]) + ", \"blind_rotate_ms\": " +
                       std::to_string(br) + "}";
            if (cd[0] == 0 && cd[1] == 0) plain_ms = br;
            if (br < best_sch && (cd[0] == 0 || br < 0.97 * plain_ms)) {
                best_sch = br;
                best_persist = cd[0];
                best_window = cd[1];
            }
        }
        c->tune.br_persist = saved.br_persist;
        c->tune.br_sync_window = saved.br_sync_window;
    }
    // the latency shape for batches up to one workgroup per CU, at the same
    // digit batching (every shape gives the same words)
    const int64_t Bs = std::min<int64_t>(c->tune.small_max, B);
    double best_s = 1e30;
    int best_ts = best_t;
    std::string rep_s;
    for (int threads : {128, 256, 512, 1024}) {
        if (threads > p.N || Bs <= 0 || (threads == 1024 && p.N < 1024)) continue;
        double br = 0;
        if (int rc = time_shape(threads, best_pb, Bs, &br)) return restore(rc);
        rep_s += std::string(rep_s.empty() ? "" : ", ") + "{\"threads\": " + std::to_string(threads) +
                 ", \"blind_rotate_ms\": " + std::to_string(br) + "}";
        if (br < best_s) {
            best_s = br;
            best_ts = threads;
        }
    }
    restore(0);
    c->tune.br_threads = best_t;
    c->tune.br_pb = best_pb;
    c->tune.br_threads_small = Bs > 0 ? best_ts : 0;
    c->tune.br_persist = best_persist;
    c->tune.br_sync_window = best_window;
    HIP_TRY(hipMemset(c->status, 0, sizeof(int)));
    rep += "], \"chosen\": {\"threads\": " + std::to_string(best_t) + ", \"polys_per_batch\": " +
           std::to_string(best_pb) + ", \"blind_rotate_ms\": " + std::to_string(best) + "}" +
           ", \"schedule\": {\"candidates\": [" + rep_sch + "], \"persist\": " + std::to_string(best_persist) +
           ", \"sync_window\": " + std::to_string(best_window) + "}" +
           ", \"small_batch\": {\"max_batch\": " + std::to_string(c->tune.small_max) + ", \"tuned_at\": " +
           std::to_string(Bs) + ", \"candidates\": [" + rep_s + "], \"threads\": " + std::to_string(best_ts) +
           "}}";
    if (report && report_len) {
        strncpy(report, rep.c_str(), report_len - 1);
        report[report_len - 1] = 0;
    }
    return 0;
}

// ---------------------------------------------------------------- HIP graphs

struct fheacc_graph {
    fheacc_ctx *ctx = nullptr;  // launches on this context's stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // one scratch buffer per set, owned by the graph: later eager calls may
    // grow (reallocate) their context's scratch
    std::vector<void *> scratch;
    // the contexts and key sets the captured kernels read (held until destroy)
    std::vector<fheacc_ctx *> ctxs;
    std::vector<fheacc_keys *> keys;
};

// Captures the gate path of every set into one graph, set after set on one
// capture stream, each set with a scratch buffer of its own.  proto[i] holds
// set i's buffers (and gathered rows, for an interleaved batch); `owned`
// device buffers pass to the graph.
static int capture_sets(int nsets, fheacc_ctx *const *ctxs, fheacc_keys *const *keys,
                        const std::vector<GateLaunch> &proto, std::vector<void *> owned, fheacc_graph **g) {
    auto drop_owned = [&](int rc) {
        for (void *p : owned) (void)hipFree(p);
        return rc;
    };
    if (!g || nsets < 1 || nsets > 16 || !ctxs || !keys) return drop_owned(fail("invalid argument"));
    *g = nullptr;
    for (int i = 0; i < nsets; i++) {
        const GateLaunch &pl = proto[i];
        if (check_ctx(ctxs[i]) || !keys[i] || !pl.gates || !pl.c1 || !pl.c2 || !pl.out || pl.B < 0 ||
            pl.B > 0x7fffffff)
            return drop_owned(fail("invalid argument for set " + std::to_string(i)));
        if (keys[i]->ctx != ctxs[i])
            return drop_owned(fail("keys of set " + std::to_string(i) + " belong to another context"));
        if (!keys[i]->bsk || !keys[i]->ksk)
            return drop_owned(fail("bootstrap and key-switch keys must be resident"));
        if (ctxs[i]->device != ctxs[0]->device) return drop_owned(fail("all sets must be on one device"));
    }
    if (hipSetDevice(ctxs[0]->device) != hipSuccess) return drop_owned(fail("hipSetDevice failed"));
    std::unique_ptr<fheacc_graph, int (*)(fheacc_graph *)> gr(new fheacc_graph(), fheacc_graph_destroy);
    gr->ctx = ctxs[0];
    gr->scratch = std::move(owned);
    for (int i = 0; i < nsets; i++) {
        ctx_retain(ctxs[i]);
        gr->ctxs.push_back(ctxs[i]);
        keys_retain(keys[i]);
        gr->keys.push_back(keys[i]);
    }
    std::vector<void *> set_scratch(nsets, nullptr);
    for (int i = 0; i < nsets; i++) {
        if (proto[i].B == 0) continue;
        HIP_TRY(hipMalloc(&set_scratch[i], gate_scratch_bytes(ctxs[i], proto[i].B)));
        gr->scratch.push_back(set_scratch[i]);
    }
    for (int i = 0; i < nsets; i++) HIP_TRY(hipStreamSynchronize(ctxs[i]->stream));
    // Capture on a stream of its own (the null stream cannot be captured).
    // The sets run one after another: every set of a mixed batch fills the
    // chip, and concurrent branches (a fork per set) were measured 3.5 %
    // slower once the 64-bit set runs 1024-thread workgroups, which share
    // CUs with the 32-bit set's and compete for the VALU.
    hipStream_t root = nullptr;
    int rc = 0;
    if (hipStreamCreateWithFlags(&root, hipStreamNonBlocking) != hipSuccess) rc = fail("stream creation");
    hipGraph_t graph = nullptr;
    if (!rc) {
        if (hipStreamBeginCapture(root, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = fail("begin capture");
        for (int i = 0; i < nsets && !rc; i++) {
            if (proto[i].B == 0) continue;
            GateLaunch gl = proto[i];
            gl.scratch = set_scratch[i];
            gl.stream = root;
            gl.profile = false;
            rc = launch_gate_path(ctxs[i], keys[i], gl);
        }
        hipError_t e = hipStreamEndCapture(root, &graph);
        if (!rc && e != hipSuccess) rc = fail(std::string("end capture: ") + hipGetErrorString(e));
    }
    if (root) (void)hipStreamDestroy(root);
    gr->graph = graph;
    if (!rc && hipGraphInstantiate(&gr->exec, graph, nullptr, nullptr, 0) != hipSuccess) rc = fail("instantiate");
    if (rc) return rc;
    *g = gr.release();
    return 0;
}

int fheacc_graph_capture_bootstrap(fheacc_ctx *c, fheacc_keys *k, const int32_t *gates, const uint64_t *c1,
                                   const uint64_t *c2, uint64_t *out, int64_t B, fheacc_graph **g) {
    return fheacc_graph_capture_mixed(1, &c, &k, &gates, &c1, &c2, &out, &B, g);
}

int fheacc_graph_capture_mixed(int nsets, fheacc_ctx *const *ctxs, fheacc_keys *const *keys,
                               const int32_t *const *gates, const uint64_t *const *c1, const uint64_t *const *c2,
                               uint64_t *const *out, const int64_t *B, fheacc_graph **g) {
    if (!g || nsets < 1 || nsets > 16 || !gates || !c1 || !c2 || !out || !B) return fail("invalid argument");
    std::vector<GateLaunch> proto(nsets);
    for (int i = 0; i < nsets; i++) {
        if (B[i] <= 0) return fail("invalid argument for set " + std::to_string(i));
        proto[i].gates = gates[i];
        proto[i].c1 = c1[i];
        proto[i].c2 = c2[i];
        proto[i].out = out[i];
        proto[i].B = B[i];
    }
    return capture_sets(nsets, ctxs, keys, proto, {}, g);
}

int fheacc_graph_capture_interleaved(int nsets, fheacc_ctx *const *ctxs, fheacc_keys *const *keys,
                                     const int32_t *set_of, const int32_t *gates, const uint64_t *c1,
                                     const uint64_t *c2, uint64_t *out, int64_t total, int64_t row_words,
                                     fheacc_graph **g) {
    if (!g || nsets < 1 || nsets > 16 || !ctxs || !keys || !set_of || !gates || !c1 || !c2 || !out)
        return fail("invalid argument");
    if (total <= 0 || total > 0x7fffffff) return fail("total out of range");
    std::vector<std::vector<int64_t>> rows(nsets);
    for (int64_t r = 0; r < total; r++) {
        if (set_of[r] < 0 || set_of[r] >= nsets) return fail("row " + std::to_string(r) + ": set out of range");
        rows[set_of[r]].push_back(r);
    }
    for (int i = 0; i < nsets; i++) {
        if (check_ctx(ctxs[i])) return -1;
        if (row_words < ctxs[i]->p.n + 1)
            return fail("row_words must hold the largest LWE ciphertext (n + 1 words of set " + std::to_string(i) +
                        ")");
    }
    // the row indices of every set, back to back, in a device buffer the graph owns
    HIP_TRY(hipSetDevice(ctxs[0]->device));
    DevBuf idx;
    HIP_TRY(hipMalloc(&idx.p, (size_t)total * 8));
    std::vector<GateLaunch> proto(nsets);
    int64_t off = 0;
    for (int i = 0; i < nsets; i++) {
        const int64_t *d = (const int64_t *)idx.p + off;
        if (!rows[i].empty())
            HIP_TRY(hipMemcpy((void *)d, rows[i].data(), rows[i].size() * 8, hipMemcpyHostToDevice));
        GateLaunch &pl = proto[i];
        pl.gates = gates;
        pl.c1 = c1;
        pl.c2 = c2;
        pl.out = out;
        pl.B = (int64_t)rows[i].size();
        pl.idx1 = pl.idx2 = pl.oidx = pl.gidx = d;
        pl.row_words = row_words;
        off += pl.B;
    }
    void *owned = idx.p;
    idx.p = nullptr;
    return capture_sets(nsets, ctxs, keys, proto, {owned}, g);
}

int fheacc_graph_launch(fheacc_graph *g) {
    if (!g || !g->exec) return fail("null graph");
    HIP_TRY(hipSetDevice(g->ctx->device));
    HIP_TRY(hipGraphLaunch(g->exec, g->ctx->stream));
    return 0;
}

int fheacc_graph_destroy(fheacc_graph *g) {
    if (!g) return 0;
    (void)hipSetDevice(g->ctx ? g->ctx->device : 0);
    if (g->exec) {
        // a replay may still be running on the context stream
        if (g->ctx) (void)hipStreamSynchronize(g->ctx->stream);
        (void)hipGraphExecDestroy(g->exec);
    }
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (void *p : g->scratch) (void)hipFree(p);
    for (fheacc_keys *k : g->keys) keys_release(k);
    for (fheacc_ctx *c : g->ctxs) ctx_release(c);
    delete g;
    return 0;
}
