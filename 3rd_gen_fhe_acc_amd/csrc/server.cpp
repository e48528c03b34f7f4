// server.cpp -- the serving loop: gate-circuit jobs from many clients,
// admitted by device-memory footprint, their ready gates coalesced into
// gate-path launches while earlier launches are still on the GPU.
//
// The reference has no serving layer: its apps evaluate one gate
// (RLWE::eval_bootstrap, RLWE.cpp:726-786) or one fixed-size batch
// (FPGA::fpga_eval_bootstrap_x{4,8,12,16}, myfpga.cpp:1003-1660) per call.
// Each gate launched here is that same evaluation, word for word.
//
// Memory: every ciphertext a job owns (its inputs and one per gate) is a row
// of (n+1) words in one device pool of hbm_budget bytes, allocated once.  A
// job is admitted, in arrival order, when a contiguous range of rows for it
// is free; it returns the range when its outputs are in host memory.
// Inside the range the output gates come last, in output order, so the
// outputs leave in one copy.
//
// Launch: a launch takes ready gates of the admitted jobs up to max_batch
// (with fair_share, each job first gets an equal share, so that a small job
// is not queued behind a large one's ready level; then oldest job first),
// orders them by client and runs the blind rotation once
// for all of them (each ciphertext indexes its client's bootstrap key) and
// the key switch once per client segment.  Up to max_inflight launches run
// at once, each on its own stream with its own scratch and descriptors.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fheacc_internal.hpp"

using namespace fheacc;

namespace {

constexpr int kMaxClients = 64;
constexpr int kMaxInflight = 4;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Job {
    int64_t id = 0;
    int32_t client = 0;
    int64_t n_in = 0, n_gates = 0, n_out = 0;
    std::vector<uint64_t> in_words;
    std::vector<int32_t> ops;  // [n_gates][3]: gate, src1, src2
    std::vector<int64_t> outputs;
    uint64_t *out = nullptr;
    // scheduling state (worker thread only)
    int64_t base = -1;           // first pool row
    std::vector<int64_t> row;    // node -> pool row
    std::vector<uint8_t> pending;
    std::vector<int64_t> users_off;
    std::vector<int32_t> users;
    std::deque<int32_t> ready;
    std::vector<double> ready_t;
    int64_t gates_done = 0;
    int64_t inflight = 0;        // gates of this job in launches not yet retired
    bool contiguous_out = false;
    // completion (guarded by the server mutex)
    double t_submit = 0, t_done = 0;
    bool finished = false;
    int rc = 0;
    std::string err;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    void *scratch = nullptr;
    void *desc_dev = nullptr, *desc_host = nullptr;  // idx1, idx2, oidx (i64); gates, key_of, status (i32)
    bool busy = false;
    double t_launch = 0;
    std::vector<std::pair<std::shared_ptr<Job>, int32_t>> entries;  // owning: a job outlives its launches
};

template <typename T>
double pct(std::vector<T> v, double p) {
    if (v.empty()) return 0.0;
    size_t k = (size_t)std::min<double>((double)v.size() - 1, p / 100.0 * (double)(v.size() - 1) + 0.5);
    std::nth_element(v.begin(), v.begin() + k, v.end());
    return (double)v[k];
}

}  // namespace

struct fheacc_server {
    fheacc_ctx *ctx = nullptr;
    fheacc_server_config cfg{};
    int64_t row_words = 0, rows = 0;
    uint64_t *pool = nullptr;
    void **key_tab_dev = nullptr;  // [kMaxClients] device bootstrap-key pointers
    std::vector<fheacc_keys *> clients;
    Slot slots[kMaxInflight];
    hipStream_t copy_stream = nullptr;
    std::thread worker;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    int64_t next_id = 1;
    std::deque<std::shared_ptr<Job>> queue;
    std::unordered_map<int64_t, std::shared_ptr<Job>> jobs;
    std::map<int64_t, int64_t> free_rows;  // start -> length
    int64_t used_rows = 0, peak_rows = 0;

    bool work = false;  // set by submit / destroy, cleared by the worker

    // worker-only (read under mu by the stats call)
    std::vector<std::shared_ptr<Job>> active;
    int64_t nready = 0;
    int ninflight = 0;
    double busy_since = 0;

    // statistics (guarded by mu)
    fheacc_server_stats st{};
    double t_reset = 0;
    std::vector<float> gate_ms, job_ms;

    int64_t alloc_rows(int64_t n) {
        for (auto it = free_rows.begin(); it != free_rows.end(); ++it) {
            if (it->second >= n) {
                const int64_t s = it->first, len = it->second;
                free_rows.erase(it);
                if (len > n) free_rows[s + n] = len - n;
                used_rows += n;
                peak_rows = std::max(peak_rows, used_rows);
                return s;
            }
        }
        return -1;
    }
    void free_range(int64_t s, int64_t n) {
        used_rows -= n;
        auto it = free_rows.emplace(s, n).first;
        auto nx = std::next(it);
        if (nx != free_rows.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_rows.erase(nx);
        }
        if (it != free_rows.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_rows.erase(it);
            }
        }
    }

    void run();
    bool admit_and_upload();
    bool poll();
    bool launch();
    void resolve(Job *j, int64_t node, double t);
    void finish(Job *j, int rc, std::string err);
    void release(Job *j);
    void retire_failed(Slot &s, const std::string &err);
};

// ------------------------------------------------------------------ worker

void fheacc_server::resolve(Job *j, int64_t node, double t) {
    for (int64_t u = j->users_off[node]; u < j->users_off[node + 1]; u++) {
        const int32_t g = j->users[u];
        if (--j->pending[g] == 0) {
            j->ready.push_back(g);
            j->ready_t[g] = t;
            nready++;
        }
    }
}

// Completes a job: outputs to host memory (rc == 0), waiters woken.  Its
// rows go back to the pool once none of its gates is in flight.
void fheacc_server::finish(Job *j, int rc, std::string err) {
    if (!rc) {
        const int64_t rb = row_words * 8;
        hipError_t e = hipSuccess;
        if (j->contiguous_out && j->n_out > 0) {
            e = hipMemcpyAsync(j->out, pool + j->row[j->outputs[0]] * row_words, j->n_out * rb,
                               hipMemcpyDeviceToHost, copy_stream);
        } else {
            for (int64_t t = 0; t < j->n_out && e == hipSuccess; t++)
                e = hipMemcpyAsync(j->out + t * row_words, pool + j->row[j->outputs[t]] * row_words, rb,
                                   hipMemcpyDeviceToHost, copy_stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(copy_stream);
        if (e != hipSuccess) {
            rc = FHEACC_ERR;
            err = std::string("output copy: ") + hipGetErrorString(e);
        }
    }
    nready -= (int64_t)j->ready.size();
    j->ready.clear();
    {
        std::lock_guard<std::mutex> lk(mu);
        j->finished = true;
        j->rc = rc;
        j->err = err;
        j->t_done = now_s();
        if (rc) st.jobs_failed++;
        else {
            st.jobs_done++;
            job_ms.push_back((float)((j->t_done - j->t_submit) * 1e3));
        }
        cv_done.notify_all();
    }
    release(j);
}

void fheacc_server::release(Job *j) {
    if (!j->finished || j->inflight > 0) return;
    std::lock_guard<std::mutex> lk(mu);
    if (j->base >= 0) free_range(j->base, j->n_in + j->n_gates);
    j->base = -1;
    for (size_t i = 0; i < active.size(); i++)
        if (active[i].get() == j) {
            active.erase(active.begin() + i);
            break;
        }
}

bool fheacc_server::admit_and_upload() {
    std::vector<Job *> admitted;
    {
        std::lock_guard<std::mutex> lk(mu);
        while (!queue.empty() && !stop) {
            Job *j = queue.front().get();
            const int64_t r = alloc_rows(j->n_in + j->n_gates);
            if (r < 0) break;
            j->base = r;
            active.push_back(queue.front());
            queue.pop_front();
            admitted.push_back(j);
        }
    }
    for (Job *j : admitted) {
        // node rows: inputs, then the non-output gates, then the outputs in order
        const int64_t nn = j->n_in + j->n_gates;
        j->row.assign(nn, -1);
        for (int64_t i = 0; i < j->n_in; i++) j->row[i] = j->base + i;
        bool contig = j->n_out > 0;
        std::vector<uint8_t> is_out(nn, 0);
        for (int64_t t = 0; t < j->n_out; t++) {
            const int64_t o = j->outputs[t];
            if (o < j->n_in || is_out[o]) contig = false;
            is_out[o] = 1;
        }
        j->contiguous_out = contig;
        int64_t next = j->base + j->n_in;
        for (int64_t g = 0; g < j->n_gates; g++)
            if (!contig || !is_out[j->n_in + g]) j->row[j->n_in + g] = next++;
        if (contig)
            for (int64_t t = 0; t < j->n_out; t++) j->row[j->outputs[t]] = next++;
        hipError_t e = hipMemcpyAsync(pool + j->base * row_words, j->in_words.data(),
                                      j->in_words.size() * 8, hipMemcpyHostToDevice, copy_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(copy_stream);
        std::vector<uint64_t>().swap(j->in_words);
        if (e != hipSuccess) {
            finish(j, FHEACC_ERR, std::string("input copy: ") + hipGetErrorString(e));
            continue;
        }
        const double t = now_s();
        for (int64_t i = 0; i < j->n_in; i++) resolve(j, i, t);
        if (j->n_gates == 0) finish(j, 0, "");
    }
    return !admitted.empty();
}

// A launch that failed, or whose completion reported an error: its jobs
// fail (nothing it wrote can be trusted) and its entries retire.
void fheacc_server::retire_failed(Slot &s, const std::string &err) {
    for (auto &en : s.entries) en.first->inflight--;
    for (auto &en : s.entries) {
        Job *j = en.first.get();
        if (!j->finished) finish(j, FHEACC_ERR, err);
        else release(j);
    }
    s.entries.clear();
}

bool fheacc_server::poll() {
    bool any = false;
    for (int si = 0; si < cfg.max_inflight; si++) {
        Slot &s = slots[si];
        if (!s.busy) continue;
        hipError_t e = hipEventQuery(s.done);
        if (e == hipErrorNotReady) continue;
        any = true;
        s.busy = false;
        ninflight--;
        const double t = now_s();
        if (ninflight == 0) {
            std::lock_guard<std::mutex> lk(mu);
            st.gpu_busy_s += t - busy_since;
        }
        if (e != hipSuccess) {
            retire_failed(s, std::string("gate launch: ") + hipGetErrorString(e));
            continue;
        }
        const int64_t B = (int64_t)s.entries.size();
        const int32_t *status = (const int32_t *)((char *)s.desc_host + B * 24 + B * 8);
        std::vector<float> lat;
        lat.reserve(B);
        std::vector<Job *> bad;
        for (int64_t i = 0; i < B; i++) {
            Job *j = s.entries[i].first.get();
            const int32_t g = s.entries[i].second;
            j->inflight--;
            if (j->finished) continue;
            if (status[i]) {
                bad.push_back(j);
                continue;
            }
            lat.push_back((float)((t - j->ready_t[g]) * 1e3));
            j->gates_done++;
            resolve(j, j->n_in + g, t);
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            st.bootstraps += B;
            st.batches++;
            st.max_batch_seen = std::max<int64_t>(st.max_batch_seen, B);
            gate_ms.insert(gate_ms.end(), lat.begin(), lat.end());
        }
        for (Job *j : bad)
            if (!j->finished)
                finish(j, FHEACC_ERR, "identical operands passed to a gate bootstrap (RLWE.cpp:727-730)");
        for (auto &en : s.entries) {
            Job *j = en.first.get();
            if (!j->finished && j->gates_done == j->n_gates) finish(j, 0, "");
            else release(j);
        }
        s.entries.clear();
    }
    return any;
}

bool fheacc_server::launch() {
    if (nready == 0 || ninflight >= cfg.max_inflight) return false;
    if (ninflight > 0 && nready < cfg.min_fill) return false;
    int si = 0;
    while (slots[si].busy) si++;
    Slot &s = slots[si];
    s.entries.clear();
    auto take = [&](const std::shared_ptr<Job> &jp, int64_t most) {
        Job *j = jp.get();
        for (int64_t t = 0; t < most && !j->ready.empty() && (int64_t)s.entries.size() < cfg.max_batch; t++) {
            s.entries.emplace_back(jp, j->ready.front());
            j->ready.pop_front();
            j->inflight++;
            nready--;
        }
    };
    if (cfg.fair_share) {
        int64_t jobs = 0;
        for (auto &jp : active) jobs += !jp->ready.empty();
        const int64_t share = std::max<int64_t>(1, cfg.max_batch / std::max<int64_t>(1, jobs));
        for (auto &jp : active) take(jp, share);
    }
    // then oldest job first
    for (auto &jp : active) {
        Job *j = jp.get();
        while (!j->ready.empty() && (int64_t)s.entries.size() < cfg.max_batch) {
            s.entries.emplace_back(jp, j->ready.front());
            j->ready.pop_front();
            j->inflight++;
            nready--;
        }
        if ((int64_t)s.entries.size() >= cfg.max_batch) break;
    }
    std::stable_sort(s.entries.begin(), s.entries.end(),
                     [](const std::pair<std::shared_ptr<Job>, int32_t> &a,
                        const std::pair<std::shared_ptr<Job>, int32_t> &b) {
                         return a.first->client < b.first->client;
                     });
    const int64_t B = (int64_t)s.entries.size();
    int64_t *idx1 = (int64_t *)s.desc_host, *idx2 = idx1 + B, *oidx = idx2 + B;
    int32_t *gates = (int32_t *)(oidx + B), *key_of = gates + B;
    std::vector<int64_t> seg_off;
    std::vector<const void *> seg_ksk;
    for (int64_t i = 0; i < B; i++) {
        Job *j = s.entries[i].first.get();
        const int32_t g = s.entries[i].second;
        const int32_t *op = &j->ops[3 * (size_t)g];
        gates[i] = op[0];
        idx1[i] = j->row[op[1]];
        idx2[i] = j->row[op[2]];
        oidx[i] = j->row[j->n_in + g];
        key_of[i] = j->client;
        if (i == 0 || j->client != s.entries[i - 1].first->client) {
            seg_off.push_back(i);
            seg_ksk.push_back(clients[j->client]->ksk);
        }
    }
    seg_off.push_back(B);
    char *dd = (char *)s.desc_dev;
    GateLaunch gl;
    gl.gates = (const int32_t *)(dd + B * 24);
    gl.c1 = pool;
    gl.c2 = pool;
    gl.out = pool;
    gl.B = B;
    gl.idx1 = (const int64_t *)dd;
    gl.idx2 = (const int64_t *)(dd + B * 8);
    gl.oidx = (const int64_t *)(dd + B * 16);
    gl.key_of = (const int32_t *)(dd + B * 24 + B * 4);
    gl.bsk_tab = (const void *const *)key_tab_dev;
    gl.ct_status = (int *)(dd + B * 24 + B * 8);
    gl.nseg = (int)seg_ksk.size();
    gl.seg_off = seg_off.data();
    gl.seg_ksk = seg_ksk.data();
    gl.scratch = s.scratch;
    gl.stream = s.stream;
    std::string err;
    int rc = 0;
    hipError_t e = hipMemcpyAsync(s.desc_dev, s.desc_host, B * 32, hipMemcpyHostToDevice, s.stream);
    if (e != hipSuccess) err = std::string("descriptor copy: ") + hipGetErrorString(e);
    if (err.empty()) {
        rc = launch_gate_path(ctx, nullptr, gl);
        if (rc) err = fheacc_last_error();
    }
    if (err.empty()) {
        e = hipMemcpyAsync((char *)s.desc_host + B * 32, gl.ct_status, B * 4, hipMemcpyDeviceToHost, s.stream);
        if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
        if (e != hipSuccess) err = std::string("status copy: ") + hipGetErrorString(e);
    }
    if (!err.empty()) {
        // nothing of this launch can be trusted: fail its jobs
        (void)hipStreamSynchronize(s.stream);
        retire_failed(s, err);
        return true;
    }
    s.busy = true;
    s.t_launch = now_s();
    if (ninflight++ == 0) busy_since = s.t_launch;
    return true;
}

void fheacc_server::run() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        bool stopping;
        {
            std::lock_guard<std::mutex> lk(mu);
            stopping = stop;
        }
        bool progressed = stopping ? false : admit_and_upload();
        progressed |= poll();
        if (!stopping) progressed |= launch();
        std::unique_lock<std::mutex> lk(mu);
        if (stop && ninflight == 0) break;
        if (progressed) continue;
        // in flight: poll completions every 200 us; idle: sleep until work
        const auto dt = ninflight > 0 ? std::chrono::microseconds(200) : std::chrono::microseconds(50000);
        cv_work.wait_for(lk, dt, [&] { return work || stop; });
        work = false;
    }
    // stopping: nothing is in flight; fail what did not finish
    std::vector<std::shared_ptr<Job>> left = active;
    for (auto &j : left)
        if (!j->finished) finish(j.get(), FHEACC_ERR, "server destroyed before the job finished");
    std::lock_guard<std::mutex> lk(mu);
    for (auto &q : queue) {
        q->finished = true;
        q->rc = FHEACC_ERR;
        q->err = "server destroyed before the job was admitted";
        st.jobs_failed++;
    }
    queue.clear();
    cv_done.notify_all();
}

// ------------------------------------------------------------------ C ABI

static void server_release(fheacc_server *s) {
    for (Slot &sl : s->slots) {
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.scratch) (void)hipFree(sl.scratch);
        if (sl.desc_dev) (void)hipFree(sl.desc_dev);
        if (sl.desc_host) (void)hipHostFree(sl.desc_host);
    }
    if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
    if (s->pool) (void)hipFree(s->pool);
    if (s->key_tab_dev) (void)hipFree(s->key_tab_dev);
    for (fheacc_keys *k : s->clients) keys_release(k);
    if (s->ctx) ctx_release(s->ctx);
    delete s;
}

int fheacc_server_create(fheacc_ctx *ctx, const fheacc_server_config *cfg, fheacc_server **out) {
    if (!ctx || !cfg || !out) return fail("null argument");
    *out = nullptr;
    if (cfg->max_batch < 1 || cfg->max_batch > (1 << 20)) return fail("max_batch must be in [1, 2^20]");
    if (cfg->max_inflight < 1 || cfg->max_inflight > kMaxInflight) return fail("max_inflight must be in [1, 4]");
    if (cfg->min_fill < 0) return fail("min_fill must be >= 0");
    if (cfg->fair_share != 0 && cfg->fair_share != 1) return fail("fair_share must be 0 or 1");
    const HostParams &p = ctx->p;
    const int64_t rw = p.n + 1;
    const int64_t rows = (int64_t)(cfg->hbm_budget / (uint64_t)(rw * 8));
    if (rows < 3) return fail("hbm_budget holds fewer than 3 ciphertexts");
    HIP_TRY(hipSetDevice(ctx->device));
    fheacc_server *s = new fheacc_server();
    ctx_retain(ctx);
    s->ctx = ctx;
    s->cfg = *cfg;
    s->row_words = rw;
    s->rows = rows;
    auto bail = [&](hipError_t e, const char *what) {
        server_release(s);
        return fail(std::string(what) + ": " + hipGetErrorString(e),
                    e == hipErrorOutOfMemory ? FHEACC_ENOMEM : FHEACC_ERR);
    };
    hipError_t e;
    if ((e = hipMalloc(&s->pool, (size_t)rows * rw * 8)) != hipSuccess) return bail(e, "pool");
    if ((e = hipMalloc(&s->key_tab_dev, kMaxClients * sizeof(void *))) != hipSuccess) return bail(e, "key table");
    if ((e = hipMemset(s->key_tab_dev, 0, kMaxClients * sizeof(void *))) != hipSuccess) return bail(e, "key table");
    if ((e = hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "stream");
    const size_t scratch = gate_scratch_bytes(ctx, cfg->max_batch);
    const size_t desc = (size_t)cfg->max_batch * 36;
    for (int i = 0; i < cfg->max_inflight; i++) {
        Slot &sl = s->slots[i];
        if ((e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "stream");
        if ((e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess) return bail(e, "event");
        if ((e = hipMalloc(&sl.scratch, scratch)) != hipSuccess) return bail(e, "scratch");
        if ((e = hipMalloc(&sl.desc_dev, desc)) != hipSuccess) return bail(e, "descriptors");
        if ((e = hipHostMalloc(&sl.desc_host, desc, hipHostMallocDefault)) != hipSuccess)
            return bail(e, "descriptors");
    }
    s->free_rows[0] = rows;
    s->t_reset = now_s();
    s->st.pool_bytes = (uint64_t)rows * rw * 8;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return 0;
}

int fheacc_server_destroy(fheacc_server *s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->work = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    // waiters that have not collected their jobs were woken with a failure;
    // the handle must not be used after this call
    server_release(s);
    return 0;
}

int fheacc_server_add_client(fheacc_server *s, fheacc_keys *keys, int32_t *client) {
    if (!s || !keys || !client) return fail("null argument");
    const HostParams &a = keys->ctx->p, &b = s->ctx->p;
    if (a.n != b.n || a.q != b.q || a.N != b.N || a.Q != b.Q || a.BGbits != b.BGbits || a.Br != b.Br ||
        a.Bks != b.Bks)
        return fail("keys belong to a different parameter set");
    if (keys->ctx->device != s->ctx->device) return fail("keys are resident on another device");
    if (!keys->bsk || !keys->ksk) return fail("bootstrap and key-switch keys must be resident");
    std::lock_guard<std::mutex> lk(s->mu);
    if ((int)s->clients.size() >= kMaxClients) return fail("too many clients (64)");
    const int32_t id = (int32_t)s->clients.size();
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(hipMemcpy(s->key_tab_dev + id, &keys->bsk, sizeof(void *), hipMemcpyHostToDevice));
    keys_retain(keys);
    s->clients.push_back(keys);
    *client = id;
    return 0;
}

int fheacc_server_submit(fheacc_server *s, int32_t client, const uint64_t *inputs, int64_t n_in,
                         const int32_t *ops, int64_t n_gates, const int64_t *outputs, int64_t n_out,
                         uint64_t *out, int64_t *job) {
    if (!s || !job || (n_in > 0 && !inputs) || (n_gates > 0 && !ops) || (n_out > 0 && (!outputs || !out)))
        return fail("null argument");
    *job = 0;
    const double t_submit = now_s();
    if (n_in < 0 || n_gates < 0 || n_out < 0 || n_gates > 0x7fffffff) return fail("negative or oversized count");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (client < 0 || client >= (int32_t)s->clients.size()) return fail("unknown client");
        if (s->stop) return fail("server is stopping");
    }
    const int64_t nn = n_in + n_gates;
    if (nn > s->rows) return fail("job needs more ciphertext rows than the whole HBM budget holds");
    const HostParams &p = s->ctx->p;
    auto j = std::make_shared<Job>();
    j->client = client;
    j->n_in = n_in;
    j->n_gates = n_gates;
    j->n_out = n_out;
    j->out = out;
    j->ops.assign(ops, ops + 3 * n_gates);
    j->outputs.assign(outputs, outputs + n_out);
    j->pending.assign(n_gates, 2);
    j->ready_t.assign(n_gates, 0.0);
    std::vector<int64_t> deg(nn + 1, 0);
    for (int64_t g = 0; g < n_gates; g++) {
        const int32_t gt = j->ops[3 * g], a = j->ops[3 * g + 1], b = j->ops[3 * g + 2];
        if (gt < 0 || gt > 5) return fail("gate " + std::to_string(g) + ": gate type out of range");
        if (a < 0 || b < 0 || a >= n_in + g || b >= n_in + g)
            return fail("gate " + std::to_string(g) + ": operands must be earlier nodes");
        if (a == b) return fail("gate " + std::to_string(g) + ": identical operands (RLWE.cpp:727-730)");
        deg[a + 1]++;
        deg[b + 1]++;
    }
    for (int64_t t = 0; t < n_out; t++)
        if (j->outputs[t] < 0 || j->outputs[t] >= nn) return fail("output node out of range");
    for (int64_t i = 0; i < nn; i++) deg[i + 1] += deg[i];
    j->users_off = deg;
    j->users.assign(deg[nn], 0);
    std::vector<int64_t> fill(deg.begin(), deg.end() - 1);
    for (int64_t g = 0; g < n_gates; g++) {
        j->users[fill[j->ops[3 * g + 1]]++] = (int32_t)g;
        j->users[fill[j->ops[3 * g + 2]]++] = (int32_t)g;
    }
    // LWE words mod q; the gate pre-combination reduces them (LWE.cpp:203-222),
    // so a previous job's outputs (which may equal q, LWE.cpp:255-260) chain
    j->in_words.assign(inputs, inputs + n_in * (int64_t)(p.n + 1));
    j->t_submit = t_submit;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j->id = s->next_id++;
        s->jobs[j->id] = j;
        s->queue.push_back(j);
        s->st.jobs_submitted++;
        s->work = true;
        *job = j->id;
    }
    s->cv_work.notify_all();
    return 0;
}

int fheacc_server_wait(fheacc_server *s, int64_t job, double *latency_ms) {
    if (!s) return fail("null server");
    std::shared_ptr<Job> j;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job);
        if (it == s->jobs.end()) return fail("unknown job (or already collected)");
        j = it->second;
        s->cv_done.wait(lk, [&] { return j->finished; });
        s->jobs.erase(it);
    }
    if (latency_ms) *latency_ms = (j->t_done - j->t_submit) * 1e3;
    if (j->rc) return fail(j->err, j->rc);
    return 0;
}

int fheacc_server_get_stats(fheacc_server *s, fheacc_server_stats *out) {
    if (!s || !out) return fail("null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    const double t = now_s();
    out->wall_s = t - s->t_reset;
    if (s->ninflight > 0) out->gpu_busy_s += t - s->busy_since;
    out->jobs_queued = (int64_t)s->queue.size();
    out->jobs_active = (int64_t)s->active.size();
    out->gate_ms_p50 = pct(s->gate_ms, 50);
    out->gate_ms_p99 = pct(s->gate_ms, 99);
    out->job_ms_p50 = pct(s->job_ms, 50);
    out->job_ms_p99 = pct(s->job_ms, 99);
    out->pool_bytes = (uint64_t)s->rows * s->row_words * 8;
    out->peak_bytes = (uint64_t)s->peak_rows * s->row_words * 8;
    return 0;
}

int fheacc_server_reset_stats(fheacc_server *s) {
    if (!s) return fail("null server");
    std::lock_guard<std::mutex> lk(s->mu);
    const int64_t sub = s->st.jobs_submitted - s->st.jobs_done - s->st.jobs_failed;
    s->st = fheacc_server_stats{};
    s->st.jobs_submitted = sub;  // jobs still in the system
    s->gate_ms.clear();
    s->job_ms.clear();
    s->peak_rows = s->used_rows;
    s->t_reset = now_s();
    if (s->ninflight > 0) s->busy_since = s->t_reset;
    return 0;
}
