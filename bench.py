"""Benchmark: batched gate bootstraps on one MI355X.

Workload (one step): a batch of 4096 seeded NAND gate bootstraps with the
benchmark parameter set -- LWE n=500, q=1024, binary LWE secret; RLWE N=1024,
ternary secret, Q = 0x7FFF801 (27-bit, = 1 mod 2048); gadget base 2^7 (d=4);
blind-rotation radix Br=32 (digitR=2); key-switch base 2^5 (digitKS=6);
sigma 3.19 -- i.e. gate pre-combination, accumulator init, blind rotation,
LWE key switch and modulus switch for every pair, with inputs and keys
already resident in HBM.  Keys come from the library's seeded device key
generator.

Also measured: the same gate bootstrap one ciphertext per launch (the plain
form), its p50/p99 latency, per-kernel HIP-event times, the serving loop under
Poisson load (8 clients with their own resident keys, k-bit set-equality
jobs, p50/p99 gate and request latency), and a CPU baseline on a bounded
sample (the reference's own eval_bootstrap compiled from its sources when
oracle/_ref was built, else the oracle's C restatement).

Prints one JSON line.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bootstraps/sec (N=1024, batch 4096) vs 1-per-launch; p50/p99 gate latency"
PARAMS = dict(n=500, q=1024, N=1024, Q=0, nbit=27, BGbits=7, Br=32, Bks=32, Bks_R_bits=9)
SEED = 20260101
HBM_PEAK_GBS = 8000.0
# kernel sources whose counters the committed profiles hold: a profile is used
# only while the digest it records matches these files
PROFILED_SOURCES = ("kernels_impl.hpp", "ntt_lds.hpp", "modarith.hpp")


def kernel_sources_digest():
    import hashlib
    h = hashlib.sha1()
    for f in PROFILED_SOURCES:
        with open(os.path.join(ROOT, "3rd_gen_fhe_acc_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def load_profile(name):
    """A committed counter summary and whether it describes the current kernels."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, {"profile": "profiles/" + name, "fresh": False, "why": "absent"}
    with open(path) as f:
        prof = json.load(f)
    cur = kernel_sources_digest()
    fresh = prof.get("kernel_sources_sha1") == cur
    return prof, {"profile": "profiles/" + name, "kernel_sources_sha1": prof.get("kernel_sources_sha1"),
                  "current_sources_sha1": cur, "fresh": fresh,
                  "why": "" if fresh else "kernel sources changed since the counters were collected"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--single", type=int, default=24, help="one-per-launch gate bootstraps to time")
    ap.add_argument("--cpu-sample", type=int, default=20, help="gates in the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tuning", default="auto",
                    help="'auto' (run the autotuner) or THREADS,POLYS_PER_BATCH for the blind-rotation kernel")
    ap.add_argument("--schedule", default="auto",
                    help="blind-rotation schedule 'persist,window' (fheacc_ctx_set_schedule), or 'auto' to keep "
                         "the autotuner's choice (with --tuning auto) / one workgroup per ciphertext")
    ap.add_argument("--no-microbench", action="store_true", help="skip the NTT / external-product microbench")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="serving-loop arrival window (0: skip)")
    ap.add_argument("--serve-qps", type=float, default=6.0, help="serving-loop Poisson arrival rate")
    ap.add_argument("--serve-curve", default="3,6,10",
                    help="offered loads (req/s) of the latency curve after the serving run ('' to skip)")
    ap.add_argument("--serve-curve-requests", type=int, default=30, help="requests per curve point")
    ap.add_argument("--serve-wave-cap", type=int, default=1024,
                    help="the curve's second launch cap: one resident wave of gates (4 per CU x 256 CUs)")
    ap.add_argument("--serve-max-batch", type=int, default=4096)
    ap.add_argument("--serve-inflight", type=int, default=1,
                    help="serving launches in flight (one gives lower gate latency at this load)")
    ap.add_argument("--serve-min-fill", type=int, default=0)
    ap.add_argument("--serve-fifo", action="store_true",
                    help="fill serving launches oldest job first (default: fair share per job, then oldest first)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-parameter-set batch")
    ap.add_argument("--no-leveled", action="store_true", help="skip the RLWE expansion / blind-rotate microbench")
    ap.add_argument("--no-fpga-order", action="store_true",
                    help="skip timing the batch in the accelerator's step orders")
    ap.add_argument("--no-copy-ceiling", action="store_true",
                    help="skip measuring the device-to-device copy rate (the practical HBM ceiling)")
    ap.add_argument("--no-app256", action="store_true",
                    help="skip the reference's 256-bit application set (29-bit Q, N=2048)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device-to-device (reference dispatch for HBM byte counters)")
    ap.add_argument("--only-set", default="", choices=["", "A", "B", "C"],
                    help="time one set of the mixed batch alone (its dense launch at its share of 8192; "
                         "--tuning / --schedule apply to it) and print that set's JSON line instead")
    return ap.parse_args()


def microbench(fhe, torch, B, reps=5):
    """Batched forward+inverse negacyclic NTT plus NTT-domain RLWE x RGSW
    external product: B RLWE ciphertexts, N=1024, Q=0x3FFFFFFFFED001 (54-bit,
    = 1 mod 2048), coefficients uniform in [0, Q) from a seed, one shared RGSW
    with gadget base 2^9 (d=6).  Times from HIP events per kernel."""
    ctx = fhe.Context(n=8, q=2048, N=1024, Q=0x3FFFFFFFFED001, BGbits=9)
    rng = np.random.default_rng(SEED + 1)
    Q = ctx.Q
    polys = torch.from_numpy(rng.integers(0, Q, (B, 2, 1024), dtype=np.uint64).view(np.int64)).cuda()
    rgsw = torch.from_numpy(rng.integers(0, Q, ctx.digitG * 4 * 1024, dtype=np.uint64).view(np.int64)).cuda()
    out = torch.zeros_like(polys)
    work = polys.clone()
    ctx.ntt_dev(work)
    ctx.ntt_dev(work, inverse=True)
    ctx.external_product_dev(polys, rgsw, out)
    ctx.sync()
    roundtrip_ok = bool(torch.equal(work, polys))
    ctx.profile(True)
    for _ in range(reps):
        ctx.ntt_dev(work)
        ctx.ntt_dev(work, inverse=True)
        ctx.external_product_dev(polys, rgsw, out)
    kms = ctx.kernel_ms()
    ctx.profile(False)
    ntt_ms = kms["ntt"][0] / max(kms["ntt"][1], 1)
    ep_ms = kms["external_product"][0] / max(kms["external_product"][1], 1)
    d, N = ctx.digitG, 1024
    ep_modmul = 2 * d * (N // 2) * 10 + 2 * (N // 2) * 10 + 2 * N + 4 * d * N
    return {
        "workload": "B=%d RLWE, N=1024, Q=0x3FFFFFFFFED001, BG=2^9 d=6, shared RGSW" % B,
        "ntt_ms_per_launch": round(ntt_ms, 4), "ntt_polys_per_launch": 2 * B,
        "ntt_polys_per_s": round(2 * B / (ntt_ms * 1e-3), 1),
        "external_product_ms_per_launch": round(ep_ms, 4),
        "external_products_per_s": round(B / (ep_ms * 1e-3), 1),
        "external_product_modmul_per_s_T": round(B * ep_modmul / (ep_ms * 1e-3) / 1e12, 4),
        "forward_inverse_roundtrip_ok": roundtrip_ok,
    }


def serving_clients(ctx):
    """8 clients with their own device-resident keys (benchmark parameters)."""
    clients = []
    t0 = time.perf_counter()
    for c in range(8):
        s = ctx.lwe_secret(SEED + 100 + c, binary=True)
        z = ctx.rlwe_secret(SEED + 100 + c)
        keys = ctx.keys()
        keys.generate(SEED + 200 + c, s, z)
        clients.append((s, keys))
    ctx.sync()
    return clients, time.perf_counter() - t0


def serving(fhe, ctx, clients, qps, seconds, max_batch, inflight, min_fill=0, fair=True, seed=SEED + 4,
            alone_check=True, n_requests=None):
    """Serving loop: Poisson arrivals at `qps` for `seconds` of k-bit
    set-equality jobs (k from {8, 16, 32}: XNOR of each item bit with the
    query bit, then an AND tree per item), each for one of the 8 clients.
    Item counts are log-uniform in [1, 1024], values uniform 32-bit from a
    seed.  Admission by HBM budget, continuous batching across jobs and
    clients (fheacc_server_*)."""
    circuits = importlib.import_module("3rd_gen_fhe_acc_amd.circuits")
    rng = np.random.default_rng(seed)
    arrivals = []
    if n_requests:
        # a fixed request list: unit-rate gaps scaled by 1/qps, so that every
        # load replays the same requests and the offered load scales with qps
        gaps = np.random.default_rng(seed + 1).exponential(1.0, n_requests)
        arrivals = list(np.cumsum(gaps) / qps)
        seconds = float(arrivals[-1])
    else:
        t = rng.exponential(1.0 / qps)
        while t < seconds:
            arrivals.append(t)
            t += rng.exponential(1.0 / qps)
    reqs = []
    cache = {}
    t0 = time.perf_counter()
    for i, ta in enumerate(arrivals):
        c = int(rng.integers(0, 8))
        k = int(rng.choice([8, 16, 32]))
        m = int(min(1024, max(1, np.floor(np.exp(rng.uniform(0.0, np.log(1025.0)))))))
        items = rng.integers(0, 2 ** 32, m, dtype=np.uint64)
        query = int(rng.integers(0, 2 ** 32, dtype=np.uint64))
        if (m, k) not in cache:
            cache[(m, k)] = circuits.equality_circuit(m, k)
        n_in, ops, outs = cache[(m, k)]
        bits = circuits.equality_inputs(items, query, k)
        inputs = ctx.lwe_encrypt(clients[c][0], bits, SEED + 300, i << 20)
        mask = (1 << k) - 1
        want = np.array([(int(v) & mask) == (query & mask) for v in items], dtype=np.uint8)
        reqs.append(dict(t=ta, client=c, k=k, m=m, inputs=inputs, ops=ops, outs=outs, want=want))
    encrypt_s = time.perf_counter() - t0
    srv = fhe.Server(ctx, hbm_budget=8 << 30, max_batch=max_batch, max_inflight=inflight, min_fill=min_fill,
                     fair_share=fair)
    ids = [srv.add_client(k) for _, k in clients]
    # warm the path once, then measure from a clean slate
    r0 = min(reqs, key=lambda r: len(r["ops"]))
    warm, _ = srv.wait(srv.submit(ids[r0["client"]], r0["inputs"], r0["ops"], r0["outs"]))
    srv.reset_stats()
    jids = []
    t0 = time.perf_counter()
    for r in reqs:
        dt = t0 + r["t"] - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        jids.append(srv.submit(ids[r["client"]], r["inputs"], r["ops"], r["outs"]))
    outs = [srv.wait(j)[0] for j in jids]
    elapsed = time.perf_counter() - t0
    st = srv.stats()
    ok = all(np.array_equal(ctx.lwe_decrypt(clients[r["client"]][0], o), r["want"]) for r, o in zip(reqs, outs))
    # the same requests alone, on an idle server: bit-identical words
    alone_eq = bool(np.array_equal(warm, outs[reqs.index(r0)]))
    if alone_check:
        for i in sorted(range(len(reqs)), key=lambda i: len(reqs[i]["ops"]))[:3]:
            r = reqs[i]
            o, _ = srv.wait(srv.submit(ids[r["client"]], r["inputs"], r["ops"], r["outs"]))
            alone_eq &= bool(np.array_equal(o, outs[i]))
    srv.close()
    gates = sum(len(r["ops"]) for r in reqs)
    return {
        "workload": "Poisson %.1f req/s for %.0f s: k-bit set equality (k in {8,16,32}, XNOR + AND tree), "
                    "items log-uniform in [1,1024], 8 clients with resident keys, benchmark parameters"
                    % (qps, seconds),
        "requests": len(reqs), "bootstraps": int(st["bootstraps"]),
        "offered_bootstraps_per_s": round(gates / seconds, 1),
        "bootstraps_per_s": round(st["bootstraps"] / elapsed, 1),
        "gate_latency_ms_p50": round(st["gate_ms_p50"], 3), "gate_latency_ms_p99": round(st["gate_ms_p99"], 3),
        "request_latency_ms_p50": round(st["job_ms_p50"], 3),
        "request_latency_ms_p99": round(st["job_ms_p99"], 3),
        "elapsed_s": round(elapsed, 3), "gpu_busy_s": round(st["gpu_busy_s"], 3),
        "launches": int(st["batches"]), "mean_launch_size": round(st["bootstraps"] / max(st["batches"], 1), 1),
        "max_launch_size": int(st["max_batch_seen"]),
        "hbm_budget_bytes": int(st["pool_bytes"]), "peak_admitted_bytes": int(st["peak_bytes"]),
        "config": {"max_batch": max_batch, "max_inflight": inflight, "min_fill": min_fill, "fair_share": fair},
        "decrypt_ok": bool(ok), "alone_equals_mix": bool(alone_eq), "encrypt_s": round(encrypt_s, 3),
    }


def serving_curve(fhe, ctx, clients, args):
    """Latency against offered load: the same --serve-curve-requests requests
    arriving at each rate of --serve-curve (Poisson gaps scaled by 1/rate),
    under the default launch cap and under a cap of one resident wave of
    gates."""
    runs = []
    for qps in (float(x) for x in args.serve_curve.split(",")):
        for cap in (args.serve_max_batch, args.serve_wave_cap):
            r = serving(fhe, ctx, clients, qps, 0.0, cap, args.serve_inflight, seed=SEED + 40, alone_check=False,
                        n_requests=args.serve_curve_requests)
            runs.append({k: r[k] for k in ("requests", "bootstraps", "offered_bootstraps_per_s", "bootstraps_per_s",
                                           "gate_latency_ms_p50", "gate_latency_ms_p99", "request_latency_ms_p50",
                                           "request_latency_ms_p99", "gpu_busy_s", "elapsed_s",
                                           "mean_launch_size", "decrypt_ok")} | {"qps": qps, "max_batch": cap})
    return {"requests_per_point": args.serve_curve_requests, "runs": runs}


MIXED = {
    "A": dict(PARAMS),
    "B": dict(n=630, q=1024, N=2048, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=32, Bks=32),
    "C": dict(n=500, q=1024, N=512, Q=0, nbit=27, BGbits=6, Br=32, Bks=32),
}
MIXED_SHARE = {"A": 0.6, "B": 0.3, "C": 0.1}


def mixed(fhe, torch, main_ctx, main_keys, total=8192, reps=3):
    """Uneven mixed batch in one launch: 60 % set A (the benchmark set), 30 %
    set B (N=2048, 54-bit Q, d=3, n=630), 10 % set C (N=512, 27-bit Q, d=5),
    8192 seeded random-bit ciphertexts uniformly interleaved in one buffer
    (row stride 631 words), captured as one HIP graph whose per-set kernels
    gather their rows on the device, each set at its autotuned launch shape."""
    rng = np.random.default_rng(SEED + 5)
    names = list(MIXED)
    counts = [int(round(total * MIXED_SHARE[k])) for k in names]
    counts[0] += total - sum(counts)
    lab = np.concatenate([np.full(c, i) for i, c in enumerate(counts)]).astype(np.int32)
    rng.shuffle(lab)
    rw = max(v["n"] for v in MIXED.values()) + 1
    c1_all = np.zeros((total, rw), dtype=np.uint64)
    c2_all = np.zeros_like(c1_all)
    g_all = np.zeros(total, dtype=np.int32)
    sets = []
    t0 = time.perf_counter()
    for i, nm in enumerate(names):
        rows = np.flatnonzero(lab == i)
        B = rows.size
        if nm == "A":
            ctx, keys = main_ctx, main_keys
            s = ctx.lwe_secret(SEED, binary=True)
        else:
            ctx = fhe.Context(**MIXED[nm])
            s, z = ctx.lwe_secret(SEED + 10 + i, binary=True), ctx.rlwe_secret(SEED + 10 + i)
            keys = ctx.keys()
            keys.generate(SEED + 20 + i, s, z)
            # tuned at the set's own batch: launch shapes that win on a full
            # chip lose on a partial last wave and the other way round
            ctx.autotune(keys, B)
        m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
        gates = rng.integers(0, 6, B).astype(np.int32)
        c1, c2 = ctx.lwe_encrypt(s, m1, SEED + 30 + i, 0), ctx.lwe_encrypt(s, m2, SEED + 30 + i, B)
        n1 = ctx.n + 1
        c1_all[rows, :n1], c2_all[rows, :n1], g_all[rows] = c1, c2, gates
        truth = {0: m1 | m2, 1: m1 & m2, 2: 1 - (m1 | m2), 3: 1 - (m1 & m2), 4: m1 ^ m2, 5: 1 - (m1 ^ m2)}
        exp = np.array([truth[int(g)][t] for t, g in enumerate(gates)])
        sets.append(dict(name=nm, ctx=ctx, keys=keys, s=s, B=B, rows=rows, n1=n1,
                         dg=torch.from_numpy(gates).cuda(), d1=torch.from_numpy(c1.view(np.int64)).cuda(),
                         d2=torch.from_numpy(c2.view(np.int64)).cuda(), exp=exp))
    setup_s = time.perf_counter() - t0
    dg = torch.from_numpy(g_all).cuda()
    d1 = torch.from_numpy(c1_all.view(np.int64)).cuda()
    d2 = torch.from_numpy(c2_all.view(np.int64)).cuda()
    dout = torch.zeros_like(d1)
    g = fhe.InterleavedGraph([st["ctx"] for st in sets], [st["keys"] for st in sets], lab, dg, d1, d2, dout)
    g.launch()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        g.launch()
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    got = dout.cpu().numpy().view(np.uint64)
    ok = True
    for st in sets:
        st["ctx"].sync()
        ok &= bool(np.array_equal(st["ctx"].lwe_decrypt(st["s"], got[st["rows"], :st["n1"]]), st["exp"]))
    # each set alone (its dense launch), back to back: what one launch saves
    per_set = {}
    dense_equal = True
    for st in sets:
        sout = torch.zeros_like(st["d1"])
        st["ctx"].bootstrap_dev(st["keys"], st["dg"], st["d1"], st["d2"], sout)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            st["ctx"].bootstrap_dev(st["keys"], st["dg"], st["d1"], st["d2"], sout)
        e1.record()
        torch.cuda.synchronize()
        per_set[st["name"]] = {"B": int(st["B"]), "alone_ms": round(e0.elapsed_time(e1) / reps, 3),
                               "tuning": st["ctx"].get_tuning(), "schedule": st["ctx"].get_schedule()}
        dense_equal &= bool(np.array_equal(sout.cpu().numpy().view(np.uint64), got[st["rows"], :st["n1"]]))
    g.close()
    return {
        "workload": "8192 ciphertexts uniformly interleaved in one buffer: 60% N=1024 27-bit d=4 n=500, 30% "
                    "N=2048 54-bit d=3 n=630, 10% N=512 27-bit d=5 n=500; one graph launch, each set's kernels "
                    "gathering their rows on the device, sets back to back",
        "ms_per_launch": round(ms, 3), "bootstraps_per_s": round(total / (ms * 1e-3), 1),
        "sets": per_set, "sum_alone_ms": round(sum(v["alone_ms"] for v in per_set.values()), 3),
        "decrypt_ok": bool(ok), "equals_per_set_launch": bool(dense_equal), "setup_s": round(setup_s, 2),
    }


def one_set(fhe, torch, args, total=8192):
    """One set of the mixed batch on its own: its share of the 8192 seeded
    random-bit ciphertexts over all six gates, as one dense gate-path launch,
    timed over --steps launches after --warmup, with per-kernel HIP events."""
    nm = args.only_set
    B = int(round(total * MIXED_SHARE[nm]))
    ctx = fhe.Context(**MIXED[nm])
    i = list(MIXED).index(nm)
    s, z = ctx.lwe_secret(SEED + 10 + i, binary=True), ctx.rlwe_secret(SEED + 10 + i)
    keys = ctx.keys()
    keys.generate(SEED + 20 + i, s, z)
    if args.tuning == "auto":
        tune = ctx.autotune(keys, B)
    else:
        th, pb = (int(x) for x in args.tuning.split(","))
        ctx.set_tuning(th, pb)
        tune = {"chosen": {"threads": th, "polys_per_batch": pb}, "source": "--tuning"}
    if args.schedule != "auto":
        ctx.set_schedule(*(int(x) for x in args.schedule.split(",")))
    rng = np.random.default_rng(SEED + 40 + i)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    gates = rng.integers(0, 6, B).astype(np.int32)
    c1, c2 = ctx.lwe_encrypt(s, m1, SEED + 30 + i, 0), ctx.lwe_encrypt(s, m2, SEED + 30 + i, B)
    truth = {0: m1 | m2, 1: m1 & m2, 2: 1 - (m1 | m2), 3: 1 - (m1 & m2), 4: m1 ^ m2, 5: 1 - (m1 ^ m2)}
    exp = np.array([truth[int(g)][t] for t, g in enumerate(gates)])
    d1 = torch.from_numpy(c1.view(np.int64)).cuda()
    d2 = torch.from_numpy(c2.view(np.int64)).cuda()
    dg = torch.from_numpy(gates).cuda()
    out = torch.zeros_like(d1)
    for _ in range(max(args.warmup, 1)):
        ctx.bootstrap_dev(keys, dg, d1, d2, out)
    ctx.sync()
    ctx.profile(True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        ctx.bootstrap_dev(keys, dg, d1, d2, out)
    e1.record()
    torch.cuda.synchronize()
    ctx.sync()
    kms = ctx.kernel_ms()
    ctx.profile(False)
    ms = e0.elapsed_time(e1) / args.steps
    ok = bool(np.array_equal(ctx.lwe_decrypt(s, out.cpu().numpy().view(np.uint64)), exp))
    info = ctx.info
    return {
        "workload": "mixed-batch set %s alone: %s, %d seeded random-bit ciphertexts over all six gates, dense "
                    "eager launches" % (nm, json.dumps(MIXED[nm]), B),
        "B": B, "ms_per_batch": round(ms, 3), "bootstraps_per_s": round(B / (ms * 1e-3), 1),
        "kernels_ms_avg": {k: round(v[0] / max(v[1], 1), 4) for k, v in kms.items() if v[1]},
        "bsk_device_bytes": info["bsk_device_bytes"], "decrypt_ok": ok,
        "tuning": tune, "schedule": ctx.get_schedule(), "steps": args.steps, "warmup": args.warmup,
    }


def leveled(fhe, torch, B=8, reps=3):
    """RLWE expansion (expand_RLWE) at the reference's expansion app size --
    N=2048, Q=0x3FFFFFFFFED001, Bks_R=2^9 (digitKS_R=6) -- B inputs per call,
    N-1 substitution + key switch steps each; the leveled blind_rotate
    (11 CMUXes, BG=2^9 d=6) over 256 RLWEs; homo_expand of one packed query
    and CMUX-tree selection at the same size.  Seeded random words and keys."""
    kw = dict(n=512, q=512, N=2048, Q=0x3FFFFFFFFED001, nbit=54, BGbits=9, Br=23, Bks=25, Bks_R_bits=9)
    ctx = fhe.Context(**kw)
    z = ctx.rlwe_secret(SEED + 6)
    keys = ctx.keys()
    keys.generate_subs(SEED + 7, z)
    rng = np.random.default_rng(SEED + 6)
    N = 2048
    x = torch.from_numpy(rng.integers(0, ctx.Q, (B, 2, N), dtype=np.uint64).view(np.int64)).cuda()
    out = torch.zeros((B, N, 2, N), dtype=torch.int64, device="cuda")
    ctx.expand_rlwe_dev(keys, x, out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ctx.expand_rlwe_dev(keys, x, out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del out
    R = 256
    rl = torch.from_numpy(rng.integers(0, ctx.Q, (R, 2, N), dtype=np.uint64).view(np.int64)).cuda()
    rg = torch.from_numpy(rng.integers(0, ctx.Q, (11, ctx.digitG * 4 * N), dtype=np.uint64).view(np.int64)).cuda()
    ro = torch.zeros_like(rl)
    ctx.blind_rotate_rlwe_dev(rl, True, rg, ro)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        ctx.blind_rotate_rlwe_dev(rl, True, rg, ro)
    e1.record()
    torch.cuda.synchronize()
    ms_br = e0.elapsed_time(e1) / reps
    del rl, ro
    # homo_expand (app.cpp:78-97) of one packed query: digitG expansions, then
    # N * digitG external products by RGSW(-s) into N RGSWs (805 MB of output)
    d = ctx.digitG
    packed = torch.from_numpy(rng.integers(0, ctx.Q, (1, d, 2, N), dtype=np.uint64).view(np.int64)).cuda()
    enc_sec = torch.from_numpy(rng.integers(0, ctx.Q, d * 4 * N, dtype=np.uint64).view(np.int64)).cuda()
    he = torch.zeros((1, N, d * 4 * N), dtype=torch.int64, device="cuda")
    ctx.homo_expand_dev(keys, packed, enc_sec, he)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        ctx.homo_expand_dev(keys, packed, enc_sec, he)
    e1.record()
    torch.cuda.synchronize()
    ms_he = e0.elapsed_time(e1) / reps
    del he
    # tree selection (test.cpp:1576-1582): T CMUX trees over 2^7 leaves, the
    # reference app's 18 - log2(N) levels
    T, L = 64, 7
    leaves = torch.from_numpy(rng.integers(0, ctx.Q, (T, 1 << L, 2, N), dtype=np.uint64).view(np.int64)).cuda()
    rgt = torch.from_numpy(rng.integers(0, ctx.Q, (L, d * 4 * N), dtype=np.uint64).view(np.int64)).cuda()
    to = torch.zeros((T, 2, N), dtype=torch.int64, device="cuda")
    ctx.cmux_tree_dev(leaves, rgt, to, L)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        ctx.cmux_tree_dev(leaves, rgt, to, L)
    e1.record()
    torch.cuda.synchronize()
    ms_tree = e0.elapsed_time(e1) / reps
    return {
        "workload": "expand_RLWE, N=2048 54-bit Q, Bks_R=2^9 (6 digits), %d inputs per call; blind_rotate (11 CMUX, "
                    "BG=2^9 d=6) over %d RLWEs; homo_expand of one packed query; %d CMUX trees of %d levels"
                    % (B, R, T, L),
        "expansion_ms_per_call": round(ms, 3), "expansions_per_s": round(B / (ms * 1e-3), 1),
        "subs_keyswitch_per_s": round(B * (N - 1) / (ms * 1e-3), 1),
        "blind_rotate_ms_per_call": round(ms_br, 3), "cmux_per_s": round(R * 11 / (ms_br * 1e-3), 1),
        "homo_expand_ms_per_query": round(ms_he, 3),
        "homo_expand_external_products_per_s": round(N * d / (ms_he * 1e-3), 1),
        "cmux_tree_ms_per_call": round(ms_tree, 3),
        "tree_cmux_per_s": round(T * ((1 << L) - 1) / (ms_tree * 1e-3), 1),
    }


def app256(fhe, torch, B=2048, reps=3):
    """The reference's 256-bit application set (apps/bootstrap_fpga_256.cpp:
    n=1024 q=1024 N=2048, 29-bit Q, BG=2^10, Br=32, Bks=25): 32-bit device
    words with Harvey's butterflies."""
    P = dict(n=1024, q=1024, N=2048, Q=0, nbit=29, BGbits=10, Br=32, Bks=25)
    ctx = fhe.Context(**P)
    keys = ctx.keys()
    s, z = ctx.lwe_secret(SEED + 5, binary=True), ctx.rlwe_secret(SEED + 5)
    keys.generate(SEED + 5, s, z)
    tune = ctx.autotune(keys, min(B, 1024))
    rng = np.random.default_rng(SEED + 5)
    m1, m2 = rng.integers(0, 2, B).astype(np.uint8), rng.integers(0, 2, B).astype(np.uint8)
    c1, c2 = ctx.lwe_encrypt(s, m1, SEED + 5, 0), ctx.lwe_encrypt(s, m2, SEED + 5, B)
    gates = np.full(B, fhe.GATES["NAND"], dtype=np.int32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int64) if a.dtype == np.uint64 else a).cuda()
    d1, d2, dg = t(c1), t(c2), t(gates)
    out = torch.zeros_like(d1)
    ctx.bootstrap_dev(keys, dg, d1, d2, out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ctx.bootstrap_dev(keys, dg, d1, d2, out)
    e1.record()
    torch.cuda.synchronize()
    ctx.sync()
    ms = e0.elapsed_time(e1) / reps
    bits = ctx.lwe_decrypt(s, out.cpu().numpy().view(np.uint64))
    info = ctx.info
    return {"workload": "batched NAND gate bootstrap, apps/bootstrap_fpga_256.cpp parameters (n=1024 q=1024 "
                        "N=2048 Q=%s BG=2^10 Br=32 Bks=25), batch %d, eager" % (hex(info["Q"]), B),
            "device_words": "u32, Harvey butterflies" if info["harvey"] else ("u32" if info["narrow"] else "u64"),
            "ms_per_batch": round(ms, 3), "bootstraps_per_s": round(B / (ms / 1e3), 1),
            "decrypt_ok": bool(np.array_equal(bits, 1 - (m1 & m2))), "tuning": tune.get("chosen")}


def copy_ceiling(torch, nbytes=1 << 30, reps=5):
    """Measured device-to-device copy rate (bytes read + bytes written per
    second): the practical HBM ceiling beside the nominal 8 TB/s."""
    src = torch.empty(nbytes // 8, dtype=torch.int64, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def fpga_orders(fhe, torch, ctx, keys, dg, d1, d2, nega, s, want, reps=3):
    """The same batch in the accelerator's step orders (fpga_eval_bootstrap_x1
    and _x4..x16, myfpga.cpp:1003-1660): x1 adds a product by btkey[0][0][0]
    when step (0, 0) has residue 0, the batched forms one by RGSW(1) at every
    residue-0 step."""
    Br, dR = PARAMS["Br"], ctx.digitR
    digits = np.stack([(nega // Br ** j) % Br for j in range(dR)], axis=-1)  # [B][n][digitR]
    soft = np.count_nonzero(digits, axis=(1, 2))
    counts = {"software": soft, "x1": soft + (digits[:, 0, 0] == 0), "xn": np.full(soft.shape, ctx.n * dR)}
    res = {"workload": "the headline batch (4096 NAND gates, benchmark parameters), eager launches"}
    out = torch.zeros_like(d1)
    for name, mode in (("x1", fhe.EVAL_FPGA_X1), ("xn", fhe.EVAL_FPGA_XN)):
        ctx.eval_mode = mode
        ctx.bootstrap_dev(keys, dg, d1, d2, out)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ctx.bootstrap_dev(keys, dg, d1, d2, out)
        e1.record()
        torch.cuda.synchronize()
        ctx.sync()
        ms = e0.elapsed_time(e1) / reps
        bits = ctx.lwe_decrypt(s, out.cpu().numpy().view(np.uint64))
        res[name] = {"ms_per_batch": round(ms, 3), "bootstraps_per_s": round(d1.shape[0] / (ms / 1e3), 1),
                     "external_products_per_bootstrap": round(float(counts[name].mean()), 2),
                     "decrypt_ok": bool(np.array_equal(bits, want))}
    ctx.eval_mode = fhe.EVAL_SOFTWARE
    res["software_external_products_per_bootstrap"] = round(float(counts["software"].mean()), 2)
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


# worker threads this process may run on the host: the GPU host's CPU share
# for one GPU is 16 threads, whatever the machine has
HOST_THREAD_SHARE = 16


def cpu_baseline(ctx, keys, c1, c2, gates, gpu_out, sample):
    """The CPU path on the same batch: the oracle's restatement of
    RLWE::eval_bootstrap (reference-faithful u128 '%' arithmetic) on one thread
    and on the host cores, each compared bit for bit with the GPU; the
    reference's own eval_bootstrap compiled from its sources (oracle/_ref),
    when present, as a labelled cross-check."""
    import oracle
    from oracle import refbind

    bsk = keys.download_bsk()
    ksk = keys.download_ksk()
    o = oracle.Oracle(**PARAMS)
    c1s, c2s, gs = c1[:sample], c2[:sample], gates[:sample]
    t0 = time.perf_counter()
    out = o.bootstrap(c1s, c2s, gs, bsk, ksk, threads=1)
    secs = time.perf_counter() - t0
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(HOST_THREAD_SHARE, affinity))
    mc = min(len(c1), 4 * threads)
    t0 = time.perf_counter()
    out_mc = o.bootstrap(c1[:mc], c2[:mc], gates[:mc], bsk, ksk, threads=threads)
    mc_secs = time.perf_counter() - t0
    rec = {
        "value": round(sample / secs, 3),
        "unit": "bootstraps/s",
        "cores": 1,
        "kind": "port",
        "sample": "%d NAND gate bootstraps of the benchmark batch, single thread, oracle/fhe_oracle.c restatement "
                  "of RLWE::eval_bootstrap (RLWE.cpp:726-786) with u128 '%%' arithmetic" % sample,
        "seconds": round(secs, 3),
        "bit_exact_vs_gpu": bool(np.array_equal(out, gpu_out[:sample])),
        "all_cores": {"value": round(mc / mc_secs, 3), "unit": "bootstraps/s", "cores": threads, "kind": "port",
                      "sample": "%d NAND gate bootstraps, one ciphertext per thread, %d threads" % (mc, threads),
                      "bit_exact_vs_gpu": bool(np.array_equal(out_mc, gpu_out[:mc]))},
        "host": {"cpu_model": cpu_model(), "hardware_concurrency": os.cpu_count(), "affinity_cpus": affinity,
                 "thread_share": HOST_THREAD_SHARE,
                 "note": "the all-cores leg runs min(affinity, the host's 16-thread share per GPU) threads"},
    }
    if refbind.available():
        rs = min(sample, 10)
        ref = refbind.Reference(ctx.n, ctx.q, ctx.N, ctx.Q, PARAMS["BGbits"], PARAMS["Br"], PARAMS["Bks"])
        ref.load_keys(bsk, ksk)
        rout, rsecs = ref.bootstrap(c1[:rs], c2[:rs], gates[:rs])
        rec["reference_cross_check"] = {
            "value": round(rs / rsecs, 3), "unit": "bootstraps/s", "cores": 1, "kind": "reference",
            "sample": "%d gates through RLWE::eval_bootstrap compiled from the reference sources (oracle/_ref)" % rs,
            "bit_exact_vs_gpu": bool(np.array_equal(rout, gpu_out[:rs]))}
    return rec


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this library has no multi-GPU path (asked for --gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import torch

    fhe = importlib.import_module("3rd_gen_fhe_acc_amd")
    torch.cuda.set_device(0)
    if args.copy_calibration:
        src = torch.empty(1 << 27, dtype=torch.int64, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.only_set:
        print(json.dumps(one_set(fhe, torch, args)))
        return
    B = args.batch
    ctx = fhe.Context(**PARAMS)
    keys = ctx.keys()
    s = ctx.lwe_secret(SEED, binary=True)
    z = ctx.rlwe_secret(SEED)
    t0 = time.perf_counter()
    keys.generate(SEED, s, z)
    ctx.sync()
    keygen_s = time.perf_counter() - t0
    if args.tuning == "auto":
        tune_report = ctx.autotune(keys, min(B, 2048))
    else:
        th, pb = (int(x) for x in args.tuning.split(","))
        ctx.set_tuning(th, pb)
        tune_report = {"chosen": {"threads": th, "polys_per_batch": pb}, "source": "--tuning"}
    if args.schedule != "auto":
        ctx.set_schedule(*(int(x) for x in args.schedule.split(",")))
    schedule = ctx.get_schedule()
    rng = np.random.default_rng(SEED)
    m1 = rng.integers(0, 2, B).astype(np.uint8)
    m2 = rng.integers(0, 2, B).astype(np.uint8)
    c1 = ctx.lwe_encrypt(s, m1, SEED, 0)
    c2 = ctx.lwe_encrypt(s, m2, SEED, B)
    gates = np.full(B, fhe.GATES["NAND"], dtype=np.int32)

    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a).view(np.int64) if a.dtype == np.uint64 else a).cuda()

    d1, d2, dg = dev(c1), dev(c2), dev(gates)
    dout = torch.zeros_like(d1)
    for _ in range(args.warmup):
        ctx.bootstrap_dev(keys, dg, d1, d2, dout)
    ctx.sync()
    torch.cuda.synchronize()
    # eager launches, with per-kernel HIP events on the launch stream
    ctx.profile(True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(args.steps):
        ctx.bootstrap_dev(keys, dg, d1, d2, dout)
    ev1.record()
    torch.cuda.synchronize()
    ctx.sync()
    kms = ctx.kernel_ms()
    ctx.profile(False)
    eager_ms = ev0.elapsed_time(ev1) / args.steps
    # the timed form: the batch captured once into a HIP graph and replayed
    gout = torch.zeros_like(d1)
    graph = ctx.capture_bootstrap(keys, dg, d1, d2, gout)
    for _ in range(args.warmup):
        graph.launch()
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(args.steps):
        graph.launch()
    ev1.record()
    torch.cuda.synchronize()
    ctx.sync()
    ms_step = ev0.elapsed_time(ev1) / args.steps
    graph.close()
    value = B / (ms_step / 1e3)
    out = gout.cpu().numpy().view(np.uint64)
    graph_equals_eager = bool(np.array_equal(out, dout.cpu().numpy().view(np.uint64)))
    bits = ctx.lwe_decrypt(s, out)
    truth_ok = bool(np.array_equal(bits, 1 - (m1 & m2)))

    # one ciphertext per launch
    lat = []
    for i in range(args.single + 2 if args.single > 0 else 0):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ctx.bootstrap_dev(keys, dg[i:i + 1], d1[i:i + 1], d2[i:i + 1], dout[i:i + 1])
        e1.record()
        torch.cuda.synchronize()
        if i >= 2:
            lat.append(e0.elapsed_time(e1))
    ctx.sync()
    lat = np.array(lat if lat else [float("nan")])
    single_rate = 1e3 / lat.mean()

    copy_gbs = None if args.no_copy_ceiling else copy_ceiling(torch)
    # roofline of the dominant kernel (blind rotation), algorithmic bytes
    info = ctx.info
    n, N, dks = PARAMS["n"], PARAMS["N"], ctx.digitKS
    br_ms, br_n = kms["blind_rotate"]
    ks_ms, ks_n = kms["lwe_keyswitch"]
    br_avg = br_ms / max(br_n, 1)
    ks_avg = ks_ms / max(ks_n, 1)
    # blind-rotation launches per batch: 1, or the tiles of a tiled schedule
    # (P,0 runs the batch as launches of one resident wave of workgroups);
    # the per-batch bytes and products below are divided by it, so that they
    # are per launch like the launch time and the per-dispatch counters
    br_per_batch = max(br_n, 1) / args.steps
    # the key bytes a batch touches: every nonzero residue (residue 0 is
    # stored but the software order never reads it)
    Br = PARAMS["Br"]
    touched_bsk = info["bsk_device_bytes"] // Br * (Br - 1)
    br_bytes = int(round((touched_bsk + 2 * B * (n + 1) * 8 + B * 4 + B * N * dks + B * 8) / br_per_batch))
    achieved = br_bytes / (br_avg * 1e-3) / 1e9
    nega = (PARAMS["q"] - ((c1[:, :n] + c2[:, :n]) % PARAMS["q"])) % PARAMS["q"]
    ext_products = int(round((np.count_nonzero(nega % 32) + np.count_nonzero((nega // 32) % 32)) / br_per_batch))
    d = ctx.digitG
    modmul_per_ep = 2 * d * (N // 2) * 10 + 2 * (N // 2) * 10 + 2 * N + 4 * d * N
    # HBM traffic per launch from the committed FETCH/WRITE summary, for the
    # schedule this run uses, only while it describes these kernel sources
    sched_key = "%d,%d" % tuple(schedule)
    tprof, traffic_src = load_profile("traffic_blind_rotate_n1024_b4096.json")
    tent = (tprof or {}).get("schedules", {}).get(sched_key)
    traffic_src["schedule"] = sched_key
    if tent is None and traffic_src["fresh"]:
        traffic_src.update(fresh=False, why="no counters for schedule " + sched_key)
    traffic = tent["hbm_bytes_per_launch"] if tent and traffic_src["fresh"] else None
    if tent:
        traffic_src["kernel_ms_in_counter_run"] = tent.get("avg_ms_in_counter_run")
    # integer-VALU roofline: the kernel's VALU instructions per launch (from
    # the committed rocprofv3 SQ counter summary of this workload, same
    # freshness rule) over the live launch time, against 256 CUs x 4 SIMDs x
    # 16 lanes x 2.4 GHz
    valu_peak = 256 * 4 * 16 * 2.4e9
    sprof, valu_src = load_profile("sq_counters_blind_rotate_n1024_b4096.json")
    valu_instr = None
    sent = None
    if sprof:
        # the full counter set is for the default schedule; the other
        # schedules the autotuner can pick have their VALU counts beside it
        if sprof.get("schedule", "0,0") == sched_key:
            sent = sprof["counters_per_launch"]
        else:
            sent = sprof.get("schedules", {}).get(sched_key)
    valu_src["schedule"] = sched_key
    if sent is None and valu_src["fresh"]:
        valu_src.update(fresh=False, why="no counters for schedule " + sched_key)
    if sent and valu_src["fresh"]:
        valu_instr = sent.get("SQ_INSTS_VALU")
    # the key switch: its key, digits, b words and outputs
    ks_bytes = info["ksk_device_bytes"] + B * N * dks + 8 * B + B * (n + 1) * 8
    ks_traffic = (tprof or {}).get("keyswitch", {}).get("hbm_bytes_per_launch") if traffic_src["fresh"] else None
    rec = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "bootstraps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "batched NAND gate bootstrap, benchmark parameters (n=500 q=1024 N=1024 "
                               "Q=0x7FFF801 BG=2^7 d=4 Br=32 Bks=2^5), batch %d" % B,
                   "global_batch": B, "parallelism": "single GPU", "Q": ctx.Q, "digitG": d,
                   "digitR": ctx.digitR, "digitKS": dks},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "k_blind_rotate", "avg_launch_ms": round(br_avg, 4),
                     "algorithmic_bytes_per_launch": br_bytes, "launches_per_batch": round(br_per_batch, 3),
                     "copy_ceiling_GBps": round(copy_gbs, 1) if copy_gbs else None,
                     "frac_of_copy_ceiling": round(achieved / copy_gbs, 5) if copy_gbs else None},
        "roofline_keyswitch": {"kernel": "k_lwe_keyswitch_staged", "bound": "hbm",
                               "achieved": round(ks_bytes / (ks_avg * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(ks_bytes / (ks_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                               "traffic": ks_traffic, "avg_launch_ms": round(ks_avg, 4),
                               "algorithmic_bytes_per_launch": ks_bytes},
        "traffic_source": traffic_src,
        "valu_source": valu_src,
        "valu": {"kernel": "k_blind_rotate", "bound": "integer VALU (the kernel's roofline; no MFMA applies)",
                 "achieved": round(valu_instr * 64 / (br_avg * 1e-3) / 1e12, 3) if valu_instr else None,
                 "peak": round(valu_peak / 1e12, 3), "unit": "T lane-ops/s",
                 "frac": round(valu_instr * 64 / (br_avg * 1e-3) / valu_peak, 4) if valu_instr else None,
                 "valu_instr_per_launch": valu_instr,
                 "external_products_per_launch": ext_products, "modmul_per_external_product": modmul_per_ep,
                 "modmul_per_s_T": round(ext_products * modmul_per_ep / (br_avg * 1e-3) / 1e12, 4)},
        "kernels_ms_avg": {"blind_rotate": round(br_avg, 4), "lwe_keyswitch": round(ks_avg, 4)},
        "single_per_launch": {"bootstraps_per_s": round(single_rate, 2),
                              "latency_ms_p50": round(float(np.percentile(lat, 50)), 4),
                              "latency_ms_p99": round(float(np.percentile(lat, 99)), 4),
                              "samples": int(lat.size)},
        "speedup_vs_single_per_launch": round(value / single_rate, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                       "p99": round(float(np.percentile(lat, 99)), 4),
                       "note": "single gate, idle GPU (one ciphertext per launch); under load: serving"},
        "batch_latency_ms": round(ms_step, 4),
        "decrypt_truth_table_ok": truth_ok,
        "graph_equals_eager": graph_equals_eager,
        "eager_ms_per_step": round(eager_ms, 4),
        "keygen_s": round(keygen_s, 3),
        "tuning": tune_report,
        "schedule": {"persist": schedule[0], "sync_window": schedule[1]},
    }
    if args.serve_seconds > 0:
        clients, keygen_s_clients = serving_clients(ctx)
        rec["serving"] = serving(fhe, ctx, clients, args.serve_qps, args.serve_seconds, args.serve_max_batch,
                                 args.serve_inflight, args.serve_min_fill, not args.serve_fifo)
        rec["serving"]["client_keygen_s"] = round(keygen_s_clients, 3)
        if args.serve_curve:
            rec["serving_curve"] = serving_curve(fhe, ctx, clients, args)
        del clients
        rec["latency_ms"] = {"p50": rec["serving"]["gate_latency_ms_p50"],
                             "p99": rec["serving"]["gate_latency_ms_p99"],
                             "request_p50": rec["serving"]["request_latency_ms_p50"],
                             "request_p99": rec["serving"]["request_latency_ms_p99"],
                             "note": "gate ready -> done and request submit -> result, serving loop under "
                                     "Poisson load; idle one-per-launch figures: single_per_launch"}
    if not args.no_fpga_order:
        rec["fpga_order"] = fpga_orders(fhe, torch, ctx, keys, dg, d1, d2, nega, s, 1 - (m1 & m2))
    if not args.no_microbench:
        rec["kernel_microbench"] = microbench(fhe, torch, B)
    if not args.no_mixed:
        rec["mixed"] = mixed(fhe, torch, ctx, keys)
    if not args.no_leveled:
        rec["leveled"] = leveled(fhe, torch)
    if not args.no_app256:
        rec["app_256"] = app256(fhe, torch)
    if not args.no_cpu_baseline and args.cpu_sample > 0:
        rec["cpu_baseline"] = cpu_baseline(ctx, keys, c1, c2, gates, out, args.cpu_sample)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
