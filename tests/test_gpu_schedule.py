"""The blind rotation's schedules (fheacc_ctx_set_schedule): one workgroup per
ciphertext, a persistent grid walking the batch round by round, and the
persistent grid paced in lock-step by device counters.  Pacing only orders the
work, so every schedule must give the words of RLWE::eval_bootstrap
(RLWE.cpp:726-786) and of the accelerator's step orders: checked against the
oracle on ragged multi-round batches, with flagged ciphertexts in the batch,
and at the benchmark's full size through graph replays."""
import numpy as np
import pytest

from fheacc_testlib import CFG0, SETS, fhe
from oracle import Oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# (workgroups per CU, lock-step window): persist 1 at B = 600 is three rounds
# on 256 CUs (256, 256, 88); 64 asks for as many as fit
SCHEDULES = [(0, 0), (1, 0), (1, 1), (2, 2), (64, 0), (64, 1)]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64) if a.dtype == np.uint64 else a).cuda()


@pytest.mark.parametrize("name", ["n12_N512_q27", "n12_N1024_q27", "n8_N2048_q54"])
@pytest.mark.parametrize("mode", [fhe.EVAL_SOFTWARE, fhe.EVAL_FPGA_XN])
def test_schedules_bit_exact(name, mode):
    kw = SETS[name]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    ctx.eval_mode = mode
    o.eval_mode = mode
    s, z = ctx.lwe_secret(81), ctx.rlwe_secret(81)
    keys = ctx.keys()
    keys.generate(81, s, z)
    B = 600
    rng = np.random.default_rng(81)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    c1, c2 = ctx.lwe_encrypt(s, m1, 81, 0), ctx.lwe_encrypt(s, m2, 81, B)
    gates = rng.integers(0, 6, B).astype(np.int32)
    exp = o.bootstrap(c1, c2, gates, keys.download_bsk(), keys.download_ksk(), threads=16)
    for persist, window in SCHEDULES:
        print(name, mode, "schedule", persist, window, flush=True)
        ctx.set_schedule(persist, window)
        assert ctx.get_schedule() == (persist, window)
        assert np.array_equal(ctx.bootstrap(keys, gates, c1, c2), exp), (persist, window)


def test_paced_schedule_with_flagged_ciphertexts():
    # a ciphertext with a bad gate skips its products but still marks its
    # steps, so its group is not held up; identical operands are reported
    kw = SETS["n12_N512_q27"]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    s, z = ctx.lwe_secret(82), ctx.rlwe_secret(82)
    keys = ctx.keys()
    keys.generate(82, s, z)
    B = 700
    rng = np.random.default_rng(82)
    c1 = ctx.lwe_encrypt(s, rng.integers(0, 2, B), 82, 0)
    c2 = ctx.lwe_encrypt(s, rng.integers(0, 2, B), 82, B)
    gates = rng.integers(0, 6, B).astype(np.int32)
    bad = gates.copy()
    bad[[5, 300, 699]] = [9, -2, 6]
    same = c2.copy()
    same[17] = c1[17]
    ctx.set_schedule(1, 1)
    d1, dg = dev(c1), dev(bad)
    out = torch.zeros_like(d1)
    ctx.bootstrap_dev(keys, dg, d1, dev(c2), out)
    with pytest.raises(fhe.FheAccError, match="gate out of range"):
        ctx.sync()
    ok = np.ones(B, dtype=bool)
    ok[[5, 300, 699]] = False
    exp = o.bootstrap(c1[ok], c2[ok], gates[ok], keys.download_bsk(), keys.download_ksk(), threads=16)
    assert np.array_equal(out.cpu().numpy().view(np.uint64)[ok], exp)
    ctx.bootstrap_dev(keys, dev(gates), d1, dev(same), out)
    with pytest.raises(fhe.FheAccError, match="identical operands"):
        ctx.sync()


def test_full_size_graph_replays_equal_eager():
    # the benchmark's workload (n=500, N=1024, 27-bit Q, batch 4096) captured
    # with the paced persistent schedule and replayed three times, new inputs
    # written into the captured buffers and eager work between replays; each
    # replay equals the eager one-workgroup-per-ciphertext result and decrypts
    # to the gates' truth tables
    ctx = fhe.Context(**CFG0)
    s, z = ctx.lwe_secret(83), ctx.rlwe_secret(83)
    keys = ctx.keys()
    keys.generate(83, s, z)
    B = 4096
    g = torch.zeros(B, dtype=torch.int32, device="cuda")
    d1 = torch.zeros((B, ctx.n + 1), dtype=torch.int64, device="cuda")
    d2 = torch.zeros_like(d1)
    gout = torch.zeros_like(d1)
    ctx.set_schedule(64, 2)
    graph = ctx.capture_bootstrap(keys, g, d1, d2, gout)
    rng = np.random.default_rng(83)
    for rep in range(3):
        print("replay", rep, flush=True)
        m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
        gates = rng.integers(0, 6, B).astype(np.int32)
        c1, c2 = ctx.lwe_encrypt(s, m1, 84, 2 * B * rep), ctx.lwe_encrypt(s, m2, 84, 2 * B * rep + B)
        g.copy_(torch.from_numpy(gates))
        d1.copy_(torch.from_numpy(c1.view(np.int64)))
        d2.copy_(torch.from_numpy(c2.view(np.int64)))
        torch.cuda.synchronize()
        graph.launch()
        ctx.sync()
        got = gout.cpu().numpy().view(np.uint64).copy()
        ctx.set_schedule(0, 0)
        eager = ctx.bootstrap(keys, gates, c1, c2)
        ctx.set_schedule(64, 2)
        assert np.array_equal(got, eager), rep
        truth = {0: m1 | m2, 1: m1 & m2, 2: 1 - (m1 | m2), 3: 1 - (m1 & m2), 4: m1 ^ m2, 5: 1 - (m1 ^ m2)}
        exp = np.array([truth[int(t)][i] for i, t in enumerate(gates)])
        assert np.array_equal(ctx.lwe_decrypt(s, got), exp), rep
    graph.close()
