"""Failure paths on the GPU that the host can force without a fault: device
allocations that cannot succeed return -ENOMEM with a message, the
accelerator driver's convention (myfpga.cpp:1011-1015), and leave the
context usable."""
import numpy as np
import pytest

from fheacc_testlib import SETS, fhe
from oracle import Oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ENOMEM = -12


def test_bootstrap_key_larger_than_the_device():
    # n = 4096, N = 2048, Br = 1024 (one digit): 2^22 RGSWs of 128 KiB, a
    # 512 GiB key
    ctx = fhe.Context(n=4096, q=1024, N=2048, Q=0, nbit=27, BGbits=7, Br=1024, Bks=32)
    assert ctx.info["bsk_device_bytes"] >= 1 << 39
    keys = ctx.keys()
    s, z = ctx.lwe_secret(1), ctx.rlwe_secret(1)
    with pytest.raises(fhe.FheAccError, match="error %d" % ENOMEM):
        keys.generate(1, s, z)


def test_serving_pool_larger_than_the_device():
    ctx = fhe.Context(**SETS["n16_N256_q27"])
    with pytest.raises(fhe.FheAccError, match="error %d" % ENOMEM):
        fhe.Server(ctx, hbm_budget=1 << 40, max_batch=64)


def test_context_still_works_after_a_failed_allocation():
    kw = SETS["n16_N256_q27"]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    with pytest.raises(fhe.FheAccError, match="error %d" % ENOMEM):
        fhe.Server(ctx, hbm_budget=1 << 40, max_batch=64)
    s, z = ctx.lwe_secret(3), ctx.rlwe_secret(3)
    keys = ctx.keys()
    keys.generate(3, s, z)
    rng = np.random.default_rng(3)
    m1, m2 = rng.integers(0, 2, 8), rng.integers(0, 2, 8)
    c1, c2 = ctx.lwe_encrypt(s, m1, 3, 0), ctx.lwe_encrypt(s, m2, 3, 8)
    gates = np.full(8, 3, dtype=np.int32)
    out = ctx.bootstrap(keys, gates, c1, c2)
    assert np.array_equal(out, o.bootstrap(c1, c2, gates, keys.download_bsk(), keys.download_ksk()))


def _small_batch(ctx, s, B, seed):
    rng = np.random.default_rng(seed)
    m1, m2 = rng.integers(0, 2, B), rng.integers(0, 2, B)
    return ctx.lwe_encrypt(s, m1, seed, 0), ctx.lwe_encrypt(s, m2, seed, B)


def test_gate_out_of_range_flagged_on_device():
    # the device form re-reads gates on every launch: a value outside enum
    # GATES is flagged on the device (no out-of-bounds read), the call's
    # sync reports it, and the other ciphertexts keep their words
    kw = SETS["n16_N256_q27"]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    s, z = ctx.lwe_secret(5), ctx.rlwe_secret(5)
    keys = ctx.keys()
    keys.generate(5, s, z)
    B = 8
    c1, c2 = _small_batch(ctx, s, B, 5)
    gates = np.full(B, 1, dtype=np.int32)
    gates[3], gates[6] = 7, -1
    dev = "cuda:%d" % ctx.device
    d1, d2 = torch.from_numpy(c1.view(np.int64)).to(dev), torch.from_numpy(c2.view(np.int64)).to(dev)
    dg = torch.from_numpy(gates).to(dev)
    dout = torch.zeros_like(d1)
    ctx.bootstrap_dev(keys, dg, d1, d2, dout)
    with pytest.raises(fhe.FheAccError, match="gate out of range"):
        ctx.sync()
    out = dout.cpu().numpy().view(np.uint64)
    ok = [i for i in range(B) if 0 <= gates[i] <= 5]
    exp = o.bootstrap(c1[ok], c2[ok], gates[ok], keys.download_bsk(), keys.download_ksk())
    assert np.array_equal(out[ok], exp)
    # the status word was cleared: a good batch after it succeeds
    dg.fill_(1)
    ctx.bootstrap_dev(keys, dg, d1, d2, dout)
    ctx.sync()
    # the host form refuses the gate before launching
    with pytest.raises(fhe.FheAccError, match="gate out of range"):
        ctx.bootstrap(keys, gates, c1, c2)


def test_keys_of_another_context_rejected():
    ctx_a = fhe.Context(**SETS["n16_N256_q27"])
    ctx_b = fhe.Context(**SETS["n12_N512_q27"])
    s, z = ctx_a.lwe_secret(6), ctx_a.rlwe_secret(6)
    keys_a = ctx_a.keys()
    keys_a.generate(6, s, z)
    sb = ctx_b.lwe_secret(6)
    c1, c2 = _small_batch(ctx_b, sb, 4, 6)
    with pytest.raises(fhe.FheAccError, match="another context"):
        ctx_b.bootstrap(keys_a, np.full(4, 3, dtype=np.int32), c1, c2)
    with pytest.raises(fhe.FheAccError, match="another context"):
        ctx_b.autotune(keys_a, 4)


def test_context_destroyed_before_its_keys():
    # key sets hold a reference to their context: destroying the context
    # first leaves the keys usable until they are destroyed themselves
    kw = SETS["n16_N256_q27"]
    ctx = fhe.Context(**kw)
    s, z = ctx.lwe_secret(8), ctx.rlwe_secret(8)
    keys = ctx.keys()
    keys.generate(8, s, z)
    before = keys.download_ksk()
    ctx.close()
    assert np.array_equal(keys.download_ksk(), before)
    keys.close()
