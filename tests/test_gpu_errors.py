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
