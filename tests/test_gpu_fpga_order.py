"""GPU parity for the accelerator's blind-rotation orders
(fheacc_ctx_set_eval_mode): fpga_eval_bootstrap_x1 multiplies at step (0, 0)
whatever the residue, the x4..x16 forms at every step with btkey[0] = RGSW(1)
for residue 0 (myfpga.cpp:1030-1071, 1180-1225).  The HIP library through its
C ABI against the oracle in the same order, word for word, and against the
trivial-key closed form (fpga_order_cases.py).  The oracle's orders are
pinned to the reference's own routines in tests/test_fpga_order.py."""
import numpy as np
import pytest

from fheacc_testlib import SETS, fhe
from fpga_order_cases import EVAL_FPGA_X1, EVAL_FPGA_XN, EVAL_SOFTWARE, MODES, expected_acc, lwe_pairs, \
    trivial_bsk
from oracle import Oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

GSETS = {
    # Br = 4 (digitR = 5): a quarter of all steps have residue 0
    "n16_N256_q27_br4": dict(n=16, q=512, N=256, Q=0, nbit=27, BGbits=9, Br=4, Bks=25),
    # wide words
    "n10_N128_q54": SETS["n10_N128_q54"],
    # the benchmark's N = 1024 shape, n reduced
    "n12_N1024_q27": SETS["n12_N1024_q27"],
}


def dev(a, dtype=np.uint64):
    a = np.ascontiguousarray(a, dtype=dtype)
    return torch.from_numpy(a.view(np.int64) if dtype == np.uint64 else a).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint64)


def keyed(name, seed):
    kw = GSETS[name]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    s, z = ctx.lwe_secret(seed), ctx.rlwe_secret(seed)
    keys = ctx.keys()
    keys.generate(seed, s, z)
    return ctx, o, keys


@pytest.mark.parametrize("name", list(GSETS))
@pytest.mark.parametrize("mode", [EVAL_FPGA_X1, EVAL_FPGA_XN])
def test_fpga_order_matches_oracle(name, mode):
    ctx, o, keys = keyed(name, 21)
    ctx.eval_mode = mode
    o.eval_mode = mode
    B = 12
    c1, c2 = lwe_pairs(o, B, 22)
    gates = (np.arange(B) % 6).astype(np.int32)
    out, acc, ks = ctx.bootstrap(keys, gates, c1, c2, capture=True)
    eo, ea, ek = o.bootstrap(c1, c2, gates, keys.download_bsk(), keys.download_ksk(), capture=True)
    assert np.array_equal(acc, ea)
    assert np.array_equal(ks, ek)
    assert np.array_equal(out, eo)


@pytest.mark.parametrize("name", ["n16_N256_q27_br4", "n10_N128_q54"])
@pytest.mark.parametrize("mode", MODES)
def test_trivial_keys_closed_form(name, mode):
    kw = GSETS[name]
    ctx, o = fhe.Context(**kw), Oracle(**kw)
    bsk, exps = trivial_bsk(o, 5)
    keys = ctx.keys()
    keys.upload_bsk(bsk)
    keys.upload_ksk(np.zeros(o.ksk_words, dtype=np.uint64))
    ctx.eval_mode = mode
    B = 6
    c1, c2 = lwe_pairs(o, B, 6)
    gates = (np.arange(B) % 6).astype(np.int32)
    _, acc, _ = ctx.bootstrap(keys, gates, c1, c2, capture=True)
    for b in range(B):
        ev = o.evaluate(c1[b], c2[b], int(gates[b]))
        assert np.array_equal(acc[b], expected_acc(o, ev, int(gates[b]), exps, mode)), b


def test_graph_keeps_the_order_it_was_captured_with():
    ctx, o, keys = keyed("n16_N256_q27_br4", 23)
    bsk, ksk = keys.download_bsk(), keys.download_ksk()
    B = 8
    c1, c2 = lwe_pairs(o, B, 24)
    gates = (np.arange(B) % 6).astype(np.int32)
    out = torch.zeros((B, o.n + 1), dtype=torch.int64, device="cuda")
    d1, d2, dg = dev(c1), dev(c2), dev(gates, np.int32)
    ctx.eval_mode = EVAL_FPGA_XN
    graph = ctx.capture_bootstrap(keys, dg, d1, d2, out)
    ctx.eval_mode = EVAL_SOFTWARE
    o.eval_mode = EVAL_FPGA_XN
    want = o.bootstrap(c1, c2, gates, bsk, ksk)
    try:
        for rep in range(3):
            out.zero_()
            eager = ctx.bootstrap(keys, gates, c1, c2)  # software order in between
            graph.launch()
            ctx.sync()
            assert np.array_equal(host(out), want), rep
    finally:
        graph.close()
    o.eval_mode = EVAL_SOFTWARE
    assert np.array_equal(eager, o.bootstrap(c1, c2, gates, bsk, ksk))
    assert not np.array_equal(eager, want)


def test_serving_loop_uses_the_context_order():
    ctx, o, keys = keyed("n16_N256_q27_br4", 25)
    ctx.eval_mode = EVAL_FPGA_X1
    o.eval_mode = EVAL_FPGA_X1
    B = 10
    c1, c2 = lwe_pairs(o, B, 26)
    gates = (np.arange(B) % 6).astype(np.int32)
    # one level: gate b reads input nodes b and B + b, writes node 2B + b
    ops = np.array([(int(g), b, B + b) for b, g in enumerate(gates)], dtype=np.int32)
    srv = fhe.Server(ctx, hbm_budget=1 << 28, max_batch=64)
    try:
        cid = srv.add_client(keys)
        got, _ = srv.wait(srv.submit(cid, np.concatenate([c1, c2]), ops, 2 * B + np.arange(B)))
    finally:
        srv.close()
    assert np.array_equal(got, o.bootstrap(c1, c2, gates, keys.download_bsk(), keys.download_ksk()))


def test_eval_mode_argument():
    ctx = fhe.Context(**GSETS["n10_N128_q54"])
    assert ctx.eval_mode == EVAL_SOFTWARE
    for bad in (-1, 3):
        with pytest.raises(fhe.FheAccError, match="evaluation mode"):
            ctx.eval_mode = bad
    ctx.eval_mode = EVAL_FPGA_XN
    assert ctx.eval_mode == EVAL_FPGA_XN


# ---- the accelerator's DDR key images (myfpga.cpp:596-712)

def from_card_buffer(img, N):
    """convert_buffer_to_RLWE (myfpga.h:226-243) on the last axis of 2N words:
    blocks of 4 words of a then 4 of b -> [..., 2, N] (a, b)."""
    blk = img.reshape(img.shape[:-1] + (N // 4, 2, 4))
    return np.moveaxis(blk, -2, -3).reshape(img.shape[:-1] + (2, N))


@pytest.mark.parametrize("name", ["n16_N256_q27_br4", "n10_N128_q54"])
def test_fpga_bsk_image(name):
    ctx, o, keys = keyed(name, 31)
    n, dR, Br, d, N = o.n, o.digitR, o.Br, o.digitG, o.N
    ref = keys.download_bsk().reshape(Br, dR, n, d, 2, 2, N)  # btkey[r][j][i].c_text[h][g].{a,b}
    img = keys.export_fpga_bsk()
    assert img.size == ref.size
    rgsw_bytes = 2 * d * N * 2 * 8  # RGSW_size (myfpga.cpp:1006,1133)
    for i in range(n):
        for j in range(dR):
            for r in range(Br):
                # the key address the evaluation issues for step (i, j, r) (myfpga.cpp:1038,1049)
                w0 = ((i * dR + j) * Br + r) * rgsw_bytes // 8
                rg = img[w0:w0 + rgsw_bytes // 8].reshape(2, d, 2 * N)  # RLWEs c_text[h][g], g-major (myfpga.cpp:624-628)
                got = from_card_buffer(rg, N)  # [g][h][a, b][N]
                assert np.array_equal(got, ref[r, j, i].transpose(1, 0, 2, 3)), (i, j, r)
    # an imported image is the same resident key
    k2 = ctx.keys()
    k2.import_fpga_bsk(img)
    k2.upload_ksk(keys.download_ksk())
    assert np.array_equal(k2.download_bsk(), ref.reshape(-1))
    assert np.array_equal(k2.export_fpga_bsk(), img)
    c1, c2 = lwe_pairs(o, 4, 32)
    gates = np.arange(4, dtype=np.int32)
    assert np.array_equal(ctx.bootstrap(k2, gates, c1, c2), ctx.bootstrap(keys, gates, c1, c2))
    with pytest.raises(ValueError):
        k2.import_fpga_bsk(img[:-1])


@pytest.mark.parametrize("N,Q,BksR", [(256, 0, 7), (128, 0x3FFFFFFFFED001, 9)])
def test_fpga_subs_image(N, Q, BksR):
    kw = dict(n=4, q=min(2 * N, 512), N=N, Q=Q, nbit=54 if Q else 27, BGbits=9, Br=16, Bks=32, Bks_R_bits=BksR)
    ctx = fhe.Context(**kw)
    keys = ctx.keys()
    keys.generate_subs(41, ctx.rlwe_secret(41))
    ref = keys.download_subs()
    img = keys.export_fpga_subs()
    # subs_key[i].RLWE_kskey[l] in order, one card buffer each (myfpga.cpp:686-700)
    assert np.array_equal(from_card_buffer(img.reshape(-1, 2 * N), N).reshape(-1), ref)
    k2 = ctx.keys()
    k2.import_fpga_subs(img)
    assert np.array_equal(k2.download_subs(), ref)
    with pytest.raises(ValueError):
        k2.import_fpga_subs(img[1:])
