"""Gate circuits of the serving loop, on plaintext bits (CPU)."""
import numpy as np
import pytest

from fheacc_testlib import fhe  # noqa: F401  (puts the repository root on sys.path)
import importlib

circuits = importlib.import_module("3rd_gen_fhe_acc_amd.circuits")


@pytest.mark.parametrize("k", [1, 3, 8, 16, 32])
@pytest.mark.parametrize("m", [1, 2, 7])
def test_equality_circuit_shape(m, k):
    n_in, ops, outs = circuits.equality_circuit(m, k)
    assert n_in == (m + 1) * k
    assert ops.shape == (m * (2 * k - 1), 3)
    assert outs.shape == (m,)
    # operands are earlier, distinct nodes
    for g, (gate, a, b) in enumerate(ops):
        assert 0 <= a < n_in + g and 0 <= b < n_in + g and a != b
        assert gate in (circuits.XNOR, circuits.AND)
    # balanced tree: one XNOR level then ceil(log2 k) AND levels
    assert circuits.circuit_depth(n_in, ops) == 1 + int(np.ceil(np.log2(k)))


@pytest.mark.parametrize("k", [8, 16, 32])
def test_equality_circuit_truth(k):
    rng = np.random.default_rng(k)
    m = 40
    items = rng.integers(0, 2 ** 32, m, dtype=np.uint64)
    query = int(rng.integers(0, 2 ** 32, dtype=np.uint64))
    mask = (1 << k) - 1
    # plant matches in the low k bits, and near misses one bit away
    items[3] = (int(items[3]) & ~mask) | (query & mask)
    items[11] = (int(items[11]) & ~mask) | (query & mask)
    items[12] = (int(items[12]) & ~mask) | ((query ^ (1 << (k - 1))) & mask)
    n_in, ops, outs = circuits.equality_circuit(m, k)
    bits = circuits.equality_inputs(items, query, k)
    got = circuits.evaluate_plain(n_in, ops, outs, bits)
    want = np.array([(int(v) & mask) == (query & mask) for v in items], dtype=np.uint8)
    assert np.array_equal(got, want)
    assert got[3] == 1 and got[11] == 1 and got[12] == 0


def test_all_gate_truth_tables():
    # each gate on each input pair (params.h:107 order)
    table = {0: [0, 1, 1, 1], 1: [0, 0, 0, 1], 2: [1, 0, 0, 0], 3: [1, 1, 1, 0], 4: [0, 1, 1, 0],
             5: [1, 0, 0, 1]}
    for g, want in table.items():
        ops = np.array([[g, 0, 1]], dtype=np.int32)
        got = [circuits.evaluate_plain(2, ops, [2], [a, b])[0] for a in (0, 1) for b in (0, 1)]
        assert got == want
