"""Gate circuits for the serving loop (fheacc_server_submit).

A circuit is (n_in, ops, outputs):
- node i < n_in is input ciphertext i;
- node n_in + g is the output of gate g;
- ops[g] = (gate, src1, src2) with both operands earlier nodes;
- outputs lists the nodes returned.

Gate numbers follow the reference's enum GATES {OR, AND, NOR, NAND, XOR, XNOR}
(FHE_ACC_runtime/include/params.h:107).
"""
import numpy as np

OR, AND, NOR, NAND, XOR, XNOR = range(6)


def equality_circuit(m, k):
    """Set equality of m k-bit items against one k-bit query.

    Inputs are the items' bits, item-major and low bit first (node i*k + b),
    then the query's bits (node m*k + b). Each item is compared bitwise with
    XNOR, then the k results are reduced with a balanced AND tree. Outputs:
    one encrypted bit per item, 1 when the item equals the query.
    Gates: m*k XNOR + m*(k-1) AND."""
    if m < 1 or k < 1:
        raise ValueError("m and k must be positive")
    n_in = (m + 1) * k
    ops = []
    outputs = []
    for i in range(m):
        level = []
        for b in range(k):
            ops.append((XNOR, i * k + b, m * k + b))
            level.append(n_in + len(ops) - 1)
        while len(level) > 1:
            nxt = []
            for t in range(0, len(level) - 1, 2):
                ops.append((AND, level[t], level[t + 1]))
                nxt.append(n_in + len(ops) - 1)
            if len(level) % 2:
                nxt.append(level[-1])
            level = nxt
        outputs.append(level[0])
    return n_in, np.array(ops, dtype=np.int32).reshape(-1, 3), np.array(outputs, dtype=np.int64)


def equality_inputs(items, query, k):
    """Plaintext input bits of equality_circuit: items' low k bits, then the query's."""
    items = np.asarray(items, dtype=np.uint64)
    bits = [(int(v) >> b) & 1 for v in items for b in range(k)]
    bits += [(int(query) >> b) & 1 for b in range(k)]
    return np.array(bits, dtype=np.uint8)


def evaluate_plain(n_in, ops, outputs, bits):
    """Evaluate a circuit on plaintext bits (the reference's gate truth tables)."""
    table = {OR: lambda a, b: a | b, AND: lambda a, b: a & b, NOR: lambda a, b: 1 - (a | b),
             NAND: lambda a, b: 1 - (a & b), XOR: lambda a, b: a ^ b, XNOR: lambda a, b: 1 - (a ^ b)}
    val = list(int(x) for x in bits[:n_in])
    for g, a, b in np.asarray(ops).reshape(-1, 3):
        val.append(table[int(g)](val[a], val[b]))
    return np.array([val[o] for o in outputs], dtype=np.uint8)


def circuit_depth(n_in, ops):
    """Longest chain of gates (bootstrap launches a job needs at least)."""
    depth = [0] * n_in
    for g, a, b in np.asarray(ops).reshape(-1, 3):
        depth.append(1 + max(depth[a], depth[b]))
    return max(depth[n_in:], default=0)
