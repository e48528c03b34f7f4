"""The committed counter summaries under profiles/ describe the kernels in the
tree: each records a digest of the kernel sources it was collected on, and
bench.py reports their figures (roofline.traffic, valu) only while that
digest matches.  A kernel change without a new profile run fails here, on the
CPU, instead of silently dropping those figures from the benchmark line."""
import glob
import json
import os

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_counter_summaries_match_the_kernel_sources():
    current = bench.kernel_sources_digest()
    stamped = {}
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.json"))):
        with open(path) as f:
            d = json.load(f)
        if "kernel_sources_sha1" in d:
            stamped[os.path.basename(path)] = d["kernel_sources_sha1"]
    # the headline's SQ counters and HBM traffic, the 54-bit set's likewise
    assert {"sq_counters_blind_rotate_n1024_b4096.json", "traffic_blind_rotate_n1024_b4096.json"} <= set(stamped)
    stale = {k: v for k, v in stamped.items() if v != current}
    assert not stale, "profiles collected on other kernel sources (current %s): %s" % (current, stale)


def test_counter_summaries_give_their_commands():
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.json"))):
        with open(path) as f:
            d = json.load(f)
        assert d.get("commands"), os.path.basename(path)
        for cmd in d["commands"]:
            # committed programs only
            assert "_work" not in cmd, (os.path.basename(path), cmd)
