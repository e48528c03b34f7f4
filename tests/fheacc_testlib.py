"""Shared helpers for the tests: importing the package (its directory name
starts with a digit) and the parameter sets used for parity."""
import importlib
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

fhe = importlib.import_module("3rd_gen_fhe_acc_amd")

# name -> (Context/Oracle keyword arguments)
SETS = {
    # narrow words, general Bks and Br (the reference's CPU-app style, N reduced)
    "n16_N256_q27": dict(n=16, q=512, N=256, Q=0, nbit=27, BGbits=9, Br=23, Bks=25),
    # the benchmark's parameter shape (BG=2^7 d=4, Br=32, Bks=2^5), N and n reduced
    "n12_N512_q27": dict(n=12, q=1024, N=512, Q=0, nbit=27, BGbits=7, Br=32, Bks=32),
    # wide words (54-bit Q), BG=2^18 d=3
    "n10_N128_q54": dict(n=10, q=256, N=128, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=16, Bks=32),
    # ring degree 2048, 54-bit Q (mixed-set B shape, n reduced)
    "n8_N2048_q54": dict(n=8, q=1024, N=2048, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=32, Bks=32),
    # the benchmark's N=1024 shape, n reduced (every launch shape, 1024 threads included)
    "n12_N1024_q27": dict(n=12, q=1024, N=1024, Q=0, nbit=27, BGbits=7, Br=32, Bks=32),
    # ring degree 512, 27-bit Q, BG=2^6 d=5 (mixed-set C shape, n reduced)
    "n8_N512_q27_bg6": dict(n=8, q=1024, N=512, Q=0, nbit=27, BGbits=6, Br=32, Bks=32),
    # 29-bit Q at N = 2048, BG=2^10 d=3, Bks=25: the reference's 256-bit app
    # shape (apps/bootstrap_fpga_256.cpp), n reduced; 32-bit words with
    # Harvey's butterflies
    "n8_N2048_q29": dict(n=8, q=1024, N=2048, Q=0, nbit=29, BGbits=10, Br=32, Bks=25),
    # 30-bit Q at N = 1024, at the 4Q < 2^32 edge of 32-bit words
    "n8_N1024_q30": dict(n=8, q=1024, N=1024, Q=0, nbit=30, BGbits=10, Br=32, Bks=32),
    # 28-bit Q below 2^32 / 21 at N = 1024: fully lazy 32-bit forward
    # butterflies, but 16 Q > 2^32 / 2, so the inverse transform's 3-stage
    # passes leave their sums unreduced and its 4-stage pass does not
    "n8_N1024_q28": dict(n=8, q=1024, N=1024, Q=0xBEB8801, nbit=28, BGbits=7, Br=32, Bks=32),
}

# full-size configurations
CFG0 = dict(n=500, q=1024, N=1024, Q=0, nbit=27, BGbits=7, Br=32, Bks=32)
CFG1 = dict(n=8, q=2048, N=1024, Q=0x3FFFFFFFFED001, nbit=54, BGbits=9, Br=32, Bks=32)

# the mixed batch: A = the benchmark set; B = N=2048, 54-bit Q, d=3, n=630;
# C = N=512, 27-bit Q, d=5
MIXED = {
    "A": CFG0,
    "B": dict(n=630, q=1024, N=2048, Q=0x3FFFFFFFFED001, nbit=54, BGbits=18, Br=32, Bks=32),
    "C": dict(n=500, q=1024, N=512, Q=0, nbit=27, BGbits=6, Br=32, Bks=32),
}
MIXED_SHARE = {"A": 0.6, "B": 0.3, "C": 0.1}
