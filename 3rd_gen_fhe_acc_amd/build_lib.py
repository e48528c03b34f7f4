"""Build libfheacc.so (HIP, gfx950) in-tree.

hipcc compiles the kernels and the C ABI into
3rd_gen_fhe_acc_amd/lib/libfheacc.so, and the same sources with the
device-side bounds checks (-DFHEACC_DEVICE_CHECKS=1) into
lib/libfheacc_checked.so.  Called by __graft_entry__.build() and usable
directly: python 3rd_gen_fhe_acc_amd/build_lib.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libfheacc.so")
CHECKED_LIB = os.path.join(LIBDIR, "libfheacc_checked.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels_br_u32.hip", "kernels_br_u64.hip", "kernels_cmux_u32.hip", "kernels_cmux_u64.hip",
           "kernels_misc.hip", "api.cpp", "params.cpp", "server.cpp", "noalloc.cpp"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 forms stay off (none of this is floating point anyway)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def build(verbose=False, jobs=8):
    compile_lib(LIB, [], verbose, jobs)
    compile_lib(CHECKED_LIB, ["-DFHEACC_DEVICE_CHECKS=1"], verbose, jobs)
    build_host()
    build_diag()
    return LIB


def compile_lib(lib, defines, verbose=False, jobs=8):
    """Compile SOURCES with `defines` and link them into the shared library `lib`."""
    os.makedirs(LIBDIR, exist_ok=True)
    tag = os.path.splitext(os.path.basename(lib))[0]
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, "%s.%s.o" % (tag, os.path.splitext(src)[0]))
        cmd = [HIPCC] + FLAGS + defines + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait() if procs[0][1].poll() is None else None
            for _, p in procs:
                if p.poll() is None:
                    p.wait()
                    break
    failed = False
    for src, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        errs = [l for l in out.splitlines() if "error" in l or "warning: " in l and "packed-fp32" not in l]
        if p.returncode != 0:
            failed = True
            sys.stderr.write("compile failed: %s\n%s\n" % (src, out))
        elif verbose and errs:
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("%s build failed" % tag)
    tmp = lib + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib)
    for o in objs:
        os.remove(o)
    return lib


def build_host():
    """The reference-shaped C++ host API (libfheacc_host.so) and gate_demo."""
    host = os.path.join(HERE, "host")
    link = ["-L" + LIBDIR, "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    so = os.path.join(LIBDIR, "libfheacc_host.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", so + ".tmp",
                           os.path.join(host, "fhe_acc.cpp"), "-lfheacc"] + link)
    os.replace(so + ".tmp", so)
    for name in ("gate_demo", "rlwe_demo", "tree_demo"):
        demo = os.path.join(LIBDIR, name)
        subprocess.check_call(["g++", "-O2", "-std=c++17", "-Wall", "-o", demo + ".tmp",
                               os.path.join(host, name + ".cpp"), "-lfheacc_host", "-lfheacc"] + link)
        os.replace(demo + ".tmp", demo)


def build_diag():
    """valu_rate: the issue rates of the integer VALU instructions the kernels use."""
    exe = os.path.join(LIBDIR, "valu_rate")
    subprocess.check_call([HIPCC] + FLAGS + ["-x", "hip", os.path.join(CSRC, "valu_rate.hip"), "-o", exe + ".tmp"])
    os.replace(exe + ".tmp", exe)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
