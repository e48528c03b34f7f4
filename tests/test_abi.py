"""CPU checks of the C-ABI library: it loads, exports every symbol
include/fheacc.h declares, and rejects invalid parameter sets before touching
a device.  No compute calls (there is no GPU here)."""
import ctypes as C
import os
import re

import pytest

from fheacc_testlib import ROOT, fhe


def declared_symbols():
    with open(os.path.join(ROOT, "include", "fheacc.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(fheacc_\w+)\s*\(", text, re.M)))


def test_header_declares_the_path():
    syms = declared_symbols()
    for s in ("fheacc_ctx_create", "fheacc_bootstrap_batch", "fheacc_bootstrap_batch_dev", "fheacc_ntt_dev",
              "fheacc_external_product_dev", "fheacc_subs_keyswitch_dev", "fheacc_keys_upload_bsk"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    L = C.CDLL(fhe.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # the binding's table covers the header too
    assert set(declared_symbols()) == set(fhe.EXPORTS)


@pytest.mark.parametrize("bad", [
    dict(n=16, q=512, N=300),            # N not a power of two
    dict(n=16, q=4096, N=1024),          # q does not divide 2N
    dict(n=16, q=512, N=256, Q=97),      # Q != 1 mod 2N
    dict(n=16, q=512, N=256, Bks=1),     # key-switch base too small
    dict(n=16, q=512, N=256, Bks=300),   # key-switch digits must fit a byte
])
def test_invalid_params_rejected(bad):
    p = fhe._Params(bad.get("n", 16), bad.get("q", 512), bad.get("N", 256), bad.get("Q", 0), 27, 7,
                    23, bad.get("Bks", 25), 9, 3.19, 18)
    h = C.c_void_p()
    rc = fhe.lib().fheacc_ctx_create(C.byref(p), 0, C.byref(h))
    assert rc == -1
    assert fhe.lib().fheacc_last_error()


def test_no_cpu_fallback(monkeypatch, tmp_path):
    # a missing library is an error, never a silent fallback
    monkeypatch.setattr(fhe, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(fhe, "_lib", None)
    with pytest.raises(fhe.FheAccError):
        fhe.lib()


def test_checked_library_loads_with_the_same_symbols():
    # the device-side bounds-check build exports the same C ABI, and says it
    # is checked: each kernel translation unit registers its violation word
    # while the library loads (no device call)
    lib_dir = os.path.join(ROOT, "3rd_gen_fhe_acc_amd", "lib")
    plain = C.CDLL(os.path.join(lib_dir, "libfheacc.so"))
    checked = C.CDLL(os.path.join(lib_dir, "libfheacc_checked.so"))
    missing = [s for s in declared_symbols() if not hasattr(checked, s)]
    assert not missing, missing
    assert plain.fheacc_device_checks_enabled() == 0
    assert checked.fheacc_device_checks_enabled() == 1
    # a null output is refused before any device is touched
    assert checked.fheacc_device_check_read(0, None) == -1
