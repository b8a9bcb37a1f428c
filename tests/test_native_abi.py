"""The C-ABI library loads and exports every entry point include/palabos_amd.h
declares (no device calls: these run without a GPU)."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "palabos_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(plbamd_\w+)\s*\(", text, re.M)))


def test_header_declares_the_binding_table():
    from palabos_amd import native
    assert declared_symbols() == sorted(native.EXPORTS)


def test_library_exports_every_declared_symbol():
    from palabos_amd import native
    lib = native.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (plbamd_\w+)", out))
    for name in declared_symbols():
        assert name in exported, name
        getattr(lib, name)


def test_unknown_handle_is_rejected_with_a_message():
    # Handle lookup happens before any device call, so this runs without a GPU.
    from palabos_amd import native
    lib = native.lib()
    assert lib.plbamd_lattice_destroy(123456) == native.EINVAL
    assert b"unknown handle" in lib.plbamd_last_error()


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from palabos_amd import native
    monkeypatch.setattr(native, "_lib", None)
    monkeypatch.setattr(native, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(native.NativeLibraryError):
        native.lib()


def test_bench_refuses_more_than_one_gpu():
    # The benchmark has a single-GPU path only; asking for more must end before
    # any device is touched, with a message that says why.
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_bandwidth_probe_rejects_bad_arguments_before_the_device():
    # Argument checks come before any device call, so this runs without a GPU.
    import ctypes
    from palabos_amd import native
    lib = native.lib()
    ms, moved = ctypes.c_double(), ctypes.c_double()
    assert lib.plbamd_bandwidth_probe(9, 1 << 20, 1, ctypes.byref(ms), ctypes.byref(moved)) == native.EINVAL
    assert b"unknown pattern" in lib.plbamd_last_error()
    assert lib.plbamd_bandwidth_probe(0, 1 << 20, 0, ctypes.byref(ms), ctypes.byref(moved)) == native.EINVAL
    assert lib.plbamd_bandwidth_probe(1, 100, 1, ctypes.byref(ms), ctypes.byref(moved)) == native.EINVAL
    assert b"size out of range" in lib.plbamd_last_error()
