"""The C-ABI library loads and exports every symbol include/rsr_amd.h declares (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rsr_amd.h")
LIB = os.path.join(ROOT, "rsr_amd", "librsr_amd.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"RSR_API\s+[\w\s\*]+?\b(rsr_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ["rsr_context_create", "rsr_dataset_create", "rsr_programs_create", "rsr_eval_mse",
                 "rsr_eval_mse_grad", "rsr_eval_points", "rsr_eval_points_dual"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build librsr_amd.so first (python -m rsr_amd.build)"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (rsr_\w+)", out))
    assert set(declared()) <= exported


def test_error_strings_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.rsr_error_string.restype = ctypes.c_char_p
    assert lib.rsr_error_string(0) == b"ok"
    assert lib.rsr_error_string(-4) == b"HBM budget exceeded"
    lib.rsr_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.rsr_version()


def test_no_cpu_fallback_in_product_bindings():
    """The product bindings never import the oracle."""
    for dirpath, _, files in os.walk(os.path.join(ROOT, "rsr_amd")):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "from oracle" not in src and "liboracle" not in src, f


def test_host_entry_points_reject_bad_arguments_without_gpu():
    """Host-only entry points return RSR_E_INVALID_ARG (-1) for null or inconsistent
    arguments instead of crashing, and name the problem."""
    lib = ctypes.CDLL(LIB)
    lib.rsr_last_error_message.restype = ctypes.c_char_p
    i64 = ctypes.c_int64
    # simplification: null offsets
    assert lib.rsr_simplify_programs(1, None, None, None, None, None, None, None, i64(0), i64(0), i64(0),
                                     None, None, None, None, None, None, None) == -1
    # complexity: null arrays for a non-empty batch
    assert lib.rsr_complexity(2, None, None, None) == -1
    # file readers: a missing file, and a null count
    n = i64()
    assert lib.rsr_read_csv(b"/nonexistent/none.csv", None, 0, -1, None, None, i64(0), ctypes.byref(n)) == -1
    assert b"cannot read" in lib.rsr_last_error_message()
    assert lib.rsr_read_json(b"/nonexistent/none.json", None, 0, None, None, i64(0), None) == -1
    # fitness: null arrays
    assert lib.rsr_fitness(3, None, None, ctypes.c_float(5e-6), None) == -1
