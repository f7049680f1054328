"""The C-ABI library loads and exports every entry point include/ffhip.h declares (no GPU
calls), and its host-side argument checks reject bad input without touching the device."""
import ctypes
import os
import re

import pytest

from finflo_amd import ffhip
from finflo_amd.layout import FIELDS, Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ffhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ff_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ffhip.lib()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"
    assert set(names) == set(ffhip.EXPORTS)


def test_field_enum_matches_header():
    src = open(os.path.join(ROOT, "include", "ffhip.h")).read()
    body = src[src.index("enum ff_field"):src.index("FF_NFIELDS")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"FF_([A-Z0-9]+)", body)
    assert names == FIELDS


def test_params_struct_size():
    # 30 doubles + 16 ints, no padding
    assert ctypes.sizeof(Params) == 30 * 8 + 16 * 4


def test_null_arguments_rejected():
    lib = ffhip.lib()
    assert lib.ff_ctx_create(None, None) == -1
    assert lib.ff_state(None, 1, 1) == -1
    assert lib.ff_residual(None, 1, 1, 1, None) == -1
    assert lib.ff_cycle(None, 1, 1, None) == -1
    assert lib.ff_last_error(None) == b"null context"
