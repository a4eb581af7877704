"""The C-ABI library loads without a GPU and exports every function include/aero_hip.h declares."""
import ctypes as C
import re
from pathlib import Path

import aero_hip
from aero_hip import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "aero_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void\s*\*?|void)\s+\**\s*(aero_\w+)\s*\(", text, flags=re.M)))


def test_header_symbols_exported():
    names = declared_functions()
    assert len(names) >= 15, names
    lib = C.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes prototype in the Python layer
    assert set(names) <= set(_lib.PROTOTYPES), set(names) - set(_lib.PROTOTYPES)


def test_struct_layouts_match_header_order():
    # spot-check the field order of the largest struct against the header text
    text = HEADER.read_text()
    body = text[text.index("typedef struct aero_case_config"): text.index("} aero_case_config;")]
    decl = re.findall(r"(?:int32_t|double)\s+([^;]+);", body)
    names = [n.strip() for d in decl for n in d.split(",")]
    assert names == [n for n, _ in _lib.CASE_FIELDS]


def test_available_reports_reason_without_gpu():
    err = C.create_string_buffer(256)
    ok = _lib.lib().aero_hip_available(err, len(err))
    if not ok:
        assert err.value, "aero_hip_available must say why no device is usable"


def test_mesh_create_rejects_bad_input():
    err = C.create_string_buffer(256)
    h = C.c_void_p()
    assert _lib.lib().aero_hip_mesh_create(None, None, C.byref(h), err, len(err)) == _lib.AERO_EINVAL
    assert b"null" in err.value
