"""The C ABI library loads without a GPU, exports every function include/ptmi.h declares,
and its record sizes match the Python mirrors (no compute calls here)."""
import ctypes as C
import os
import re

from pathtracerglsl_amd import _lib, abi

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ptmi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\*?\s+\**(pt_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_parse_finds_entry_points():
    names = declared_functions()
    assert "pt_render_frames" in names and "pt_trace_closest" in names
    assert set(_lib.DECLARED) <= set(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_record_sizes_match_mirrors():
    lib = _lib.load()
    sizes = [lib.pt_abi_sizeof(i) for i in range(8)]
    assert sizes[0] == abi.OBJECT_DTYPE.itemsize
    assert sizes[1] == abi.MATERIAL_DTYPE.itemsize
    assert sizes[2] == abi.LIGHT_DTYPE.itemsize
    assert sizes[3] == C.sizeof(abi.PtEnv)
    assert sizes[4] == C.sizeof(abi.PtSceneDesc)
    assert sizes[5] == C.sizeof(abi.PtProgram)
    assert sizes[6] == C.sizeof(abi.PtParams)
    assert sizes[7] == C.sizeof(abi.PtInstTable)


def test_null_arguments_are_errors_not_crashes():
    lib = _lib.load()
    assert lib.pt_ctx_create(0, None) == 1
    assert b"null" in lib.pt_last_error()
    assert lib.pt_scene_upload(None, None) == 1
    assert lib.pt_render_frames(None, None, 1) == 1
