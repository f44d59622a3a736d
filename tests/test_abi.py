"""The C ABI library loads without a GPU, exports every function include/ptmi.h and
include/ptmi_debug.h declare and nothing else, and its record sizes match the Python mirrors (no
compute calls here)."""
import ctypes as C
import os
import re

from pathtracerglsl_amd import _lib, abi

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
HEADER = os.path.join(INCLUDE, "ptmi.h")
DEBUG_HEADER = os.path.join(INCLUDE, "ptmi_debug.h")


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:[a-z_0-9]+\s+)*[a-z_0-9]+\*?\s+\**(pt_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_parse_finds_entry_points():
    names = declared_functions()
    assert "pt_render_frames" in names and "pt_trace_closest" in names
    assert set(_lib.DECLARED) <= set(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() + declared_functions(DEBUG_HEADER) if not hasattr(lib, n)]
    assert not missing, missing


def test_debug_header_lists_the_instrumentation():
    debug = declared_functions(DEBUG_HEADER)
    assert sorted(_lib.DECLARED_DEBUG) == debug
    assert not set(debug) & set(declared_functions())


def test_library_exports_nothing_undeclared():
    """Every exported pt_ symbol is declared in one of the two headers."""
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.split()[-2:-1] == ["T"] and ln.split()[-1].startswith("pt_")}
    assert exported == set(declared_functions()) | set(declared_functions(DEBUG_HEADER))


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
