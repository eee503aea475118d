"""C-ABI checks that need no GPU: the library loads, exports every symbol include/gprt_hip.h and
include/gprt_hip_ext.h declare, and validates arguments before any HIP call (reference error convention:
gprt/gprt.cpp:247-251 prints "#gprt error:" and raises SIGINT; tests use the recording mode)."""
import ctypes as C
import os
import re

import pytest

from gprt_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = "".join(open(os.path.join(ROOT, "include", h)).read() for h in ("gprt_hip.h", "gprt_hip_ext.h"))
    return sorted(set(re.findall(r"GPRT_API\s+[\w\s\*]+?\b(gprt\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ["gprtContextCreate", "gprtBufferSort", "gprtBufferSortPayload", "gprtBufferExclusiveSum",
                 "gprtBufferPartition", "gprtBufferSelect", "gprtTriangleAccelCreate", "gprtAccelBuild",
                 "gprtAccelUpdate", "gprtAccelTraceClosestHit", "gprtBeginProfile", "gprtEndProfile"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(api.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert sorted(api.symbols()) == _declared()


@pytest.fixture()
def rec():
    lib = api.load()
    lib.gprtSetErrorMode(api.GPRT_ERROR_RECORD)
    lib.gprtClearLastError()
    yield lib
    lib.gprtClearLastError()


def test_multi_device_context_rejected(rec):
    ids = (C.c_int32 * 2)(0, 1)
    assert not rec.gprtContextCreate(ids, 2)
    assert "numDevices=2" in api.last_error()


def test_null_arguments_rejected(rec):
    rec.gprtBufferSort(None, None, None)
    assert "null argument" in api.last_error()
    rec.gprtClearLastError()
    rec.gprtBufferSortPayload(None, None, None, None)
    assert "null argument" in api.last_error()
    rec.gprtClearLastError()
    assert rec.gprtBufferExclusiveSum(None, None, None, None) == 0xFFFFFFFF
    assert "null argument" in api.last_error()
    rec.gprtClearLastError()
    rec.gprtAccelBuild(None, None, api.GPRTBuildParams())
    assert "null argument" in api.last_error()
    rec.gprtClearLastError()
    rec.gprtAccelTraceClosestHit(None, None, None, None, 0)
    assert "null argument" in api.last_error()


def test_build_params_layout():
    # GPRTBuildParams is passed by value: enum + five bools, 12 bytes with padding.
    assert C.sizeof(api.GPRTBuildParams) == 12
    assert api.GPRTBuildParams().buildMode == api.GPRT_BUILD_MODE_FAST_TRACE_NO_UPDATE


def test_version_string():
    assert b"gfx950" in api.load().gprtVersion()


def test_instance_record_layout(tmp_path):
    """GPRTInstance as a C compiler lays it out equals api.INSTANCE_DTYPE / api.GPRTInstance
    (gprt::Instance, gprt/gprt_shared.h:92-101: 48 B transform, two bitfield words, address)."""
    import shutil
    import subprocess
    import numpy as np
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <string.h>\n#include <stddef.h>\n#include "gprt_hip.h"\n'
        "int main(void) { GPRTInstance x; memset(&x, 0, sizeof x);\n"
        "  x.transform[1][3] = 2.5f; x.instanceCustomIndex = 0x123456; x.mask = 0xAB;\n"
        "  x.geometryIndex = 7; x.flags = 3; x.accelAddress = 0x1122334455667788ull;\n"
        "  fwrite(&x, sizeof x, 1, stdout); return 0; }\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    raw = subprocess.run([str(exe)], check=True, capture_output=True).stdout
    assert len(raw) == 64 == api.INSTANCE_DTYPE.itemsize == C.sizeof(api.GPRTInstance)
    rec = np.frombuffer(raw, dtype=api.INSTANCE_DTYPE)[0]
    assert rec["transform"][1, 3] == 2.5
    assert rec["word0"] == 0xAB123456 and rec["word1"] == 0x03000007
    assert rec["accelAddress"] == 0x1122334455667788
    c = api.GPRTInstance.from_buffer_copy(raw)
    assert (c.instanceCustomIndex, c.mask, c.geometryIndex, c.flags) == (0x123456, 0xAB, 7, 3)


def test_width_explicit_sorts_validate(rec):
    for name in ("gprtBufferSort32", "gprtBufferSort64"):
        getattr(rec, name)(None, None, None)
        assert "null argument" in api.last_error()
        rec.gprtClearLastError()
    for name in ("gprtBufferSortPayload32", "gprtBufferSortPayload64"):
        getattr(rec, name)(None, None, None, None)
        assert "null argument" in api.last_error()
        rec.gprtClearLastError()


def test_extension_header_compiles_as_c(tmp_path):
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    src = tmp_path / "ext.c"
    src.write_text('#include "gprt_hip_ext.h"\nvoid (*f)(GPRTContext, GPRTBuffer, GPRTBuffer) = gprtBufferSort32;\nint main(void) { return f != gprtBufferSort32; }\n')
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    str(src)], check=True)


def test_compute_entry_points_validate(rec):
    g = (C.c_uint32 * 3)(1, 1, 1)
    assert not rec.gprtModuleCreate(None, None, 0)
    assert "null argument" in api.last_error()
    rec.gprtClearLastError()
    assert not rec.gprtComputeCreate(None, None, b"x")
    assert "null argument" in api.last_error()
    rec.gprtClearLastError()
    assert rec.gprtComputeLaunchArgs(None, g, g, None, 0) == 0
    assert "null argument" in api.last_error()


def test_cpp_argument_packing_matches_python(tmp_path):
    """gprtPackArgs (C++ header) and api.pack_args lay arguments out identically: natural alignment,
    in order (a uint32 before a pointer leaves 4 bytes of padding, as a HIP kernel's argument
    block does); and an argument list over 256 bytes does not compile."""
    import shutil
    import subprocess
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    src = tmp_path / "pack.cpp"
    src.write_text(
        '#include <cstdio>\n#include "gprt_hip.h"\n'
        "struct Rec { void *a; float b; };\n"
        "int main() { alignas(16) unsigned char blk[256] = {};\n"
        "  Rec r{(void *)0x1122334455667788ull, 2.5f};\n"
        "  size_t n = gprtPackArgs(blk, (uint32_t)7, (void *)0x0102030405060708ull, (uint16_t)9, 1.5, (uint8_t)3, r);\n"
        "  static_assert(gprtPackedSize<uint32_t, void *, uint16_t, double, uint8_t, Rec>() == 56, \"size\");\n"
        "  fwrite(blk, 1, n, stdout); return 0; }\n")
    exe = tmp_path / "pack"
    inc = os.path.join(ROOT, "include")
    subprocess.run(["g++", "-std=c++17", "-I", inc, str(src), "-o", str(exe)], check=True)
    raw = subprocess.run([str(exe)], check=True, capture_output=True).stdout

    class Rec(C.Structure):
        _fields_ = [("a", C.c_void_p), ("b", C.c_float)]

    want = api.pack_args(C.c_uint32(7), C.c_void_p(0x0102030405060708), C.c_uint16(9), C.c_double(1.5),
                         C.c_uint8(3), Rec(0x1122334455667788, 2.5))
    assert len(raw) == 56 and raw == want
    big = tmp_path / "big.cpp"
    big.write_text('#include "gprt_hip.h"\nstruct B { char c[250]; };\n'
                   "int main() { B b{}; return (int)gprtComputeLaunch((GPRTCompute)0, {1, 1, 1}, {1, 1, 1}, 1.0, b); }\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", inc, str(big)], capture_output=True, text=True)
    assert r.returncode != 0 and "PUSH_CONSTANTS_LIMIT" in r.stderr


def test_sync_timeout_validated(rec):
    rec.gprtSetSyncTimeout(None, 1.0)
    assert "null context" in api.last_error()
