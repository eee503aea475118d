"""C-ABI checks that need no GPU: the library loads, exports every symbol include/gprt_hip.h
declares, and validates arguments before any HIP call (reference error convention:
gprt/gprt.cpp:247-251 prints "#gprt error:" and raises SIGINT; tests use the recording mode)."""
import ctypes as C
import os
import re

import pytest

from gprt_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "gprt_hip.h")).read()
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
