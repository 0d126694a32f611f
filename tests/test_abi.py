"""CPU checks of the C ABI: the library loads, exports every function include/pb_hip.h declares,
and the ctypes structures match the C layouts (compiled here with gcc)."""
import os
import re
import subprocess
import tempfile

import ctypes as C
import pytest

from plabundle_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pb_hip.h")


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:pb_status|void|int|const char\*)\s+(pb_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_abi.LIBRARY_PATH):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "plabundle_amd", "csrc")], check=True)
    return _abi.load_library()


def test_header_functions_are_exported(lib):
    declared = _declared_functions()
    assert set(declared) == set(_abi.EXPORTED_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIBRARY_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(pb_\w+)$", out, re.M))
    assert set(declared) <= exported
    for name in declared:
        assert hasattr(lib, name)


def test_struct_layouts_match_header():
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "pb_hip.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(pb_problem_view), sizeof(pb_solve_options), sizeof(pb_solve_report));
  printf("%zu %zu %zu %zu %zu %zu\n", offsetof(pb_problem_view, gauge_baseline),
         offsetof(pb_solve_options, step_rms_tolerance), offsetof(pb_solve_report, message),
         offsetof(pb_problem_view, cp_matrix), offsetof(pb_problem_view, lp_weight),
         offsetof(pb_solve_options, control_point_huber_delta));
  printf("%zu %zu %zu %zu\n", sizeof(pb_batch_view), offsetof(pb_batch_view, windows), sizeof(pb_window_report),
         offsetof(pb_window_report, device_seconds));
  printf("%zu %zu %zu %zu\n", sizeof(pb_quality_options), offsetof(pb_quality_options, min_accepted_valid_track_ratio),
         sizeof(pb_quality_report), offsetof(pb_quality_report, message));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = [int(x) for x in lines[0].split()]
    offs = [int(x) for x in lines[1].split()]
    assert sizes == [C.sizeof(_abi.ProblemView), C.sizeof(_abi.SolveOptionsC), C.sizeof(_abi.SolveReportC)]
    assert offs == [_abi.ProblemView.gauge_baseline.offset, _abi.SolveOptionsC.step_rms_tolerance.offset,
                    _abi.SolveReportC.message.offset, _abi.ProblemView.cp_matrix.offset,
                    _abi.ProblemView.lp_weight.offset, _abi.SolveOptionsC.control_point_huber_delta.offset]
    batch = [int(x) for x in lines[2].split()]
    assert batch == [C.sizeof(_abi.BatchView), _abi.BatchView.windows.offset, C.sizeof(_abi.WindowReportC),
                     _abi.WindowReportC.device_seconds.offset]
    quality = [int(x) for x in lines[3].split()]
    assert quality == [C.sizeof(_abi.QualityOptionsC), _abi.QualityOptionsC.min_accepted_valid_track_ratio.offset,
                       C.sizeof(_abi.QualityReportC), _abi.QualityReportC.message.offset]


def test_default_options_are_the_reference_defaults(lib):
    from oracle import oracle as O

    o = _abi.default_options(lib)
    ref = O.default_options()
    for name, _ in _abi.SolveOptionsC._fields_:
        a, b = getattr(o, name), getattr(ref, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        elif name not in ("precision",):
            assert a == b, name


def test_invalid_calls_fail_cleanly_without_a_device(lib):
    # no context -> invalid; a missing device is reported, never a crash
    assert lib.pb_hip_upload(None, None) == _abi.PB_ERR_INVALID
    assert lib.pb_hip_solve_batch(None, None, None, None, None, None) == _abi.PB_ERR_INVALID
    assert lib.pb_hip_batch_download(None, 0, None, None, None, None, None) == _abi.PB_ERR_INVALID
    assert lib.pb_hip_destroy(None) == _abi.PB_OK
    ctx = C.c_void_p()
    if not lib.pb_hip_available(0):
        assert lib.pb_hip_create(0, C.byref(ctx)) == _abi.PB_ERR_DEVICE


def test_batch_factor_shapes_and_tuning_queries(lib):
    # the launch shapes of the dense window factorisation: every shape covers the largest window system
    # (threads x rows per thread = 512) with an even panel width no wider than a wave; the tuning query needs a
    # context and a precision index 0 or 1
    t, p = C.c_int32(), C.c_int32()
    shapes = []
    for v in range(_abi.PB_BATCH_FACTOR_SHAPES):
        assert lib.pb_hip_batch_factor_shape(v, C.byref(t), C.byref(p)) == _abi.PB_OK
        assert t.value in (128, 256, 512) and 512 % t.value == 0
        assert p.value % 2 == 0 and 2 <= p.value <= 64
        shapes.append((t.value, p.value))
    assert shapes[0] == (256, 16) and len(set(shapes)) == len(shapes)
    assert lib.pb_hip_batch_factor_shape(_abi.PB_BATCH_FACTOR_SHAPES, C.byref(t), C.byref(p)) == _abi.PB_ERR_INVALID
    assert lib.pb_hip_batch_factor_shape(-1, None, None) == _abi.PB_ERR_INVALID
    shape = C.c_int32()
    assert lib.pb_hip_batch_tuning(None, 1, C.byref(shape), None) == _abi.PB_ERR_INVALID
