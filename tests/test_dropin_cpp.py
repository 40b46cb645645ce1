"""The drop-in C++ surface (speckle-stacker_amd/smip/*.h) exercised through
tests/cpp/dropin_test.cpp, whose cases follow the reference's own test programs
(tests/array2_test.cpp, tests/range_test.cpp, tests/bispectrum_test.cpp,
tests/correl_test.cpp).  CPU cases run everywhere; GPU cases drive the device
through libsmip_mi355x.so."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "speckle-stacker_amd", "lib", "smip_dropin_test")


def _binary():
    if not os.path.exists(BIN):
        import sys
        sys.path.insert(0, os.path.join(REPO, "speckle-stacker_amd"))
        import build
        build.build_cli()
    return BIN


def _cases(kind):
    out = subprocess.run([_binary(), "list"], check=True, capture_output=True, text=True).stdout
    return [l.split()[0] for l in out.splitlines() if l.split()[1] == kind]


def _run(case):
    r = subprocess.run([_binary(), case], capture_output=True, text=True, timeout=300, cwd=REPO)
    assert r.returncode == 0, r.stdout + r.stderr


CPU_CASES = ["array2_arithmetic", "array2_bounds", "array2_construction", "array2_convert_import",
             "array2_copy_move_resize", "array2_fill_index", "array2_ranges", "array2_reference_assignment",
             "array2_shift", "next_reco_index_spiral", "range_contains", "window_functions"]
GPU_CASES = ["bispectrum_accumulate_from_fft", "bispectrum_arithmetic", "bispectrum_construction",
             "bispectrum_element_access", "bispectrum_typed", "cross_correlation_shifts",
             "phase_recovery_stepwise"]


def test_case_lists_are_complete():
    assert sorted(_cases("cpu")) == sorted(CPU_CASES)
    assert sorted(_cases("gpu")) == sorted(GPU_CASES)


@pytest.mark.parametrize("case", CPU_CASES)
def test_dropin_cpu(case):
    _run(case)


@pytest.mark.gpu
@pytest.mark.parametrize("case", GPU_CASES)
def test_dropin_gpu(case):
    _run(case)
