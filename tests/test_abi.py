"""CPU checks of the C ABI boundary: the engine library loads and exports every
function include/*.h declares, and the host-only entry points behave like the
reference.  No kernel is launched here."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import oracle as O
import smip_amd as S

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"SMIP_API\s+[\w\s\*]+?\b(smip_\w+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared()
    assert len(names) > 40
    L = C.CDLL(S.library_path())
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_the_abi():
    assert _declared() == set(S._SIGS)


def test_next_reco_index_matches_oracle():
    """Host restatement of NextRecoIndex (src/phasereco.cpp:17-36) in the product."""
    st_p = (0.0, 0.0, 0, 0)
    st_o = (0.0, 0.0, 0, 0)
    for _ in range(20000):
        st_p = S.next_reco_index(*st_p)
        st_o = O.next_reco_index(*st_o)
        assert st_p == st_o


def test_window_matches_oracle():
    for (x, y, a, al) in [(16, 16, 20.0, 0.5), (9, 7, 6.0, 0.54), (64, 32, 128.0, 0.5)]:
        assert np.array_equal(S.general_hamming(x, y, a, al), O.general_hamming(x, y, a, al))
    with pytest.raises(S.SmipError) as e:
        S.general_hamming(4, 4, -1.0, 0.5)
    assert e.value.code == S.ERR_DOMAIN


def test_job_footprint_counts_the_bispectrum():
    p = S.job_params(256, 256, 48, 128, np.uint16, S.PRECISION_MIXED)
    fp = S.job_footprint(p)
    assert fp > 40_612_425 * 16  # the FP64 fold buffer alone
    assert fp < 4 * 2 ** 30
    # any side FFTW takes, up to the 2048 the engine's transforms hold in LDS
    assert S.job_footprint(S.job_params(100, 90, 8, 8, np.uint8)) > 0
    assert S.job_footprint(S.job_params(2, 2039, 2, 2, np.uint8)) > 0
    for ncols, nrows in ((1, 64), (64, 2049), (4096, 4096)):
        with pytest.raises(S.SmipError) as e:
            S.job_footprint(S.job_params(ncols, nrows, 8, 8, np.uint8))
        assert e.value.code == S.ERR_INVALID_ARGUMENT
    # bispectrum depth up to 255: phase recovery's (2(b/2)+1)^2 <= 65536 candidates
    assert S.job_footprint(S.job_params(64, 64, 255, 8, np.uint8)) > 0
    with pytest.raises(S.SmipError) as e:
        S.job_footprint(S.job_params(64, 64, 256, 8, np.uint8))
    assert e.value.code == S.ERR_INVALID_ARGUMENT


def test_errors_without_device_are_reported():
    with pytest.raises(S.SmipError):
        S.lib()  # loads
        S.Context(10_000)
