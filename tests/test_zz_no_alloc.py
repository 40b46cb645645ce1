"""smip_bispec_get_elements_noalloc, the allocation-free form of get_elements
(include/bispectrum.h:482-523), against smip_bispec_get_elements through the C
ABI: bit-identical values and the same errors, on consecutive calls of
decreasing size that share the per-GPU scratch array."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
import smip_amd as S

pytestmark = pytest.mark.gpu

MAX_LOOKUPS = 1 << 16


def _noalloc(b: S.Bispectrum, idx: np.ndarray, fill=np.nan):
    a = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, 4))
    out = np.full(a.shape[0], fill, dtype=np.complex128)
    rc = S.lib().smip_bispec_get_elements_noalloc(b.h, a.ctypes.data_as(C.c_void_p), a.shape[0],
                                                  out.ctypes.data_as(C.c_void_p))
    return rc, out


def _lookups(rng, N, b, n):
    h = b // 2
    ux, uy = rng.integers(-N // 2, N // 2, size=(2, n))
    vx, vy = rng.integers(-h, h + 1, size=(2, n))
    return np.stack([ux, uy, vx, vy], axis=1).astype(np.int32)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_noalloc_matches_get_elements_on_consecutive_calls(ctx, dtype):
    rng = np.random.default_rng(31)
    N, b = 32, 12
    d = O.desc(N, N, b)
    vals = (rng.normal(size=int(d.base_size)) + 1j * rng.normal(size=int(d.base_size))).astype(dtype)
    B = S.Bispectrum(ctx, (N, N, b, b), dtype)
    B.upload(vals)
    for n in (MAX_LOOKUPS, 5000, 37):  # each call smaller than the one before
        idx = _lookups(rng, N, b, n)
        want = B.get_elements(idx)
        rc, got = _noalloc(B, idx)
        assert rc == S.OK, S.lib().smip_last_error()
        assert np.array_equal(got, want)
        if n == MAX_LOOKUPS:  # the aliased (T3/T9) lookups are among them
            assert sum(O.lookup(d, i)[2] for i in idx[:2000]) > 0
    B.close()


def test_noalloc_errors_match_get_elements(ctx):
    rng = np.random.default_rng(5)
    N, b = 16, 8
    B = S.Bispectrum(ctx, (N, N, b, b))
    idx = _lookups(rng, N, b, 300)
    idx[211] = (N, 0, 0, 0)  # outside the stored range: std::out_of_range
    idx[250] = (N, 1, 0, 0)
    with pytest.raises(S.SmipError) as e:
        B.get_elements(idx)
    want_code, want_msg = e.value.code, S.lib().smip_last_error()
    rc, _ = _noalloc(B, idx)
    assert rc == want_code == S.ERR_OUT_OF_RANGE
    assert S.lib().smip_last_error() == want_msg
    assert b"[16, 0, 0, 0]" in want_msg  # the first failing lookup is named
    # more lookups than the scratch holds: refused before anything runs
    rc, _ = _noalloc(B, np.zeros((MAX_LOOKUPS + 1, 4), dtype=np.int32))
    assert rc == S.ERR_INVALID_ARGUMENT
    assert b"at most 65536" in S.lib().smip_last_error()
    # a good call after the failures
    ok = _lookups(rng, N, b, 100)
    rc, got = _noalloc(B, ok)
    assert rc == S.OK and np.array_equal(got, B.get_elements(ok))
    B.close()
