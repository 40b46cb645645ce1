"""Pins the CPU oracle (oracle/smip_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a dropped
conjugate, a swapped index, a wrong sign or wrap) fails it: the reference's own
tests, brute force over explicit enumerations, closed forms and invariants.
"""
import numpy as np
import pytest

import oracle as O


# --------------------------------------------------------------------------
# reference tests: tests/bispectrum_test.cpp, tests/array2_test.cpp
# --------------------------------------------------------------------------
def test_construction_sizes():
    """tests/bispectrum_test.cpp:27-42"""
    d = O.desc(2, 3, 4, 5)
    assert list(d.sizes) == [3, 3, 5, 5]
    assert d.totalsize == 3 * 3 * 5 * 5
    assert d.base_size > 0
    B = O.new_bispectrum(d)
    assert O.get_element(d, B, (0, 0, 0, 0)) == 0


def test_element_access_roundtrip():
    """tests/bispectrum_test.cpp:44-62: put/get and calc_indices <-> get_element."""
    d = O.desc(3, 4, 5, 6)
    B = O.new_bispectrum(d)
    O.put_element(d, B, (-1, 1, -1, 1), 42 + 1j)
    assert O.get_element(d, B, (-1, 1, -1, 1)) == 42 + 1j
    for a in range(int(d.base_size)):
        B[a] = a
        assert O.get_element(d, B, O.calc_indices(d, a)) == a


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_complex_conjugation(dtype):
    """tests/bispectrum_test.cpp:64-71 (T7)."""
    d = O.desc(4, 4, 4, 4)
    B = O.new_bispectrum(d, dtype)
    O.put_element(d, B, (-1, 2, -1, 1), 1 + 2j)
    assert O.get_element(d, B, (1, -2, 1, -1)) == np.conj(O.get_element(d, B, (-1, 2, -1, 1)))
    assert O.get_element(d, B, (1, -2, 1, -1)) == 1 - 2j


def test_zero_size_throws():
    """tests/bispectrum_test.cpp:73-79: empty bispectrum -> std::out_of_range."""
    with pytest.raises(O.OracleError) as e:
        O.get_element(O.empty_desc(), np.zeros(1, np.complex128), (0, 0, 0, 0))
    assert e.value.code == O.OUT_OF_RANGE


def test_multiple_get():
    """tests/bispectrum_test.cpp:106-119"""
    d = O.desc(2, 2, 2, 2)
    B = O.new_bispectrum(d)
    O.put_element(d, B, (0, 0, 0, 0), 1.5 + 1j)
    O.put_element(d, B, (-1, 1, -1, 1), 2.5 - 1j)
    assert O.get_element(d, B, (0, 0, 0, 0)) == 1.5 + 1j
    assert O.get_element(d, B, (-1, 1, -1, 1)) == 2.5 - 1j


def test_shift():
    """tests/array2_test.cpp:329-345: shifted({-1, 1}) of [[1,2,3],[4,5,6],[7,8,9]]."""
    a = np.arange(1, 10, dtype=float).reshape(3, 3)
    s = O.shifted(a, -1, 1)
    assert s[1].tolist() == [2, 3, 0]
    assert s[:, 1].tolist() == [0, 3, 6]


# --------------------------------------------------------------------------
# descriptor and layout
# --------------------------------------------------------------------------
@pytest.mark.parametrize("dims", [(128, 128, 32, 32), (256, 256, 48, 48), (512, 512, 64, 64), (7, 9, 3, 5)])
def test_descriptor_closed_form(dims):
    """sizes = 2 floor(d/2) + 1; base = [s0 - s0/2, s1, s2 - s2/2, s3]."""
    d = O.desc(*dims)
    s = [2 * (x // 2) + 1 for x in dims]
    assert list(d.sizes) == s
    assert list(d.base_sizes) == [s[0] - s[0] // 2, s[1], s[2] - s[2] // 2, s[3]]
    assert list(d.min_idx) == [-(x // 2) for x in s]
    assert list(d.max_idx) == [x // 2 for x in s]


def test_known_base_sizes():
    """Cell counts of the three frame sizes the engine is measured at."""
    assert O.desc(128, 128, 32).base_size == 4_703_985
    assert O.desc(256, 256, 48).base_size == 40_612_425
    assert O.desc(512, 512, 64).base_size == 282_798_945


def test_calc_offset_wraps_modulo_2_64():
    """calc_offset has no range check: out-of-range x indices wrap in size_t."""
    d = O.desc(8, 8, 4)
    # ux = +1 -> -1 + sizes[0] after negation: an address beyond the stored half
    a = O.calc_offset(d, (1, 0, 0, 0))
    assert a == 8 * d.base_sizes[1] * d.base_sizes[2] * d.base_sizes[3]
    # a negative wrapped index far below zero becomes a huge unsigned value
    assert O.calc_offset(d, (0, -40, 0, 0)) >= 2 ** 63


# --------------------------------------------------------------------------
# FFT
# --------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(8, 8), (16, 32), (12, 10), (9, 7), (64, 64)])
def test_fft_matches_numpy_and_dft(shape):
    rng = np.random.default_rng(0)
    x = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    want = np.fft.fft2(x)
    got = O.fft2(x, -1)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    back = O.fft2(got, +1)  # unnormalised backward transform
    assert np.max(np.abs(back - x * x.size)) < 1e-11 * x.size
    if x.size <= 100:  # brute-force DFT by definition
        nr, nc = shape
        r, c = np.arange(nr), np.arange(nc)
        W = np.exp(-2j * np.pi * np.outer(r, r) / nr)
        V = np.exp(-2j * np.pi * np.outer(c, c) / nc)
        assert np.allclose(W @ x @ V.T, got, rtol=0, atol=1e-10)


def test_rfft_irfft():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(16, 32))
    h = O.rfft2(x)
    assert np.allclose(h, np.fft.rfft2(x), atol=1e-11)
    assert np.allclose(O.irfft2(h, 32), x * x.size, atol=1e-10)


def test_fft_delta_and_parseval():
    x = np.zeros((16, 16))
    x[3, 5] = 2.0
    F = O.fft2(x)
    assert np.allclose(np.abs(F), 2.0)
    rng = np.random.default_rng(2)
    y = rng.normal(size=(32, 32))
    assert np.isclose(np.sum(np.abs(O.fft2(y)) ** 2), y.size * np.sum(y ** 2))


# --------------------------------------------------------------------------
# accumulate_from_fft
# --------------------------------------------------------------------------
def _valid_tuples(nrows, ncols, b):
    """Explicit enumeration of the accumulate domain from Array2 / descriptor ranges."""
    d = O.desc(ncols, nrows, b)
    fmin = (-(ncols // 2), -(nrows // 2))
    fmax = (ncols - 1 - ncols // 2, nrows - 1 - nrows // 2)
    out = []
    for i in range(max(fmin[0], d.min_idx[0]), 1):
        for j in range(max(fmin[1], d.min_idx[1]), min(fmax[1], d.max_idx[1]) + 1):
            for k in range(max(fmin[0], d.min_idx[2]), 1):
                for l in range(max(fmin[1], d.min_idx[3]), min(fmax[1], d.max_idx[3]) + 1):
                    if fmin[0] <= i + k <= fmax[0] and fmin[1] <= j + l <= fmax[1]:
                        out.append((i, j, k, l))
    return d, out


@pytest.mark.parametrize("shape,b", [((8, 8), 4), ((6, 5), 3), ((10, 12), 6)])
def test_accumulate_brute_force(shape, b):
    """Every valid (u, v) cell holds F(u) F(v) conj(F(u+v)); every other cell is 0."""
    rng = np.random.default_rng(5)
    nrows, ncols = shape
    x = rng.normal(size=shape)
    F = np.fft.fft2(x)
    d, tuples = _valid_tuples(nrows, ncols, b)
    B = O.new_bispectrum(d)
    O.accumulate_from_fft(d, B, F)
    seen = np.zeros(B.size, dtype=bool)
    for (i, j, k, l) in tuples:
        want = F[j % nrows, i % ncols] * F[l % nrows, k % ncols] * np.conj(F[(j + l) % nrows, (i + k) % ncols])
        a = O.calc_offset(d, (i, j, k, l))
        assert not seen[a]
        seen[a] = True
        assert abs(B[a] - want) <= 1e-12 * max(1.0, abs(want))
    assert np.all(B[~seen] == 0)
    assert len(tuples) == O.count_triples(d, ncols, nrows)


@pytest.mark.parametrize("N,b,expected", [(128, 32, 3_829_488), (256, 48, 34_936_200)])
def test_triple_counts(N, b, expected):
    """Valid triples per frame of the loop bounds (include/bispectrum.h:383-398)."""
    assert O.count_triples(O.desc(N, N, b), N, N) == expected


def test_delta_frame_gives_unit_bispectrum():
    """A delta frame has |F| = 1 and a flat closure phase: B == 1 on every written cell."""
    N, b = 16, 6
    x = np.zeros((N, N))
    x[4, 9] = 1.0
    d = O.desc(N, N, b)
    B = O.new_bispectrum(d)
    O.accumulate_from_fft(d, B, O.fft2(x))
    nz = B != 0
    assert nz.sum() == O.count_triples(d, N, N)
    assert np.max(np.abs(B[nz] - 1)) < 1e-12


def test_translation_invariance():
    """The bispectrum of a circularly shifted frame is the same (phase terms cancel)."""
    rng = np.random.default_rng(9)
    N, b = 16, 8
    x = rng.normal(size=(N, N))
    d = O.desc(N, N, b)
    B1, B2 = O.new_bispectrum(d), O.new_bispectrum(d)
    O.accumulate_from_fft(d, B1, O.fft2(x))
    O.accumulate_from_fft(d, B2, O.fft2(np.roll(x, (3, -5), axis=(0, 1))))
    assert np.max(np.abs(B1 - B2)) < 1e-10 * np.max(np.abs(B1))


def test_two_delta_closed_form():
    """f = d(x) + a d(x - s): B(u,v) = sum over the 8 sign patterns (closed form)."""
    N, b, a = 16, 6, 0.6
    s = np.array([3, 2])  # (x, y)
    x = np.zeros((N, N))
    x[0, 0] = 1.0
    x[s[1], s[0]] = a
    d = O.desc(N, N, b)
    B = O.new_bispectrum(d)
    O.accumulate_from_fft(d, B, O.fft2(x))

    def Fk(kx, ky):
        return 1 + a * np.exp(-2j * np.pi * (kx * s[0] + ky * s[1]) / N)

    for (i, j, k, l) in _valid_tuples(N, N, b)[1][::37]:
        want = Fk(i, j) * Fk(k, l) * np.conj(Fk(i + k, j + l))
        assert abs(O.get_element(d, B, (i, j, k, l)) - want) < 1e-12


def test_f32_accumulate_matches_f64_to_float_precision():
    rng = np.random.default_rng(4)
    N, b = 16, 6
    d = O.desc(N, N, b)
    B64, B32 = O.new_bispectrum(d), O.new_bispectrum(d, np.complex64)
    for _ in range(4):
        F = O.fft2(rng.normal(size=(N, N)))
        O.accumulate_from_fft(d, B64, F)
        O.accumulate_from_fft(d, B32, F)
    assert np.max(np.abs(B32 - B64)) < 1e-5 * np.max(np.abs(B64))


# --------------------------------------------------------------------------
# get_element symmetries
# --------------------------------------------------------------------------
def test_get_element_symmetries_against_direct_products():
    """For one real frame, every canonical (non-aliased) lookup equals F(u)F(v)F*(u+v).

    This covers the six symmetry cases T1/T7/T6/T9/T3/T12 of canonicalize_indices
    (include/bispectrum.h:553-589) with the pre- and post-swap of get_element."""
    rng = np.random.default_rng(12)
    N, b = 16, 8
    x = rng.normal(size=(N, N))
    F = np.fft.fft2(x)
    d = O.desc(N, N, b)
    B = O.new_bispectrum(d)
    O.accumulate_from_fft(d, B, F)
    h = b // 2
    checked = aliased = 0
    cases = set()
    for ux in range(-N // 2, N // 2 + 1):
        for uy in range(-N // 2, N // 2 + 1):
            for vx in range(-h, h + 1):
                for vy in range(-h, h + 1):
                    try:
                        addr, conj, alias = O.lookup(d, (ux, uy, vx, vy))
                    except O.OracleError:
                        continue
                    if alias:
                        aliased += 1
                        continue
                    wx, wy = ux + vx, uy + vy
                    if not (-N // 2 <= wx <= N // 2 - 1 and -N // 2 <= wy <= N // 2 - 1):
                        continue  # u+v outside the stored FFT range: never written
                    if not (-N // 2 <= ux <= N // 2 - 1 and -N // 2 <= uy <= N // 2 - 1):
                        continue
                    want = F[uy % N, ux % N] * F[vy % N, vx % N] * np.conj(F[wy % N, wx % N])
                    got = O.get_element(d, B, (ux, uy, vx, vy))
                    if got == 0:
                        continue  # canonical cell outside the accumulated domain
                    assert abs(got - want) < 1e-9 * max(1, abs(want)), (ux, uy, vx, vy)
                    cases.add((ux > 0, vx > 0, ux + vx > 0))
                    checked += 1
    assert checked > 10000 and aliased > 0
    assert len(cases) == 6  # every symmetry case exercised


# --------------------------------------------------------------------------
# phase reconstruction
# --------------------------------------------------------------------------
def test_next_reco_index_first_visits():
    v = O.spiral_visits(2.0)
    assert v[:13].tolist() == [[1, 0], [0, 0], [2, 0], [1, 0], [1, 1], [0, 1], [-1, 1], [-1, 0], [-1, -1],
                               [0, -1], [1, -1], [1, 0], [3, 0]]
    r, phi, i, j = O.next_reco_index(0.0, 0.0, 0, 0)  # dphi = +inf at r = 0
    assert (r, phi, i, j) == (1.0, 0.0, 1, 0)


def test_delta_object_phase_recovery():
    """B == 1 (delta object): every spiral point in the disc is flagged with phase 1."""
    N, b, p = 64, 16, 30
    d = O.desc(N, N, b)
    B = O.new_bispectrum(d)
    O.accumulate_from_fft(d, B, np.ones((N, N), dtype=np.complex128))
    r = O.reconstruct_phases(d, B, N, N, p)
    vis = {tuple(v) for v in O.spiral_visits(p).tolist()
           if -N // 2 <= v[0] < N // 2 and -N // 2 <= v[1] < N // 2}
    fl = {(x, y) for y in range(-N // 2, N // 2) for x in range(-N // 2, N // 2) if r.flags[y % N, x % N]}
    assert fl == vis
    m = r.flags.astype(bool)
    assert np.max(np.abs(r.phases[m] - 1)) < 1e-14
    assert np.max(np.abs(r.consistency[m] - 1)) < 1e-14


def test_calc_phase_call_count_c1():
    """12,835 calc_phase calls at 128^2, b=32, p=64 with every call succeeding."""
    N, b, p = 128, 32, 64
    d = O.desc(N, N, b)
    B = O.new_bispectrum(d, np.complex64)
    O.accumulate_from_fft(d, B, np.ones((N, N), dtype=np.complex128))
    r = O.reconstruct_phases(d, B, N, N, p)
    assert r.stats["calls"] == 12_835
    assert r.stats["flagged"] == 12_835
    assert r.stats["aliased"] > 0


def test_single_frame_recovers_object_phases():
    """Noise-free single frame: recovered phase = object phase up to the tilt fixed
    by the seeds phi(1,0) = phi(0,1) = 0.

    With p <= b/2 every flagged u has |uy| <= b/2, so no lookup goes through the
    aliasing T3/T9 post-swap and the recursion is exact."""
    rng = np.random.default_rng(21)
    N, b, p = 32, 24, 12
    x = np.zeros((N, N))
    for _ in range(4):
        x[rng.integers(0, N), rng.integers(0, N)] += rng.uniform(0.5, 1.0)
    x += 0.05  # keeps |F| away from zero
    F = np.fft.fft2(x)
    d = O.desc(N, N, b)
    B = O.new_bispectrum(d)
    O.accumulate_from_fft(d, B, F)
    r = O.reconstruct_phases(d, B, N, N, p)
    Fh = F / np.abs(F)
    t10, t01 = Fh[0, 1], Fh[1, 0]
    m = r.flags.astype(bool)
    worst = 0.0
    for yy in range(N):
        for xx in range(N):
            if not m[yy, xx]:
                continue
            wx = xx if xx < N // 2 else xx - N
            wy = yy if yy < N // 2 else yy - N
            want = Fh[yy, xx] * np.conj(t10) ** wx * np.conj(t01) ** wy
            worst = max(worst, abs(r.phases[yy, xx] - want))
    assert r.stats["aliased"] == 0
    assert m.sum() > 400 and worst < 1e-8


# --------------------------------------------------------------------------
# cross-correlation, shift-and-add, window
# --------------------------------------------------------------------------
def test_xcorr_brute_force():
    rng = np.random.default_rng(3)
    ref, fr = rng.normal(size=(8, 8)), rng.normal(size=(8, 8))
    c = O.xcorr(ref, fr)
    want = np.zeros((8, 8))
    for dy in range(8):
        for dx in range(8):
            want[dy, dx] = 64 * np.sum(ref * np.roll(fr, (-dy, -dx), axis=(0, 1)))
    assert np.allclose(c, want, atol=1e-10)


def test_displacement_recovers_shifts_and_first_max():
    rng = np.random.default_rng(8)
    N = 32
    ref = np.zeros((N, N))
    ref[10:14, 7:12] = rng.uniform(1, 2, size=(4, 5))
    for dx, dy in [(3, -5), (-7, 2), (0, 0), (15, -16), (-16, 15)]:
        fr = np.roll(ref, (dy, dx), axis=(0, 1))
        assert O.displacement(O.xcorr(ref, fr)) == (dx, dy)
    c = np.zeros((4, 4))
    c[1, 3] = c[2, 0] = 5.0  # tie: the lowest linear index wins, then the signed wrap
    assert O.displacement(c) == (-1, 1)


def test_hann_closed_form():
    w = O.hann(8, 6, 10.0)
    for j in range(-3, 3):
        for i in range(-4, 4):
            want = (0.5 + 0.5 * np.cos(2 * np.pi * i / 10)) * (0.5 + 0.5 * np.cos(2 * np.pi * j / 10))
            assert w[j % 6, i % 8] == pytest.approx(want, abs=1e-15)
    with pytest.raises(O.OracleError) as e:
        O.hann(4, 4, 0.0)
    assert e.value.code == O.DOMAIN_ERROR
