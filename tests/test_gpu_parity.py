"""GPU parity of the HIP engine against the CPU oracle (through the C ABI).

Bars (stated per test):
  * integer / index work -- shifts, flags, element addresses: bit-exact;
  * the drop-in accumulate_from_fft on identical FFT input: bit-exact (both sides
    evaluate the reference's operation order without FMA contraction);
  * FP64 engine (FP64 FFT of our own, FP64 triple products): relative 1e-11;
  * mixed engine (FP32 FFT and triple products per batch, FP64 fold): bispectrum
    relative error 2e-5 of max|B|, image relative L2 1e-5 (the engine's stated accuracy target).
"""
import numpy as np
import pytest

import oracle as O
import smip_amd as S

pytestmark = pytest.mark.gpu


def _frames(n, N, dtype=np.uint8, seed=1, hi=255):
    rng = np.random.default_rng(seed)
    # a bright blob pair on a noisy background keeps the correlation peak unique
    y, x = np.mgrid[0:N, 0:N]
    out = np.empty((n, N, N), dtype=dtype)
    for f in range(n):
        cx, cy = rng.integers(N // 4, 3 * N // 4, size=2)
        img = 0.6 * hi * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / 8.0)
        img += 0.35 * hi * np.exp(-((x - cx - 5) ** 2 + (y - cy - 3) ** 2) / 6.0)
        img += rng.uniform(0, 0.05 * hi, size=(N, N))
        out[f] = np.clip(np.rint(img), 0, hi).astype(dtype) if np.issubdtype(dtype, np.integer) else img.astype(dtype)
    return out


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape,b", [((16, 16), 8), ((12, 10), 6), ((9, 7), 5), ((32, 32), 12)])
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_accumulate_from_fft_bitexact(ctx, shape, b, dtype):
    """Bispectrum<T>::accumulate_from_fft, include/bispectrum.h:356-409, any frame shape."""
    rng = np.random.default_rng(7)
    nrows, ncols = shape
    d = O.desc(ncols, nrows, b)
    ref = O.new_bispectrum(d, dtype)
    gpu = S.Bispectrum(ctx, (ncols, nrows, b, b), dtype)
    for _ in range(3):
        fft = O.fft2(rng.normal(size=shape) * 10)
        O.accumulate_from_fft(d, ref, fft)
        gpu.accumulate_from_fft(fft)
    got = gpu.download()
    assert got.dtype == ref.dtype
    assert np.array_equal(got, ref)


def test_get_elements_bitexact_with_aliasing(ctx):
    """get_element with all six symmetry cases and calc_offset's wrapping lookups."""
    rng = np.random.default_rng(3)
    N, b = 32, 12
    d = O.desc(N, N, b)
    B = (rng.normal(size=int(d.base_size)) + 1j * rng.normal(size=int(d.base_size))).astype(np.complex128)
    gpu = S.Bispectrum(ctx, (N, N, b, b))
    gpu.upload(B)
    h = b // 2
    idx = []
    for _ in range(4000):  # the lookups calc_phase makes: u anywhere, v in the stored v range
        ux, uy = rng.integers(-N // 2, N // 2, size=2)
        vx, vy = rng.integers(-h, h + 1, size=2)
        idx.append((ux, uy, vx, vy))
    ok = [i for i in idx]
    want = []
    aliased = 0
    for i in ok:
        want.append(O.get_element(d, B, i))
        aliased += O.lookup(d, i)[2]
    got = gpu.get_elements(ok)
    assert aliased > 0
    assert np.array_equal(got, np.array(want))
    # an out-of-range lookup raises std::out_of_range's code
    with pytest.raises(S.SmipError) as e:
        gpu.get_element((N, 0, 0, 0))
    assert e.value.code == S.ERR_OUT_OF_RANGE


@pytest.mark.parametrize("N,b,p", [(32, 8, 16), (64, 16, 32)])
def test_reconstruct_phases_dropin(ctx, N, b, p):
    """reconstruct_phases on the same bispectrum: flags bit-exact, phases to 1e-9."""
    frames = _frames(6, N).astype(np.float64)
    B, _ = O.power_and_bispectrum(frames, b)
    d = O.desc(N, N, b)
    ref = O.reconstruct_phases(d, B, N, N, p)
    gpu = S.Bispectrum(ctx, (N, N, b, b))
    gpu.upload(B)
    ph, fl, co = gpu.reconstruct_phases(N, N, p)
    assert np.array_equal(fl, ref.flags)
    m = ref.flags.astype(bool)
    assert np.max(np.abs(ph[m] - ref.phases[m])) < 1e-9
    assert np.max(np.abs(co - ref.consistency)) < 1e-9
    assert np.all(ph[~m] == 0)


def test_reconstruct_phases_radius_beyond_the_map(ctx):
    """Past the map's half-diagonal the spiral (src/phasereco.cpp:17-36) only visits
    points outside the PhaseMap, which reconstruct_phases skips; the visit list is
    cut there, so a radius of 1e6 finishes at once and gives the oracle's result at
    a radius that already covers the map."""
    N, b = 16, 8
    frames = _frames(4, N).astype(np.float64)
    B, _ = O.power_and_bispectrum(frames, b)
    d = O.desc(N, N, b)
    ref = O.reconstruct_phases(d, B, N, N, 2.0 * N)
    gpu = S.Bispectrum(ctx, (N, N, b, b))
    gpu.upload(B)
    ph, fl, co = gpu.reconstruct_phases(N, N, 1e6)
    assert np.array_equal(fl, ref.flags)
    m = ref.flags.astype(bool)
    assert m.sum() > N * N // 2
    assert np.max(np.abs(ph[m] - ref.phases[m])) < 1e-9
    assert np.max(np.abs(co - ref.consistency)) < 1e-9


def test_reconstruct_phases_c64_bispectrum(ctx):
    """The CLI's Bispectrum<complex<float>> input (smip-cli/main.cpp:319)."""
    N, b, p = 32, 8, 16
    frames = _frames(5, N).astype(np.float64)
    B, _ = O.power_and_bispectrum(frames, b, bispec_f32=True)
    d = O.desc(N, N, b)
    ref = O.reconstruct_phases(d, B, N, N, p)
    gpu = S.Bispectrum(ctx, (N, N, b, b), np.complex64)
    gpu.upload(B)
    ph, fl, co = gpu.reconstruct_phases(N, N, p)
    assert np.array_equal(fl, ref.flags)
    assert np.max(np.abs(ph - ref.phases)) < 1e-9


def test_xcorr_displacement(ctx):
    """CrossCorrelation<double>: correlation array and exact integer shifts."""
    rng = np.random.default_rng(11)
    N = 64
    ref = _frames(1, N, np.float64, seed=5)[0]
    xc = S.CrossCorrelation(ctx, ref)
    with pytest.raises(S.SmipError) as e:
        xc.get_displacement()
    assert e.value.code == S.ERR_BAD_FUNCTION_CALL
    for _ in range(12):
        dx, dy = (int(v) for v in rng.integers(-20, 21, size=2))
        fr = np.roll(ref, (dy, dx), axis=(0, 1))
        want = O.displacement(O.xcorr(ref, fr))
        assert xc(fr) == want == (dx, dy)
    xc.correlate(fr)
    c = xc.get_correlation_array()
    assert _rel(c, O.xcorr(ref, fr)) < 1e-12
    with pytest.raises(S.SmipError) as e:
        xc.correlate(np.zeros((N, N // 2)))
    assert e.value.code == S.ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("N,b,p,n", [(32, 8, 16, 9), (64, 16, 32, 20)])
def test_engine_fp64(ctx, N, b, p, n):
    """The plain FP64 engine against the FP64 oracle pipeline."""
    frames = _frames(n, N, seed=N)
    want = O.pipeline(frames.astype(np.float64), b, p)
    got = S.run_pipeline(ctx, frames, b, p, precision=S.PRECISION_FP64, frames_per_batch=4)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.sum_image, want.sum_image)
    assert _rel(got.bispectrum, want.bispectrum) < 1e-11
    assert _rel(got.powerspec, want.powerspec) < 1e-11
    assert np.array_equal(got.flags, want.flags)
    assert np.max(np.abs(got.phases - want.phases)) < 1e-9
    assert _rel_l2(got.image, want.image) < 1e-9


@pytest.mark.parametrize("N,b,p,n,batch", [(32, 8, 16, 9, 4), (64, 16, 32, 20, 8), (64, 12, 30, 17, 0)])
def test_engine_mixed(ctx, N, b, p, n, batch):
    """The batched mixed-precision engine against the FP64 oracle pipeline."""
    frames = _frames(n, N, seed=N + 1)
    want = O.pipeline(frames.astype(np.float64), b, p)
    got = S.run_pipeline(ctx, frames, b, p, precision=S.PRECISION_MIXED, frames_per_batch=batch)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.sum_image, want.sum_image)
    assert _rel(got.bispectrum, want.bispectrum) < 2e-5
    # cells the reference never writes stay exactly zero
    assert np.array_equal(got.bispectrum == 0, want.bispectrum == 0)
    assert _rel(got.powerspec, want.powerspec) < 1e-5
    assert np.array_equal(got.flags, want.flags)
    assert _rel_l2(got.image, want.image) < 1e-5


@pytest.mark.parametrize("precision", [S.PRECISION_FP64, S.PRECISION_MIXED])
@pytest.mark.parametrize("b", [32, 40])
def test_depth_at_and_beyond_half_frame(ctx, precision, b):
    """b/2 >= N/2: the accumulate loops clip v to the frame (include/bispectrum.h:383-391)."""
    N, p, n = 32, 12, 6
    frames = _frames(n, N, seed=23)
    want = O.pipeline(frames.astype(np.float64), b, p)
    got = S.run_pipeline(ctx, frames, b, p, precision=precision)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.flags, want.flags)
    assert np.array_equal(got.bispectrum == 0, want.bispectrum == 0)
    assert _rel(got.bispectrum, want.bispectrum) < (1e-11 if precision == S.PRECISION_FP64 else 2e-5)
    assert _rel_l2(got.image, want.image) < (1e-9 if precision == S.PRECISION_FP64 else 1e-5)


@pytest.mark.parametrize("vy_tile", [9, 13, 17, 25, 33])
def test_tile_shapes(ctx, vy_tile):
    """Every instantiated register-tile shape gives the same bispectrum."""
    N, b, n = 64, 24, 6
    frames = _frames(n, N, seed=3)
    B, _ = O.power_and_bispectrum(frames.astype(np.float64), b)
    got = S.run_pipeline(ctx, frames, b, 20, precision=S.PRECISION_MIXED, saa=False, vy_tile=vy_tile)
    assert _rel(got.bispectrum, B) < 2e-5
    assert np.array_equal(got.bispectrum == 0, B == 0)


@pytest.mark.parametrize("N", [8, 2048])
def test_transform_size_extremes(ctx, N):
    """The smallest and largest frame sides the engine takes: every transform (the
    FP32 forward of the bispectrum, the FP64 ones of shift-and-add and synthesis)
    at lengths 8 and 2048, whose workgroup shapes and LDS budgets differ from the
    usual sizes, through a whole job against the oracle."""
    frames = _frames(3, N, seed=23)
    b, p = 2, 4
    want = O.pipeline(frames.astype(np.float64), b, p)
    got = S.run_pipeline(ctx, frames, b, p, precision=S.PRECISION_MIXED)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.sum_image, want.sum_image)
    assert np.array_equal(got.flags, want.flags)
    assert np.array_equal(got.bispectrum == 0, want.bispectrum == 0)
    assert _rel(got.bispectrum, want.bispectrum) < 2e-5
    assert _rel(got.powerspec, want.powerspec) < 1e-5
    assert _rel_l2(got.image, want.image) < 1e-5


def test_reco_windowed_matches_sequential(ctx, monkeypatch):
    """The windowed schedule against the one-visit-at-a-time kernel and the oracle:
    identical flags, phases equal up to the FP64 summation order."""
    N, b, p = 128, 32, 64
    frames = _frames(3, N, seed=17).astype(np.float64)
    B, _ = O.power_and_bispectrum(frames, b)
    gpu = S.Bispectrum(ctx, (N, N, b, b))
    gpu.upload(B)
    ph_w, fl_w, co_w = gpu.reconstruct_phases(N, N, p)
    monkeypatch.setenv("SMIP_RECO_SEQUENTIAL", "1")
    ph_s, fl_s, co_s = gpu.reconstruct_phases(N, N, p)
    monkeypatch.delenv("SMIP_RECO_SEQUENTIAL")
    ref = O.reconstruct_phases(O.desc(N, N, b), B, N, N, p)
    assert np.array_equal(fl_w, fl_s) and np.array_equal(fl_w, ref.flags)
    assert np.max(np.abs(ph_w - ph_s)) < 1e-10
    assert np.max(np.abs(ph_w - ref.phases)) < 1e-10
    assert np.max(np.abs(co_w - ref.consistency)) < 1e-10


@pytest.mark.parametrize("precision", [S.PRECISION_MIXED, S.PRECISION_FP64])
def test_binary128_full_config(ctx, precision):
    """40 frames 128x128 u8 binary star, b=32, p=64: the reference's CPU-runnable case."""
    W = S.WORKLOADS["binary128"]
    frames = S.synth_frames(S.synth_params("binary128"), 0, W["n"])
    want = O.pipeline(frames.astype(np.float64), W["b"], W["p"])
    got = S.run_pipeline(ctx, frames, W["b"], W["p"], precision=precision)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.sum_image, want.sum_image)
    assert np.array_equal(got.flags, want.flags)
    assert np.array_equal(got.bispectrum == 0, want.bispectrum == 0)
    tol_b = 2e-5 if precision == S.PRECISION_MIXED else 1e-11
    assert _rel(got.bispectrum, want.bispectrum) < tol_b
    assert _rel_l2(got.image, want.image) < (1e-5 if precision == S.PRECISION_MIXED else 1e-9)
    assert got.stats["reco_calls"] == want.stats["calls"]


def test_triple256_sample(ctx):
    """8 frames of the 256x256 u16 triple-star workload at b=48, p=128 (the metric's
    shape) through the mixed engine against the FP64 oracle."""
    W = S.WORKLOADS["triple256"]
    frames = S.synth_frames(S.synth_params("triple256"), 0, 8)
    want = O.pipeline(frames.astype(np.float64), W["b"], W["p"])
    got = S.run_pipeline(ctx, frames, W["b"], W["p"], precision=S.PRECISION_MIXED)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.flags, want.flags)
    assert _rel(got.bispectrum, want.bispectrum) < 2e-5
    err = _rel_l2(got.image, want.image)
    assert err < 1e-5, err


def test_assign_converts_element_type(ctx):
    """operator= keeps the destination's element type (complex<double> <-> complex<float>)."""
    rng = np.random.default_rng(3)
    a = S.Bispectrum(ctx, (16, 16, 8, 8), np.complex128)
    vals = (rng.standard_normal(a.download().shape) + 1j * rng.standard_normal(a.download().shape))
    a.upload(vals)
    f = S.Bispectrum(ctx, (8, 8, 4, 4), np.complex64)  # resized by the assignment
    lib = S.lib()
    S._check(lib.smip_bispec_assign(f.h, a.h))
    got = f.download()
    assert got.dtype == np.complex64 and np.array_equal(got, vals.astype(np.complex64))
    back = S.Bispectrum(ctx, (16, 16, 8, 8), np.complex128)
    S._check(lib.smip_bispec_assign(back.h, f.h))
    assert np.array_equal(back.download(), vals.astype(np.complex64).astype(np.complex128))


def _frames_rect(n, ncols, nrows, dtype=np.uint8, seed=1, hi=255):
    """_frames on an nrows x ncols grid (blob pair plus noise, a unique correlation peak)."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:nrows, 0:ncols]
    out = np.empty((n, nrows, ncols), dtype=dtype)
    for f in range(n):
        cx = rng.integers(ncols // 4, 3 * ncols // 4)
        cy = rng.integers(nrows // 4, 3 * nrows // 4)
        img = 0.6 * hi * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / 8.0)
        img += 0.35 * hi * np.exp(-((x - cx - 5) ** 2 + (y - cy - 3) ** 2) / 6.0)
        img += rng.uniform(0, 0.05 * hi, size=(nrows, ncols))
        out[f] = np.clip(np.rint(img), 0, hi).astype(dtype) if np.issubdtype(dtype, np.integer) else img.astype(dtype)
    return out


@pytest.mark.parametrize("ncols,nrows", [(64, 128), (128, 32)])
@pytest.mark.parametrize("precision", [S.PRECISION_MIXED, S.PRECISION_FP64])
def test_engine_non_square(ctx, ncols, nrows, precision):
    """Frames with ncols != nrows through the whole job (row and column transforms,
    the half-plane layout, the tile items, the phase map and the synthesis), so an
    x/y mix-up anywhere shows: shifts, sum image and flags bit-exact, bispectrum
    and image to the engine's tolerances."""
    b, p, n = 16, 24, 9
    frames = _frames_rect(n, ncols, nrows, seed=ncols + nrows)
    want = O.pipeline(frames.astype(np.float64), b, p)
    got = S.run_pipeline(ctx, frames, b, p, precision=precision, frames_per_batch=4)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.sum_image, want.sum_image)
    assert np.array_equal(got.flags, want.flags)
    assert np.array_equal(got.bispectrum == 0, want.bispectrum == 0)
    mixed = precision == S.PRECISION_MIXED
    assert _rel(got.bispectrum, want.bispectrum) < (2e-5 if mixed else 1e-11)
    assert _rel(got.powerspec, want.powerspec) < (1e-5 if mixed else 1e-11)
    assert _rel_l2(got.image, want.image) < (1e-5 if mixed else 1e-9)


@pytest.mark.parametrize("ncols,nrows", [(128, 64), (32, 64)])
def test_xcorr_non_square(ctx, ncols, nrows):
    """CrossCorrelation<double> on ncols != nrows: exact integer shifts and the
    correlation array against the oracle."""
    rng = np.random.default_rng(ncols * 3 + nrows)
    ref = _frames_rect(1, ncols, nrows, np.float64, seed=5)[0]
    xc = S.CrossCorrelation(ctx, ref)
    for _ in range(8):
        dx = int(rng.integers(-ncols // 3, ncols // 3 + 1))
        dy = int(rng.integers(-nrows // 3, nrows // 3 + 1))
        fr = np.roll(ref, (dy, dx), axis=(0, 1))
        want = O.displacement(O.xcorr(ref, fr))
        assert xc(fr) == want == (dx, dy)
    c = xc.get_correlation_array()
    assert c.shape == (nrows, ncols)
    assert _rel(c, O.xcorr(ref, fr)) < 1e-12


@pytest.mark.parametrize("ncols,nrows,b,p", [(100, 100, 16, 30), (96, 128, 16, 40), (120, 90, 20, 30),
                                              (45, 27, 8, 10), (131, 67, 12, 20), (7, 9, 4, 3)])
@pytest.mark.parametrize("precision", [S.PRECISION_MIXED, S.PRECISION_FP64])
def test_engine_any_size(ctx, ncols, nrows, b, p, precision):
    """Frame extents FFTW takes but a power-of-two FFT does not (smip-cli/main.cpp:256-259):
    even, odd, prime and mixed extents through the mixed-radix transforms of the
    whole job, against the oracle (whose FFT is a direct DFT at such lengths)."""
    n = 7
    frames = _frames_rect(n, ncols, nrows, seed=ncols * 7 + nrows)
    want = O.pipeline(frames.astype(np.float64), b, p)
    got = S.run_pipeline(ctx, frames, b, p, precision=precision, frames_per_batch=3)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.sum_image, want.sum_image)
    assert np.array_equal(got.flags, want.flags)
    assert np.array_equal(got.bispectrum == 0, want.bispectrum == 0)
    mixed = precision == S.PRECISION_MIXED
    assert _rel(got.bispectrum, want.bispectrum) < (2e-5 if mixed else 1e-11)
    assert _rel(got.powerspec, want.powerspec) < (1e-5 if mixed else 1e-11)
    assert _rel_l2(got.image, want.image) < (1e-5 if mixed else 1e-9)


@pytest.mark.parametrize("ncols,nrows", [(2039, 6), (2048, 10), (6, 2039)])
def test_longest_lengths_not_powers_of_two(ctx, ncols, nrows):
    """The largest prime length under the 2048 limit (one 2039-point stage of
    per-output sums) and the 2048-point mixed-radix transform beside a short
    non-power-of-two side, in rows and in columns."""
    frames = _frames_rect(3, ncols, nrows, seed=ncols + 3 * nrows)
    b, p = 2, 2
    want = O.pipeline(frames.astype(np.float64), b, p)
    got = S.run_pipeline(ctx, frames, b, p, precision=S.PRECISION_MIXED)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.sum_image, want.sum_image)
    assert np.array_equal(got.flags, want.flags)
    assert _rel(got.bispectrum, want.bispectrum) < 2e-5
    assert _rel(got.powerspec, want.powerspec) < 1e-5
    assert _rel_l2(got.image, want.image) < 1e-5


@pytest.mark.parametrize("ncols,nrows", [(100, 60), (131, 67), (9, 7)])
def test_xcorr_any_size(ctx, ncols, nrows):
    """CrossCorrelation<double> at extents that are not powers of two: exact shifts
    and the correlation array against the oracle."""
    rng = np.random.default_rng(ncols + nrows)
    ref = _frames_rect(1, ncols, nrows, np.float64, seed=11)[0]
    xc = S.CrossCorrelation(ctx, ref)
    for _ in range(6):
        dx = int(rng.integers(-(ncols // 3), ncols // 3 + 1))
        dy = int(rng.integers(-(nrows // 3), nrows // 3 + 1))
        fr = np.roll(ref, (dy, dx), axis=(0, 1))
        want = O.displacement(O.xcorr(ref, fr))
        assert xc(fr) == want
    c = xc.get_correlation_array()
    assert _rel(c, O.xcorr(ref, fr)) < 1e-12
