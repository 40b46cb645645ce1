"""The engine at the measured configurations' own sizes, against answers the
oracle wrote beforehand (tests/golden/make_goldens.py; the oracle needs minutes
per case on one CPU core, the GPU milliseconds).

* 512x512 float32 extended object, b = 64, p = 256 (frames 0-2 of the stack),
  mixed path in 256-frame batches and plain FP64 path:
    shifts, sum image and phase-consistency flags bit-exact;
    the set of zero bispectrum cells exact, B within 2e-5 (mixed) / 1e-11 (FP64)
    of max|B| at 8192 seeded cells and the 256 largest;
    |image| within 1e-5 rel L2, FP64 phases within 1e-9 rad, mixed phases within
    1e-4 rad, each relaxed to 10 x the recursion's own conditioning (the FP64
    path on the same frames moved by one float32 ulp) should that be larger.
* 256x256 uint16 triple star, b = 48, p = 128, all 2000 frames (four FP64 folds
  of 512-frame batches):
    every shift and the sum image bit-exact against the oracle, flags of the
    mixed path equal to the FP64 path's, image within 1e-5 rel L2 and B within
    2e-5 of max|B| of the FP64 path (which matches the oracle to 1e-11 at the
    sizes the oracle finishes).
"""
import hashlib
import os

import numpy as np
import pytest

import smip_amd as S

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _zero_sha(B) -> str:
    h = hashlib.sha256()
    step = 1 << 24
    for k in range(0, B.size, step):
        h.update(np.packbits(B[k:k + step] == 0).tobytes())
    return h.hexdigest()


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _run(ctx, frames, b, p, precision, batch=0):
    n, nrows, ncols = frames.shape
    job = S.Job(ctx, S.job_params(ncols, nrows, b, p, frames.dtype, precision, saa=True, frames_per_batch=batch))
    try:
        job.add_frames(S.Frames(ctx, frames))
        job.finalize()
        return job.result(with_bispectrum=True)
    finally:
        job.close()


@pytest.mark.timeout(600)
def test_extended512_against_oracle(ctx):
    g = np.load(os.path.join(GOLDEN, "extended512_f3.npz"))
    b, p = int(g["b"]), int(g["p"])
    frames = S.synth_frames(S.synth_params("extended512"), 0, int(g["frames"][1]))
    want_flags = np.unpackbits(g["flags_bits"])[: 512 * 512].reshape(512, 512)
    want_img = g["image_abs"].astype(np.float64)
    res = {}
    for name, prec, batch in (("mixed", S.PRECISION_MIXED, 256), ("fp64", S.PRECISION_FP64, 0)):
        r = _run(ctx, frames, b, p, prec, batch)
        assert np.array_equal(r.shifts, g["shifts"]), name
        assert _sha(r.sum_image) == str(g["sum_sha"]), name
        assert np.array_equal(r.flags, want_flags), name
        assert _zero_sha(r.bispectrum) == str(g["zero_sha"]), name
        berr = float(np.max(np.abs(r.bispectrum[g["cells"]] - g["cell_values"])) / float(g["bmax"]))
        res[name] = dict(berr=berr, phases=r.phases.ravel()[g["points"]].copy(),
                         image=np.abs(r.image), flags=r.flags)
        del r
    assert res["mixed"]["berr"] < 2e-5, res["mixed"]["berr"]
    assert res["fp64"]["berr"] < 1e-11, res["fp64"]["berr"]

    # conditioning: the FP64 path on frames moved by one float32 ulp
    pert = _run(ctx, np.nextafter(frames, np.float32(np.inf)), b, p, S.PRECISION_FP64)
    sens_img = _rel_l2(np.abs(pert.image), res["fp64"]["image"])
    sens_ph = float(np.max(np.abs(pert.phases.ravel()[g["points"]] - res["fp64"]["phases"])))
    del pert
    out = {"sens_image": sens_img, "sens_phase": sens_ph}
    for name in ("mixed", "fp64"):
        out[name + "_image"] = _rel_l2(res[name]["image"], want_img)
        out[name + "_phase"] = float(np.max(np.abs(res[name]["phases"] - g["point_phases"])))
        out[name + "_bispectrum"] = res[name]["berr"]
    print("extended512 f3:", {k: f"{v:.3g}" for k, v in out.items()})
    # Image: within 1e-5 rel L2 of the oracle (the engine's target), or 10 x the
    # conditioning figure if rounding alone (one ulp of input) moves the FP64
    # result further.  Phases: the FP64 path within 1e-9 rad (or 10 x sens), the
    # mixed path within 1e-4 rad, the spread its FP32 triple products leave in the
    # bispectrum phases.  All figures are printed above.
    for name in ("mixed", "fp64"):
        assert out[name + "_image"] <= max(1e-5, 10 * sens_img), (name, out)
    assert out["fp64_phase"] <= max(1e-9, 10 * sens_ph), out
    assert out["mixed_phase"] <= 1e-4, out


@pytest.mark.timeout(600)
def test_triple256_full_stack(ctx):
    g = np.load(os.path.join(GOLDEN, "triple256_saa.npz"))
    W = S.WORKLOADS["triple256"]
    frames = S.synth_frames(S.synth_params("triple256"), 0, W["n"])
    rm = _run(ctx, frames, W["b"], W["p"], S.PRECISION_MIXED)
    rd = _run(ctx, frames, W["b"], W["p"], S.PRECISION_FP64)
    for r in (rm, rd):
        assert r.stats["frames"] == W["n"]
        assert np.array_equal(r.shifts, g["shifts"])
        assert _sha(r.sum_image) == str(g["sum_sha"])
    assert np.array_equal(rm.flags, rd.flags)
    assert np.array_equal(rm.bispectrum == 0, rd.bispectrum == 0)
    berr = float(np.max(np.abs(rm.bispectrum - rd.bispectrum)) / np.max(np.abs(rd.bispectrum)))
    img = _rel_l2(np.abs(rm.image), np.abs(rd.image))
    print(f"triple256 2000 frames: mixed vs FP64 bispectrum {berr:.3g} of max|B|, image rel L2 {img:.3g}")
    assert berr < 2e-5
    assert img < 1e-5
