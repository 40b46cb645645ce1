"""The batched kernel's autotuner: its choice is one of the supported shapes,
later jobs of the same geometry take it, and results stay within the mixed
path's tolerance against the oracle."""
import numpy as np
import pytest

import oracle as O
import smip_amd as S

pytestmark = pytest.mark.gpu


def test_autotune_choice_is_used_and_correct(ctx):
    S.autotune_forget()
    base = S.job_params(128, 128, 32, 64, np.uint8)
    assert S.resolve_params(base).frames_per_batch == 512  # built-in default
    r = S.autotune(ctx, 128, 128, 32)
    assert r["vy_tile"] in (9, 13, 17, 25, 33)
    assert r["frames_per_batch"] in (32, 64, 128, 256, 512, 1024)
    assert r["frames_per_s"] > 0 and r["tile_tflops"] > 0 and r["candidates"] >= 6
    q = S.resolve_params(base)
    assert (q.vy_tile, q.frames_per_batch) == (r["vy_tile"], r["frames_per_batch"])
    # explicit parameters win over the tuned ones
    assert S.resolve_params(S.job_params(128, 128, 32, 64, np.uint8, frames_per_batch=48)).frames_per_batch == 48
    W = S.WORKLOADS["binary128"]
    frames = S.synth_frames(S.synth_params("binary128"), 0, W["n"])
    got = S.run_pipeline(ctx, frames, 32, 64)
    want = O.pipeline(frames.astype(np.float64), 32, 64)
    assert np.array_equal(got.flags, want.flags) and np.array_equal(got.shifts, want.shifts)
    assert np.linalg.norm(got.image - want.image) / np.linalg.norm(want.image) < 1e-5
    S.autotune_forget()
    assert S.resolve_params(base).frames_per_batch == 512


def test_autotune_memory_limit(ctx):
    with pytest.raises(S.SmipError) as e:
        S.autotune(ctx, 256, 256, 48, max_bytes=1 << 20)
    assert e.value.code == S.ERR_NO_MEMORY
