"""Object lifetimes and error paths of the C ABI on the GPU.

* A context is not destroyed while objects made on it are alive (each of them
  uses it until it is destroyed); the Python Context closes its live objects
  first.
* Finalisation divides the bispectrum by n once (smip-cli/main.cpp:312): when
  phase recovery fails (forced from the host with SMIP_INJECT_FAIL, before any
  kernel of the stage is launched), a retry recovers from the same normalised
  bispectrum and the results equal those of a job finalised once.
"""
import numpy as np
import pytest

import smip_amd as S

pytestmark = pytest.mark.gpu


def _frames(n, N, seed=5):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:N, 0:N]
    out = np.empty((n, N, N), dtype=np.uint8)
    for f in range(n):
        cx, cy = rng.integers(N // 4, 3 * N // 4, size=2)
        img = 150 * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / 6.0) + rng.uniform(0, 12, size=(N, N))
        out[f] = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    return out


def test_context_refuses_destroy_while_objects_live():
    ctx = S.Context(0)
    b = S.Bispectrum(ctx, (16, 16, 4, 4))
    rc = S.lib().smip_ctx_destroy(ctx.h)
    assert rc == S.ERR_LOGIC
    assert b"live objects" in S.lib().smip_last_error()
    b.close()
    ctx.close()  # nothing left: destroyed
    assert ctx.h is None


def test_context_close_closes_its_objects_first():
    ctx = S.Context(0)
    frames = _frames(6, 32)
    dev = S.Frames(ctx, frames)
    job = S.Job(ctx, S.job_params(32, 32, 8, 12, np.uint8))
    job.add_frames(dev)
    xc = S.CrossCorrelation(ctx, frames[0].astype(np.float64))
    srv = S.Server(ctx, workers=2)
    ctx.close()
    assert ctx.h is None
    assert dev.h is None and job.h is None and xc.h is None and srv.h is None


@pytest.mark.parametrize("grouped", [False, True])
def test_finalize_retry_after_failed_recovery(ctx, monkeypatch, grouped):
    frames = _frames(10, 64, seed=9)
    p = S.job_params(64, 64, 16, 24, np.uint8, frames_per_batch=4)
    ref = S.Job(ctx, p)
    ref.push_frames_host(frames)
    S.finalize_many([ref]) if grouped else ref.finalize()
    want = ref.result()
    ref.close()

    job = S.Job(ctx, p)
    job.push_frames_host(frames)
    monkeypatch.setenv("SMIP_INJECT_FAIL", "reco")
    with pytest.raises(S.SmipError) as e:
        S.finalize_many([job]) if grouped else job.finalize()
    assert e.value.code == S.ERR_OUT_OF_RANGE
    monkeypatch.delenv("SMIP_INJECT_FAIL")
    # the sum is already divided by n: no frame may join it any more
    with pytest.raises(S.SmipError) as e:
        job.push_frames_host(frames[:1])
    assert e.value.code == S.ERR_LOGIC
    S.finalize_many([job]) if grouped else job.finalize()
    got = job.result()
    job.close()
    for k in ("bispectrum", "powerspec", "sum_image", "shifts", "flags"):
        assert np.array_equal(getattr(got, k), getattr(want, k)), k
    # the recovery's FP64 sums run in one fixed order, so the phases repeat exactly
    assert np.array_equal(got.phases, want.phases)
    assert np.array_equal(got.image, want.image)
    # and a finalised job refuses a second finalisation
    job2 = S.Job(ctx, p)
    job2.push_frames_host(frames)
    job2.finalize()
    with pytest.raises(S.SmipError) as e:
        job2.finalize()
    assert e.value.code == S.ERR_LOGIC
    job2.close()
