"""The continuous-batching request loop (smip_server_*) and the streaming frame
entry point it drives (smip_job_push_frames_host).

The property the serving form must keep: a job's result is the same alone and
in any mix of concurrent jobs.  Batch boundaries follow a job's own frame
sequence and every output cell is owned by one work item, so "the same" is
checked bit for bit here, which is stronger than the engine's stated tolerance
against the FP64 oracle (image rel L2 <= 1e-5, also checked)."""
import threading

import numpy as np
import pytest

import oracle as O
import smip_amd as S

pytestmark = pytest.mark.gpu


def _frames(kind, N, n, seed_offset=0):
    W = dict(S.WORKLOADS[kind])
    prm = S.synth_params(kind)
    prm.ncols = prm.nrows = N
    prm.seed = W["seed"] + seed_offset
    prm.pupil_diameter = max(8.0, N / 4)
    return S.synth_frames(prm, 0, n, threads=8)


def _outputs(job: S.Job):
    r = job.result(with_bispectrum=True)
    return dict(bispectrum=r.bispectrum, powerspec=r.powerspec, sum_image=r.sum_image, shifts=r.shifts,
                phases=r.phases, flags=r.flags, consistency=r.consistency, image=r.image)


def _assert_identical(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_push_is_independent_of_chunking(ctx):
    frames = _frames("binary128", 64, 200)
    p = S.job_params(64, 64, 16, 24, frames.dtype, frames_per_batch=32)
    ref = S.Job(ctx, p)
    ref.add_frames(S.Frames(ctx, frames))  # device stack: eager launches, batches of 32 and a last of 8
    ref.finalize()
    want = _outputs(ref)
    ref.close()
    for cuts in ([200], [1, 7, 33, 2, 47, 110], [31, 1, 31, 1, 26, 40, 70]):
        job = S.Job(ctx, p)
        pos = 0
        for c in cuts:
            job.push_frames_host(frames[pos:pos + c])
            pos += c
        job.finalize()
        _assert_identical(_outputs(job), want)
        job.close()


@pytest.mark.parametrize("precision", [S.PRECISION_MIXED, S.PRECISION_FP64])
def test_graph_replays_match_eager_launches(ctx, monkeypatch, precision):
    """Full batches of pushed frames replay captured HIP graphs; every replay
    (five here, with host work between them) must leave what the eager launches
    leave."""
    frames = _frames("triple256", 64, 6 * 24 + 5, seed_offset=7)
    p = S.job_params(64, 64, 20, 30, frames.dtype, precision, frames_per_batch=24)
    outs = {}
    for mode in ("graph", "eager"):
        if mode == "eager":
            monkeypatch.setenv("SMIP_NO_GRAPH", "1")
        job = S.Job(ctx, p)
        for k in range(0, len(frames), 13):
            job.push_frames_host(frames[k:k + 13])
            job.stats()  # synchronises: eager host work between the replays
        job.finalize()
        st = job.stats()
        assert st["frames"] == len(frames)
        assert st["graph_launches"] == (10 if mode == "graph" else 0)  # 5 batches x (accumulate, shift-and-add)
        outs[mode] = _outputs(job)
        job.close()
    monkeypatch.delenv("SMIP_NO_GRAPH")
    _assert_identical(outs["graph"], outs["eager"])


def test_host_batches_of_a_large_job_match_device_stack(ctx):
    """A job large enough for its host batches to overlap their copies and FFT with
    the previous batch's triple products (256x256, b=48, 128-frame batches): pushed
    in uneven chunks, added in one host call, and from a device stack, the outputs
    are identical."""
    frames = _frames("triple256", 256, 5 * 128 + 37, seed_offset=3)
    p = S.job_params(256, 256, 48, 128, frames.dtype, frames_per_batch=128)
    ref = S.Job(ctx, p)
    ref.add_frames(S.Frames(ctx, frames))
    ref.finalize()
    want = _outputs(ref)
    ref.close()
    pushed = S.Job(ctx, p)
    for a, b in ((0, 1), (1, 150), (150, 151), (151, 400), (400, len(frames))):
        pushed.push_frames_host(frames[a:b])
    pushed.finalize()
    _assert_identical(_outputs(pushed), want)
    pushed.close()
    added = S.Job(ctx, p)
    added.add_frames_host(frames)
    added.finalize()
    _assert_identical(_outputs(added), want)
    added.close()
    # closing a job whose pushed frames are still being copied
    left = S.Job(ctx, p)
    left.push_frames_host(frames[:200])
    left.close()


def test_released_memory_is_reused_without_changing_results(ctx):
    """A released job's device and pinned memory is kept for later jobs instead of
    going back to HIP, whose free waits for every stream of the device.  A job built
    from the blocks of released jobs that held other data gives the same outputs, bit
    for bit, and the kept memory still counts as free."""
    a = _frames("triple256", 64, 100, seed_offset=5)
    b = _frames("triple256", 64, 100, seed_offset=6)

    def run(fr, precision):
        j = S.Job(ctx, S.job_params(64, 64, 24, 32, fr.dtype, precision, frames_per_batch=32))
        j.push_frames_host(fr)
        j.finalize()
        out = _outputs(j)
        j.close()
        return out

    free0, _ = ctx.mem_info()
    for precision in (S.PRECISION_MIXED, S.PRECISION_FP64):
        want = run(a, precision)
        other = run(b, precision)
        assert not np.array_equal(other["bispectrum"], want["bispectrum"])
        _assert_identical(run(a, precision), want)
    free1, _ = ctx.mem_info()
    assert free1 >= free0 - (64 << 20)


def test_servers_made_and_destroyed_in_turn(ctx):
    """Each server makes one own-queue recovery stream per worker and destroys them
    with itself: servers made and destroyed in turn, each finalising a job, keep
    working and give the same result."""
    fr = _frames("binary128", 64, 12, seed_offset=9)
    prm = S.job_params(64, 64, 16, 24, fr.dtype, frames_per_batch=8)
    outs = []
    for i in range(6):
        srv = S.Server(ctx, max_active=2, workers=1 + i % 3)
        jid = srv.submit(prm)
        srv.push(jid, fr)
        srv.close_job(jid)
        assert srv.wait(jid, 120.0) == S.JOB_DONE
        outs.append(_outputs(srv.job(jid)))
        srv.release(jid)
        srv.close()
    for o in outs[1:]:
        _assert_identical(o, outs[0])


def test_push_mixing_rules(ctx):
    frames = _frames("binary128", 64, 8)
    job = S.Job(ctx, S.job_params(64, 64, 16, 24, frames.dtype, frames_per_batch=32))
    job.push_frames_host(frames[:5])
    with pytest.raises(S.SmipError) as e:
        job.add_frames_host(frames[5:])
    assert e.value.code == S.ERR_LOGIC
    job.finalize()
    with pytest.raises(S.SmipError):
        job.push_frames_host(frames)
    job.close()


def test_served_job_matches_oracle(ctx):
    frames = _frames("binary128", 128, 40)
    srv = S.Server(ctx)
    jid = srv.submit(S.job_params(128, 128, 32, 64, frames.dtype))
    for k in range(0, 40, 9):
        srv.push(jid, frames[k:k + 9])
    srv.close_job(jid)
    assert srv.wait(jid, 120.0) == S.JOB_DONE
    got = srv.job(jid).result(with_bispectrum=False)
    want = O.pipeline(frames.astype(np.float64), 32, 64)
    assert np.array_equal(got.shifts, want.shifts)
    assert np.array_equal(got.sum_image, want.sum_image)
    assert np.array_equal(got.flags, want.flags)
    err = np.linalg.norm(got.image - want.image) / np.linalg.norm(want.image)
    assert err < 1e-5, err
    info = srv.info(jid)
    assert info["frames_done"] == 40 and info["latency_ms"] >= info["tail_ms"] >= 0
    srv.release(jid)
    srv.close()


# jobs of uneven frame size, count, dtype and bispectrum radius
MIX = [("binary128", 64, 70, 16, 32), ("triple256", 128, 55, 24, 48), ("binary128", 128, 40, 32, 64),
       ("extended512", 64, 33, 20, 40), ("triple256", 256, 24, 48, 96), ("binary128", 32, 100, 12, 16),
       ("extended512", 128, 45, 16, 32)]


def _run_alone(ctx, frames, p):
    srv = S.Server(ctx, max_active=1, workers=1)
    jid = srv.submit(p)
    srv.push(jid, frames)
    srv.close_job(jid)
    assert srv.wait(jid, 120.0) == S.JOB_DONE
    out = _outputs(srv.job(jid))
    srv.release(jid)
    srv.close()
    return out


def test_result_is_the_same_alone_and_in_the_mix(ctx):
    data = []
    for i, (kind, N, n, b, p) in enumerate(MIX):
        fr = _frames(kind, N, n, seed_offset=100 + i)
        prm = S.job_params(N, N, b, p, fr.dtype, frames_per_batch=16 + 8 * i)
        data.append((fr, prm))
    alone = [_run_alone(ctx, fr, prm) for fr, prm in data]
    # each sequence against an FP64 reference: the oracle where it finishes in
    # seconds (N <= 128), the engine's FP64 path beyond (the oracle takes minutes
    # per 256x256 run on one core; the FP64 path is pinned to it elsewhere).
    # Shifts and flags exact; the image within max(1e-5, 20 x the reference's own
    # change when every frame pixel moves by one float32 ulp), so that sequences
    # whose phase recursion amplifies rounding (low-contrast extended objects,
    # few frames) are held to their conditioning, not to a fixed number.
    for (fr, prm), a in zip(data, alone):
        up = np.nextafter(fr.astype(np.float32), np.float32(np.inf))
        if prm.ncols <= 128:
            want = O.pipeline(fr.astype(np.float64), prm.bdepth, prm.reco_radius)
            moved = O.pipeline(up.astype(np.float64), prm.bdepth, prm.reco_radius)
            ref = "oracle"
        else:
            want = S.run_pipeline(ctx, fr, prm.bdepth, prm.reco_radius, precision=S.PRECISION_FP64,
                                  with_bispectrum=False)
            moved = S.run_pipeline(ctx, up, prm.bdepth, prm.reco_radius, precision=S.PRECISION_FP64,
                                   with_bispectrum=False)
            ref = "FP64 engine"
        norm = np.linalg.norm(want.image)
        sens = np.linalg.norm(moved.image - want.image) / norm
        err = np.linalg.norm(a["image"] - want.image) / norm
        print(f"{prm.ncols}x{prm.nrows} n={len(fr)} b={prm.bdepth} vs {ref}: image rel L2 {err:.3g}, "
              f"one-ulp sensitivity {sens:.3g}, bound {max(1e-5, 20 * sens):.3g}")
        assert np.array_equal(a["shifts"], want.shifts)
        assert np.array_equal(a["flags"], want.flags)
        assert err <= max(1e-5, 20 * sens)

    srv = S.Server(ctx, max_active=4, workers=4)
    ids = [None] * len(data)
    errors = []

    def client(i):
        try:
            fr, prm = data[i]
            ids[i] = srv.submit(prm)
            rng = np.random.default_rng(i)
            pos = 0
            while pos < len(fr):
                c = int(rng.integers(1, 23))
                srv.push(ids[i], fr[pos:pos + c])
                pos += c
            srv.close_job(ids[i])
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=client, args=(i,)) for i in range(len(data))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i, jid in enumerate(ids):
        assert srv.wait(jid, 300.0) == S.JOB_DONE, srv.info(jid)
        _assert_identical(_outputs(srv.job(jid)), alone[i])
        srv.release(jid)
    st = srv.stats()
    assert st["completed"] == len(data) and st["failed"] == 0 and st["bytes_in_use"] == 0
    assert st["finalized_in_groups"] == len(data) and 1 <= st["finalize_groups"] <= len(data)
    srv.close()


def test_admission_respects_budget_and_slots(ctx):
    fr = _frames("binary128", 128, 20)
    prm = S.job_params(128, 128, 32, 32, fr.dtype, frames_per_batch=16)
    fp = S.job_footprint(prm)
    srv = S.Server(ctx, hbm_budget=2 * fp + fp // 2, max_active=8, workers=4)
    ids = [srv.submit(prm) for _ in range(5)]
    st = srv.stats()
    assert st["active"] == 2 and st["waiting"] == 3 and st["bytes_in_use"] == 2 * fp
    for jid in ids:
        srv.push(jid, fr)
        srv.close_job(jid)
    peak = 0
    for jid in ids:
        assert srv.wait(jid, 120.0) == S.JOB_DONE
        peak = max(peak, srv.stats()["bytes_in_use"])
        srv.release(jid)  # frees the slot for the next waiting job
    assert peak <= 2 * fp
    st = srv.stats()
    assert st["completed"] == 5 and st["waiting"] == 0 and st["active"] == 0
    # a job that can never fit is refused at submit
    big = S.job_params(512, 512, 64, 128, np.uint16)
    with pytest.raises(S.SmipError) as e:
        srv.submit(big)
    assert e.value.code == S.ERR_NO_MEMORY
    srv.close()


def test_server_errors(ctx):
    srv = S.Server(ctx, max_active=2, workers=2)
    fr = _frames("binary128", 64, 4)
    prm = S.job_params(64, 64, 16, 16, fr.dtype)
    jid = srv.submit(prm)
    with pytest.raises(S.SmipError) as e:
        srv.close_job(jid)  # no frames yet
    assert e.value.code == S.ERR_RUNTIME
    srv.push(jid, fr)
    srv.close_job(jid)
    with pytest.raises(S.SmipError) as e:
        srv.push(jid, fr)
    assert e.value.code == S.ERR_LOGIC
    with pytest.raises(S.SmipError) as e:
        srv.wait(jid + 1000)
    assert e.value.code == S.ERR_INVALID_ARGUMENT
    bad = S.job_params(4096, 48, 16, 16, fr.dtype)  # a side beyond the 2048 the transforms hold
    with pytest.raises(S.SmipError):
        srv.submit(bad)
    assert srv.wait(jid, 60.0) == S.JOB_DONE
    srv.release(jid)
    with pytest.raises(S.SmipError):
        srv.info(jid)
    srv.close()


def test_captures_concurrent_with_allocation_and_copies(ctx):
    """Graph captures (second full batch of every job) overlap with other threads'
    job creation, plan uploads, frame uploads and frees; every job must succeed
    and equal its result computed alone.  HIP refuses legacy-stream copies while a
    stream captures, so the library never issues them and serialises capture
    against allocation."""
    frames = _frames("binary128", 64, 40, seed_offset=11)
    p = S.job_params(64, 64, 16, 20, frames.dtype, frames_per_batch=8)
    ref = S.Job(ctx, p)
    ref.add_frames(S.Frames(ctx, frames))
    ref.finalize()
    want = ref.image()
    ref.close()
    errors, bad = [], []

    def worker(t):
        try:
            for k in range(6):
                job = S.Job(ctx, p)
                dev = S.Frames(ctx, frames)  # an upload and an allocation between captures
                for c in range(0, 40, 5):
                    job.push_frames_host(frames[c:c + 5])
                job.finalize()
                if not np.array_equal(job.image(), want):
                    bad.append((t, k))
                dev.close()
                job.close()
        except Exception as e:  # reported below
            errors.append(repr(e))

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors[:3]
    assert not bad, bad


def test_finalize_many_matches_single(ctx):
    """Grouped phase recovery (one launch per stage and window index for all jobs)
    gives each job exactly what its own finalize gives, and what a group of that
    job alone gives."""
    spec = [("binary128", 64, 30, 16, 32), ("triple256", 128, 20, 32, 64), ("extended512", 64, 25, 20, 40),
            ("binary128", 32, 40, 12, 16), ("triple256", 128, 18, 24, 24)]
    data = [(_frames(kind, N, n, seed_offset=40 + i), S.job_params(N, N, b, p, np.dtype(
        {"binary128": np.uint8, "triple256": np.uint16, "extended512": np.float32}[kind]), frames_per_batch=16))
        for i, (kind, N, n, b, p) in enumerate(spec)]
    single, calls = [], []
    for fr, prm in data:
        job = S.Job(ctx, prm)
        job.add_frames(S.Frames(ctx, fr))
        job.finalize()
        single.append(_outputs(job))
        calls.append(job.stats()["reco_calls"])
        job.close()
    jobs = []
    for fr, prm in data:
        job = S.Job(ctx, prm)
        job.push_frames_host(fr)
        jobs.append(job)
    S.finalize_many(jobs)
    for k, (job, (fr, prm)) in enumerate(zip(jobs, data)):
        got = _outputs(job)
        alone = S.Job(ctx, prm)
        alone.push_frames_host(fr)
        S.finalize_many([alone])
        _assert_identical(got, _outputs(alone))
        alone.close()
        _assert_identical(got, single[k])
        assert job.stats()["reco_calls"] == calls[k]
        job.close()


def test_finalize_many_rules(ctx):
    fr = _frames("binary128", 64, 8)
    prm = S.job_params(64, 64, 16, 16, fr.dtype)
    a, b = S.Job(ctx, prm), S.Job(ctx, prm)
    a.push_frames_host(fr)
    S.finalize_many([])  # nothing to do
    with pytest.raises(S.SmipError) as e:
        S.finalize_many([a, a])
    assert e.value.code == S.ERR_INVALID_ARGUMENT
    with pytest.raises(S.SmipError) as e:
        S.finalize_many([a, b])  # b has no frames: nothing is changed
    assert e.value.code == S.ERR_RUNTIME
    S.finalize_many([a])
    with pytest.raises(S.SmipError) as e:
        S.finalize_many([a])
    assert e.value.code == S.ERR_LOGIC
    a.close()
    b.close()


def test_add_frames_many_matches_single_jobs(ctx):
    """smip_job_add_frames_many: one launch per stage for the batches of many jobs of
    uneven frame side, frame count, depth, sample type and tile width -- plus an
    FP64 job and a 96x64 job, which take their own launches -- gives each job
    exactly what smip_job_add_frames gives it, also when the frames arrive over
    several grouped calls (both sides recover their phases in one group)."""
    specs = [("binary128", 64, 150, 16, 0, S.PRECISION_MIXED), ("triple256", 128, 90, 24, 9, S.PRECISION_MIXED),
             ("extended512", 64, 40, 33, 0, S.PRECISION_MIXED), ("triple256", 256, 70, 48, 25, S.PRECISION_MIXED),
             ("binary128", 32, 13, 12, 17, S.PRECISION_MIXED), ("extended512", 128, 30, 20, 13, S.PRECISION_MIXED),
             ("triple256", 64, 25, 16, 0, S.PRECISION_FP64)]
    stacks, params = [], []
    for k, (kind, N, n, b, vy, prec) in enumerate(specs):
        fr = _frames(kind, N, n, seed_offset=k)
        stacks.append(S.Frames(ctx, fr))
        params.append(S.job_params(N, N, b, 2 * b, fr.dtype, prec, frames_per_batch=32, vy_tile=vy))
    odd = np.random.default_rng(5).integers(0, 255, size=(20, 64, 96), dtype=np.uint8)
    stacks.append(S.Frames(ctx, odd))
    params.append(S.job_params(96, 64, 12, 20, np.uint8, S.PRECISION_MIXED, frames_per_batch=8))

    # the pieces each call gets: whole stacks, then uneven pieces with some jobs idle
    rng = np.random.default_rng(11)
    pos = [0] * len(stacks)
    rounds = []
    while any(p < s.shape[0] for p, s in zip(pos, stacks)):
        cnt = [int(min(s.shape[0] - p, rng.integers(0, 70))) for p, s in zip(pos, stacks)]
        rounds.append(list(zip(pos, cnt)))
        pos = [p + c for p, c in zip(pos, cnt)]
    for split in ([[(0, s.shape[0]) for s in stacks]], rounds):
        want = []
        for k, (st, p) in enumerate(zip(stacks, params)):
            j = S.Job(ctx, p)
            for piece in split:
                if piece[k][1]:
                    j.add_frames(st, piece[k][0], piece[k][1])
            S.finalize_many([j])
            want.append(_outputs(j))
            j.close()
        jobs = [S.Job(ctx, p) for p in params]
        for piece in split:
            S.add_frames_many(jobs, stacks, [f for f, _ in piece], [c for _, c in piece])
        S.finalize_many(jobs)
        for j, w in zip(jobs, want):
            _assert_identical(_outputs(j), w)
            j.close()
