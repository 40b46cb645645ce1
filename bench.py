#!/usr/bin/env python
"""Bispectrum frames/s of the MI355X speckle-masking engine.

Workload (the metric's configuration): triple256 -- 2000 synthetic speckle
frames of 256x256 uint16 (12-bit), triple-star object, pupil 64 px, D/r0 = 20,
1e5 photons/frame; bispectrum depth b = 48, phase radius p = 128.

A step is one pass of the hot path (FFT + |F|^2 + triple-product accumulation)
over the whole 2000-frame stack, already resident in HBM, in the batched
mixed-precision form (FP32 triple products per batch folded into the FP64
bispectrum).  Besides the JSON contract fields the line carries the plain
one-frame-at-a-time FP64 GPU path's frames/s and the speedup over it, the p50 /
p99 latency of complete jobs (accumulate, shift-and-add, phase recovery,
synthesis), the image error of the mixed path against the FP64 GPU path, the
roofline position of the triple-product kernel and the CPU baseline.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import threading
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "speckle-stacker_amd", "python"))

import smip_amd as S  # noqa: E402

VALU_FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak (dense)
HBM_PEAK_GBS = 8000.0


_T0 = time.perf_counter()


def log(msg: str) -> None:
    print(f"[bench +{time.perf_counter() - _T0:.1f}s] {msg}", file=sys.stderr, flush=True)


def traffic_from_profile(path: str):
    """HBM bytes per launch of the triple-product kernel from a committed
    rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE lines), if present."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(frames: np.ndarray, b: int, p: float, seconds: float) -> dict:
    """The reference's per-frame CPU path (FP64 FFT, complex<float> accumulate,
    smip-cli/main.cpp:256-305) as restated by the oracle, single thread; then
    its phase recovery (include/phasereco.h:33-123) once on the sample's
    normalised bispectrum."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # the checker, used here only as the CPU baseline

    f64 = frames.astype(np.float64)
    t0 = time.perf_counter()
    O.power_and_bispectrum(f64[:1], b, bispec_f32=True, fast=True)
    one = time.perf_counter() - t0
    k = int(max(1, min(len(f64) - 1, seconds / max(one, 1e-3))))
    t0 = time.perf_counter()
    B, _ = O.power_and_bispectrum(f64[1:1 + k], b, bispec_f32=True, fast=True)
    dt = time.perf_counter() - t0
    N = frames.shape[-1]
    B /= np.complex64(k)
    t0 = time.perf_counter()
    reco = O.reconstruct_phases(O.desc(N, N, b), B, N, N, p)
    dreco = time.perf_counter() - t0
    return {"value": k / dt, "unit": "frames/s", "cores": 1, "kind": "port",
            "sample": f"{k} of the workload's frames through the oracle's FP64 FFT + complex<float> "
                      f"accumulate (-O3 -ffast-math), 1 thread, {dt:.1f} s on {platform.processor() or 'host'}",
            "phase_recovery_s": round(dreco, 3), "phase_recovery_calls": int(reco.stats["calls"]),
            "phase_recovery_note": "the oracle's calc_phase over the (2h+1)^2 candidates of each visit "
                                   "(the reference scans all N^2 u per call), 1 thread"}


def streamed_leg(ctx, frames: np.ndarray, params) -> dict:
    """The metric's accumulate with the frames in pageable host memory, so the
    PCIe copy is inside the timed region: one smip_job_add_frames_host call for
    the stack, and smip_job_push_frames_host in camera-sized chunks of 100."""
    batch = S.resolve_params(params).frames_per_batch
    # at least 8 batches per timing, so that filling the copy / FFT / triple-product
    # pipeline at the start is a small part of it
    reps = max(1, -(-8 * batch // len(frames)))
    stack = np.concatenate([frames] * reps) if reps > 1 else frames
    out = {"frames": len(stack), "frames_per_batch": batch, "host_memory": "pageable (numpy)"}
    j = S.Job(ctx, params)
    j.add_frames_host(frames[:batch])  # allocations
    ctx.synchronize()
    t0 = time.perf_counter()
    j.add_frames_host(stack)
    ctx.synchronize()
    out["add_frames_host_frames_per_s"] = round(len(stack) / (time.perf_counter() - t0), 1)
    j.close()
    del stack
    def push(job, count, chunk=100):  # count frames, cycling through the stack
        done = 0
        while done < count:
            s0 = done % len(frames)
            c = min(chunk, count - done, len(frames) - s0)
            job.push_frames_host(frames[s0:s0 + c])
            done += c

    whole = 8 * batch  # whole batches: nothing left pending
    j = S.Job(ctx, params)
    push(j, 2 * batch)  # the eager first batch and the graph capture
    j.synchronize()
    t0 = time.perf_counter()
    push(j, whole)
    j.synchronize()
    out["push_frames_host_frames_per_s"] = round(whole / (time.perf_counter() - t0), 1)
    out["push_frames_timed"] = whole
    out["push_chunk"] = 100
    j.close()
    return out


def workload_leg(ctx, name: str, nframes: int, passes: int, latency_runs: int) -> dict:
    """Another configuration of the reference's: accumulate frames/s on a
    device-resident sample of its stack (mixed path), and the latency of complete
    jobs (accumulate, shift-and-add, phases, synthesis) on that sample."""
    W = S.WORKLOADS[name]
    n = min(nframes, W["n"])
    fr = S.synth_frames(S.synth_params(name), 0, n, threads=min(16, os.cpu_count() or 1))
    dev = S.Frames(ctx, fr)
    p = S.job_params(W["N"], W["N"], W["b"], W["p"], W["dtype"], S.PRECISION_MIXED, saa=False)
    job = S.Job(ctx, p)
    job.add_frames(dev)  # warm-up
    ctx.synchronize()
    st0 = job.stats()
    t0 = time.perf_counter()
    for _ in range(passes):
        job.add_frames(dev)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    st1 = job.stats()
    job.close()
    tile_s = (st1["tile_ms"] - st0["tile_ms"]) * 1e-3
    flops = 14.0 * st1["triples_per_frame"] * n * passes
    lat = []
    stages = None
    for _ in range(latency_runs):
        t = time.perf_counter()
        jl = S.Job(ctx, S.job_params(W["N"], W["N"], W["b"], W["p"], W["dtype"], S.PRECISION_MIXED, saa=True))
        jl.add_frames(dev)
        jl.finalize()
        lat.append((time.perf_counter() - t) * 1e3)
        stages = jl.stats()
        jl.close()
    return {"workload": f"{name}: {n} of {W['n']} frames {W['N']}x{W['N']} {np.dtype(W['dtype']).name}, "
                        f"b={W['b']}, p={W['p']}",
            "frames_per_s": round(n * passes / dt, 1), "frames_timed": n * passes,
            "tile_tflops": round(flops / tile_s / 1e12, 2) if tile_s > 0 else None,
            "job_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 2), "runs": latency_runs,
                               "frames_per_job": n},
            "stages_ms": {k: round(stages[k], 2) for k in ("accumulate_ms", "saa_ms", "reco_ms", "synth_ms")}}


def run_latency_jobs(ctx, dev, W, njobs: int, batch: int, vy: int, grouped: bool = False,
                     stagger_ms: float = 0.0):
    """njobs complete reconstructions; job i arrives i * stagger_ms after the
    first and one thread drives it.  Each job finalises on its own, or (grouped)
    hands its accumulated bispectrum to one finalizer thread that recovers the
    phases of every job ready at that moment in one group
    (smip_job_finalize_many), as the serving loop does.  Latency is from a job's
    own arrival to its completion."""
    lat = [0.0] * njobs
    errs = []
    stats = [None] * njobs
    ready = []
    cv = threading.Condition()
    t_start = time.perf_counter()
    t_arrive = [t_start + i * stagger_ms * 1e-3 for i in range(njobs)]

    def worker(i):
        try:
            delay = t_arrive[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            p = S.job_params(W["N"], W["N"], W["b"], W["p"], W["dtype"], S.PRECISION_MIXED, saa=True,
                             frames_per_batch=batch, vy_tile=vy)
            job = S.Job(ctx, p)
            job.add_frames(dev)
            if grouped:
                job.synchronize()
                with cv:
                    ready.append((i, job))
                    cv.notify()
                return
            job.finalize()
            lat[i] = (time.perf_counter() - t_arrive[i]) * 1e3
            stats[i] = job.stats()
            job.close()
        except Exception as e:  # reported, not hidden
            errs.append(repr(e))
            if grouped:
                with cv:
                    ready.append((i, None))
                    cv.notify()

    def finalizer():
        done = 0
        while done < njobs:
            with cv:
                cv.wait_for(lambda: ready)
                group = ready[:]
                ready.clear()
            jobs = [j for _, j in group if j is not None]
            try:
                if jobs:
                    S.finalize_many(jobs)
            except Exception as e:
                errs.append(repr(e))
            t = time.perf_counter()
            for i, j in group:
                lat[i] = (t - t_arrive[i]) * 1e3
                if j is not None:
                    stats[i] = j.stats()
                    j.close()
            done += len(group)

    th = [threading.Thread(target=worker, args=(i,)) for i in range(njobs)]
    if grouped:
        th.append(threading.Thread(target=finalizer))
    for t in th:
        t.start()
    for t in th:
        t.join()
    if errs:
        raise RuntimeError("; ".join(errs))
    return lat, stats


def serve_leg(ctx, seconds: float, rate: float, seed: int, chunk: int = 100, camera_hz: float = 1000.0,
              max_active: int = 64, drain_s: float = 90.0, pushers: int = 4, workers: int = 0) -> dict:
    """The request loop under Poisson arrivals: jobs drawn as the multi-sequence
    mix (frame side from {64, 128, 256, 512}, frame count log-uniform in
    [40, 5000], b uniform in [16, 64], p = 2b, object and sample type per job),
    each streaming its frames in chunks of `chunk` at `camera_hz` per job.
    Frames cycle through a pool of pre-generated frames per (side, type).
    `pushers` client threads push the jobs' chunks (job i by thread i mod
    pushers), as separate cameras would: one thread's pushes from pageable memory
    reach about 8.5 GB/s, about 40 k frames/s of this mix.  The run stops
    `drain_s` after the last arrival; jobs not finished by then are reported as
    unfinished."""
    import heapq

    rng = np.random.default_rng(seed)
    pools = frame_pools(ctx, host=True)
    # arrivals
    jobs = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        q = draw_sequences(rng, 1)[0]
        q["t"] = t
        jobs.append(q)
    srv = S.Server(ctx, max_active=max_active, workers=workers)
    pending = [[] for _ in range(pushers)]  # per pusher: (time, seq, job index, first frame)
    for i, j in enumerate(jobs):
        heapq.heappush(pending[i % pushers], (j["t"], i, i, 0))
    closed = []
    results, failures = {}, {}
    stop = threading.Event()

    def reaper():
        # releases every closed job as soon as it is done, in whatever order they
        # finish: a job's slot and memory return only at release, so waiting on the
        # oldest closed job while it waits for a slot held by finished ones would
        # stall the server
        k = 0
        outstanding = []
        while True:
            while k < len(closed):
                outstanding.append(closed[k])
                k += 1
            if not outstanding:
                if stop.is_set():
                    return
                time.sleep(0.005)
                continue
            still = []
            for i, jid in outstanding:
                try:
                    st = srv.wait(jid, 0.0)
                    if st not in (S.JOB_DONE, S.JOB_FAILED):
                        still.append((i, jid))
                        continue
                    results[i] = srv.info(jid)
                except Exception as e:  # reported, not hidden
                    failures[i] = repr(e)
                try:
                    srv.release(jid)
                except Exception:
                    pass
            outstanding = still
            if stop.is_set():
                return
            time.sleep(0.002)

    th = threading.Thread(target=reaper)
    th.start()
    t0 = time.perf_counter()
    pushed = [0] * pushers
    lag = [0.0] * pushers
    deadline = t0 + seconds + drain_s

    def pusher(q):
        heap = pending[q]
        seq = len(jobs)
        try:
            while heap and time.perf_counter() < deadline:
                when, _, i, first = heapq.heappop(heap)
                delay = t0 + when - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                lag[q] = max(lag[q], time.perf_counter() - t0 - when)
                j = jobs[i]
                pool = pools[(j["N"], j["kind"])]
                if first == 0:
                    j["id"] = srv.submit(S.job_params(j["N"], j["N"], j["b"], j["p"], pool.dtype))
                cnt = min(chunk, j["n"] - first)
                pos = 0
                while pos < cnt:  # contiguous pieces of the cycled pool, pushed without a host copy here
                    a = (j["off"] + first + pos) % len(pool)
                    c = min(cnt - pos, len(pool) - a)
                    srv.push(j["id"], pool[a:a + c])
                    pos += c
                pushed[q] += cnt
                if first + cnt >= j["n"]:
                    srv.close_job(j["id"])
                    closed.append((i, j["id"]))
                else:
                    seq += 1
                    heapq.heappush(heap, (when + cnt / camera_hz, seq, i, first + cnt))
        except Exception as e:  # reported, not hidden
            failures[-1 - q] = repr(e)

    pts = [threading.Thread(target=pusher, args=(q,)) for q in range(pushers)]
    for p_ in pts:
        p_.start()
    last_log = t0
    served = None  # (frames processed, seconds) at the end of the arrival window
    while any(p_.is_alive() for p_ in pts):
        time.sleep(0.25)
        now = time.perf_counter()
        if served is None and now - t0 >= seconds:
            served = (srv.stats()["frames"], now - t0)
        if now - last_log >= 5.0:
            st = srv.stats()
            log(f"serving t={now - t0:.0f}s: submitted {st['submitted']}, done {st['completed']}, active "
                f"{st['active']}, waiting {st['waiting']}, frames {st['frames']}/{sum(pushed)}, "
                f"lag {max(lag):.2f}s")
            last_log = now
    for p_ in pts:
        p_.join()
    frames_pushed = sum(pushed)
    if served is None:
        served = (srv.stats()["frames"], time.perf_counter() - t0)
    while time.perf_counter() < deadline and len(results) + len(failures) < len(closed):
        time.sleep(1.0)
        st = srv.stats()
        log(f"serving drain: done {st['completed']}/{len(jobs)}, active {st['active']}, waiting {st['waiting']}")
    stop.set()
    th.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    done = sorted(results)
    lat = np.array([results[i]["latency_ms"] for i in done]) if done else np.zeros(1)
    tail = np.array([results[i]["tail_ms"] for i in done]) if done else np.zeros(1)
    queue = np.array([results[i]["queue_ms"] for i in done]) if done else np.zeros(1)
    pct = lambda a, q: round(float(np.percentile(a, q)), 2)
    frames_done = int(sum(jobs[i]["n"] for i in done))
    push_end_lag = round(max(lag), 2)
    return {"arrival": f"Poisson, {rate} jobs/s for {seconds} s", "jobs": len(jobs), "completed": len(done),
            "window": f"{seconds} s of arrivals (the configured serving run is 600 s; the same rate, mix and "
                      "camera model over a shorter window)",
            "failed": len(failures), "unfinished": len(jobs) - len(done) - len(failures),
            "failures": list(failures.values())[:3], "frames_of_completed_jobs": frames_done,
            "frames_pushed": int(frames_pushed), "wall_s": round(wall, 2),
            "frames_per_s": round(st["frames"] / wall, 1),
            "served_frames_per_s_in_window": round(served[0] / served[1], 1), "max_active": max_active,
            "workers": workers or 8,
            "latency_ms": {"p50": pct(lat, 50), "p95": pct(lat, 95), "p99": pct(lat, 99)},
            "after_last_frame_ms": {"p50": pct(tail, 50), "p95": pct(tail, 95), "p99": pct(tail, 99)},
            "admission_wait_ms": {"p50": pct(queue, 50), "p99": pct(queue, 99)},
            "camera": f"chunks of {chunk} frames at {camera_hz:.0f} frames/s per job, {pushers} pushing threads",
            "push_lag_s": push_end_lag,
            "mix": "side {64,128,256,512}, frames log-uniform 40-5000, b U[16,64], p=2b, "
                   "binary/u8, triple/u16, extended/f32"}


def serve_sweep(ctx, rates: list, seconds: float, plain_fps: float | None) -> dict:
    """Serving capacity: serve_leg at rising Poisson rates, each over `seconds` of
    arrivals with up to 60 s to drain.  Below capacity the completed frames/s
    follow the offered load and the after-last-frame latency stays flat; past it
    the jobs queue, that latency grows with the run and jobs are left unfinished.
    The knee is the highest rate whose jobs all finish with an after-last-frame p99
    under 1 s while the pushing threads keep to the arrival schedule (lag < 1 s)."""
    runs = []
    for k, rate in enumerate(rates):
        r = serve_leg(ctx, seconds, rate, seed=0x5EC1E100 + k, drain_s=60.0)
        offered = r["frames_pushed"] / seconds
        runs.append({"rate_jobs_per_s": rate, "jobs": r["jobs"], "completed": r["completed"],
                     "unfinished": r["unfinished"], "failed": r["failed"], "push_lag_s": r["push_lag_s"],
                     "offered_frames_per_s": round(offered, 1), "frames_per_s": r["frames_per_s"],
                     "served_frames_per_s_in_window": r["served_frames_per_s_in_window"],
                     "after_last_frame_ms": r["after_last_frame_ms"], "latency_ms": r["latency_ms"],
                     "admission_wait_ms": r["admission_wait_ms"]})
        log(f"serving sweep {rate} jobs/s: {runs[-1]}")
    ok = [x for x in runs if x["unfinished"] == 0 and x["failed"] == 0 and x["after_last_frame_ms"]["p99"] < 1000.0
          and x["push_lag_s"] < 1.0]
    knee = max(ok, key=lambda x: x["rate_jobs_per_s"]) if ok else None
    out = {"window_s_per_rate": seconds, "runs": runs,
           "knee_rate_jobs_per_s": knee["rate_jobs_per_s"] if knee else None,
           "capacity_frames_per_s": knee["offered_frames_per_s"] if knee else None,
           "rule": "knee = highest rate with every job finished, after-last-frame p99 < 1 s and the pushing "
                   "threads within 1 s of the arrival schedule",
           # above the knee the server runs flat out: the frames it has processed by the end of the
           # arrival window, per second, no longer follow the offered load
           "saturated_frames_per_s": max(x["served_frames_per_s_in_window"] for x in runs) if runs else None}
    if knee and plain_fps:
        out["plain_fp64_serial_frames_per_s"] = round(plain_fps, 1)
        out["capacity_vs_plain_fp64_serial"] = round(knee["offered_frames_per_s"] / plain_fps, 2)
        out["saturated_vs_plain_fp64_serial"] = round(out["saturated_frames_per_s"] / plain_fps, 2)
        out["plain_note"] = ("the plain FP64 path run one sequence at a time over the same mix (multi-sequence "
                             "leg's device-time estimate); its images are the FP64 reference the mixed path "
                             "is held to (image rel L2 within 1e-5, or the FP64 path's own one-ulp sensitivity "
                             "where that is larger)")
    return out


def draw_sequences(rng, n: int) -> list:
    """Sequences of the multi-sequence mix: side {64,128,256,512}, frame count
    log-uniform in [40, 5000], b uniform in [16, 64] (at most 2N-1), p = 2b,
    object and sample type per sequence."""
    kinds = ["binary128", "triple256", "extended512"]
    out = []
    for _ in range(n):
        N = int(rng.choice([64, 128, 256, 512]))
        b = min(int(rng.integers(16, 65)), 2 * N - 1)
        out.append(dict(N=N, kind=kinds[int(rng.integers(0, 3))],
                        n=int(round(np.exp(rng.uniform(np.log(40), np.log(5000))))), b=b, p=2 * b,
                        off=int(rng.integers(0, 64))))
    return out


def frame_pools(ctx, host: bool, length: int = 128):
    """64 distinct frames per (side, type); on the device repeated to `length`
    frames, so that frames off .. off + n - 1 (off < 64) of a pool are the
    sequence's frames off, off + 1, ... cycling through the 64."""
    pools = {}
    for N in (64, 128, 256, 512):
        for kind in ("binary128", "triple256", "extended512"):
            prm = S.synth_params(kind)
            prm.ncols = prm.nrows = N
            prm.pupil_diameter = max(8.0, N / 4)
            f = S.synth_frames(prm, 0, 64, threads=min(16, os.cpu_count() or 1))
            if host:
                pools[(N, kind)] = f
                continue
            dev = S.Frames(ctx, dtype=f.dtype, shape=(length, N, N))
            for at in range(0, length, 64):
                dev.upload(at, f[:min(64, length - at)])
            pools[(N, kind)] = dev
    return pools


def batch_leg(ctx, nseq: int, threads: int, plain_sample: int) -> dict:
    """nseq independent sequences processed together: `threads` host threads each
    drive one sequence at a time through its own job and stream, frames from
    device-resident pools (64 distinct frames per side and type, cycled)."""
    rng = np.random.default_rng(0x5EC1E004)
    seqs = draw_sequences(rng, nseq)
    pools = frame_pools(ctx, host=False, length=max(q["n"] for q in seqs) + 64)
    lat = [0.0] * nseq
    errs = []
    nxt = [0]
    lock = threading.Lock()

    def run(q, precision, limit=None):
        dev = pools[(q["N"], q["kind"])]
        job = S.Job(ctx, S.job_params(q["N"], q["N"], q["b"], q["p"], dev.dtype, precision))
        n = q["n"] if limit is None else min(limit, q["n"])
        job.add_frames(dev, q["off"], n)
        job.finalize()
        st = job.stats()
        job.close()
        return st

    ready = []  # (sequence, job) accumulated and waiting for phase recovery
    cv = threading.Condition()
    groups = []

    # longest first: the batch's last sequences to start are its shortest, so the
    # makespan does not end on one large sequence started late (the triple products
    # grow about as N^2 b^2 per frame)
    order = sorted(range(nseq), key=lambda k: -seqs[k]["n"] * (seqs[k]["N"] * seqs[k]["b"]) ** 2)

    def worker():
        while True:
            with lock:
                if nxt[0] >= nseq:
                    return
                i = order[nxt[0]]
                nxt[0] += 1
            try:
                q = seqs[i]
                dev = pools[(q["N"], q["kind"])]
                job = S.Job(ctx, S.job_params(q["N"], q["N"], q["b"], q["p"], dev.dtype, S.PRECISION_MIXED))
                job.add_frames(dev, q["off"], q["n"])  # batch by batch, FFT of the next beside the current
                job.synchronize()
                with cv:
                    ready.append((i, job))
                    cv.notify()
            except Exception as e:  # reported, not hidden
                errs.append(repr(e))
                with cv:
                    ready.append((i, None))
                    cv.notify()

    def finalizer():
        done = 0
        while done < nseq:
            with cv:
                cv.wait_for(lambda: ready)
                batch = ready[:]
                ready.clear()
            jobs = [j for _, j in batch if j is not None]
            try:
                if jobs:
                    S.finalize_many(jobs)  # every sequence ready now recovers its phases in one group
                    groups.append(len(jobs))
            except Exception as e:
                errs.append(repr(e))
            t = (time.perf_counter() - t0) * 1e3
            for i, j in batch:
                lat[i] = t
                if j is not None:
                    j.close()
            done += len(batch)

    for q in seqs[:4]:  # warm-up: module load, plans
        run(q, S.PRECISION_MIXED, limit=8)
    t0 = time.perf_counter()
    th = [threading.Thread(target=worker) for _ in range(threads)] + [threading.Thread(target=finalizer)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    if errs:
        raise RuntimeError("; ".join(errs))
    total = sum(q["n"] for q in seqs)
    # plain serial FP64 path: one sequence at a time, first plain_sample frames of
    # each, scaled to its full count (extrapolation from the timed sample)
    est = 0.0
    for q in seqs:
        m = min(plain_sample, q["n"])
        st = run(q, S.PRECISION_FP64, limit=m)
        # per-frame stages scale with the frame count, recovery and synthesis do not
        est += ((st["accumulate_ms"] + st["saa_ms"]) * q["n"] / m + st["reco_ms"] + st["synth_ms"]) * 1e-3
    lat = np.array(lat)
    return {"sequences": nseq, "frames": total, "host_threads": threads, "makespan_s": round(wall, 3),
            "phase_recovery_groups": groups,
            "frames_per_s": round(total / wall, 1),
            "sequence_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1),
                                    "p99": round(float(np.percentile(lat, 99)), 1)},
            "plain_fp64_serial_s_estimate": round(est, 2),
            "speedup_vs_plain_fp64_serial": round(est / wall, 2),
            "plain_estimate": f"device time of the plain FP64 path, one sequence at a time: accumulate and "
                              f"shift-and-add timed on the first {plain_sample} frames of each sequence and scaled "
                              "to its frame count, plus its phase recovery and synthesis",
            "mix": "side {64,128,256,512}, frames log-uniform 40-5000, b U[16,64], p=2b, "
                   "binary/u8, triple/u16, extended/f32; frames cycle through 64 per side and type",
            "order": "longest first (frames x N^2 b^2)",
            "form": "16 host threads, each driving one sequence at a time through smip_job_add_frames on the "
                    "sequence's own streams; grouped phase recovery"}


def batch_leg_grouped(ctx, nseq: int, batches_per_call: int = 1, window: int = 0) -> dict:
    """The multi-sequence mix of batch_leg through smip_job_add_frames_many: one
    host thread advances every unfinished sequence by its next batch in one call,
    so each stage is one launch for all of them; sequences whose frames are all
    in go to one finalizer thread, which recovers the phases of every sequence
    then ready in one group."""
    rng = np.random.default_rng(0x5EC1E004)
    seqs = draw_sequences(rng, nseq)  # the same sequences as batch_leg
    pools = frame_pools(ctx, host=False, length=max(q["n"] for q in seqs) + 64)
    lat = [0.0] * nseq
    errs = []
    ready = []
    cv = threading.Condition()
    groups = []
    calls = [0]
    window = window or nseq  # sequences advanced together at most

    def finalizer():
        done = 0
        while done < nseq:
            with cv:
                cv.wait_for(lambda: ready)
                batch = ready[:]
                ready.clear()
            jobs = [j for _, j in batch if j is not None]
            try:
                if jobs:
                    S.finalize_many(jobs)
                    groups.append(len(jobs))
            except Exception as e:  # reported, not hidden
                errs.append(repr(e))
            t = (time.perf_counter() - t0) * 1e3
            for i, j in batch:
                lat[i] = t
                if j is not None:
                    j.close()
            done += len(batch)

    warm = [S.Job(ctx, S.job_params(q["N"], q["N"], q["b"], q["p"], pools[(q["N"], q["kind"])].dtype,
                                    S.PRECISION_MIXED)) for q in seqs[:4]]
    S.add_frames_many(warm, [pools[(q["N"], q["kind"])] for q in seqs[:4]], [0] * 4, [8] * 4)
    for j in warm:
        j.close()
    pos = [0] * nseq
    batch = [S.resolve_params(S.job_params(q["N"], q["N"], q["b"], q["p"], pools[(q["N"], q["kind"])].dtype,
                                           S.PRECISION_MIXED)).frames_per_batch for q in seqs]
    # jobs are created on a thread of their own (their device allocations are
    # synchronous) while the driver already advances the ones made, longest first
    # as in batch_leg
    order = sorted(range(nseq), key=lambda k: -seqs[k]["n"] * (seqs[k]["N"] * seqs[k]["b"]) ** 2)
    jobs = [None] * nseq
    made = []
    made_cv = threading.Condition()

    def creator():
        for i in order:
            q = seqs[i]
            try:
                j = S.Job(ctx, S.job_params(q["N"], q["N"], q["b"], q["p"], pools[(q["N"], q["kind"])].dtype,
                                            S.PRECISION_MIXED))
            except Exception as e:
                errs.append(repr(e))
                j = None
            with made_cv:
                jobs[i] = j
                made.append(i)
                made_cv.notify()

    t0 = time.perf_counter()
    th = threading.Thread(target=finalizer)
    th.start()
    mk = threading.Thread(target=creator)
    mk.start()
    taken = 0
    active = []
    sent = set()
    try:
        while True:
            with made_cv:
                if not active and taken < nseq:
                    made_cv.wait_for(lambda: len(made) > taken)
                new = made[taken:taken + max(0, window - len(active))]
                taken += len(new)
            failed = [i for i in new if jobs[i] is None]
            if failed:
                sent.update(failed)
                with cv:
                    ready.extend((i, None) for i in failed)
                    cv.notify()
            active += [i for i in new if jobs[i] is not None]
            if not active:
                if taken >= nseq:
                    break
                continue
            firsts = [seqs[i]["off"] + pos[i] for i in active]
            counts = [min(batches_per_call * batch[i], seqs[i]["n"] - pos[i]) for i in active]
            S.add_frames_many([jobs[i] for i in active], [pools[(seqs[i]["N"], seqs[i]["kind"])] for i in active],
                              firsts, counts)
            calls[0] += 1
            finished = []
            for i, c in zip(active, counts):
                pos[i] += c
                if pos[i] >= seqs[i]["n"]:
                    finished.append(i)
            if finished:
                sent.update(finished)
                with cv:
                    ready.extend((i, jobs[i]) for i in finished)
                    cv.notify()
            active = [i for i in active if pos[i] < seqs[i]["n"]]
    except Exception as e:
        errs.append(repr(e))
        mk.join()
        with cv:  # every sequence not yet handed over ends now (its job closed unfinished)
            ready.extend((i, jobs[i]) for i in range(nseq) if i not in sent)
            cv.notify()
    mk.join()
    th.join()
    wall = time.perf_counter() - t0
    if errs:
        raise RuntimeError("; ".join(errs))
    total = sum(q["n"] for q in seqs)
    lat = np.array(lat)
    return {"sequences": nseq, "frames": total, "makespan_s": round(wall, 3), "grouped_calls": calls[0],
            "phase_recovery_groups": groups, "frames_per_s": round(total / wall, 1),
            "sequence_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1),
                                    "p99": round(float(np.percentile(lat, 99)), 1)},
            "form": f"smip_job_add_frames_many: one launch per stage and round for the next {batches_per_call} "
                    f"batch(es) of up to {window} unfinished sequences per call (admitted longest first), one host "
                    "thread; grouped phase recovery"}


def fft_passes_leg(ctx) -> dict:
    """The transform passes alone: b = 2 leaves the triple products negligible,
    so a batch's accumulate time is its forward FFT and |F|^2, and with
    shift-and-add on its SAA time is the FP64 FFT, correlation, inverse with
    argmax and the shifted sum.  512 frames per batch of the 256x256 uint16 and
    512x512 float32 stacks."""
    out = {}
    for name, N, passes in (("triple256", 256, 4), ("extended512", 512, 2)):
        fr = S.synth_frames(S.synth_params(name), 0, 512, threads=min(16, os.cpu_count() or 1))
        dev = S.Frames(ctx, fr)
        for saa in (False, True):
            job = S.Job(ctx, S.job_params(N, N, 2, 2, fr.dtype, S.PRECISION_MIXED, saa=saa, frames_per_batch=512))
            job.add_frames(dev)
            ctx.synchronize()
            st0 = job.stats()
            t0 = time.perf_counter()
            for _ in range(passes):
                job.add_frames(dev)
            ctx.synchronize()
            wall = time.perf_counter() - t0
            st = job.stats()
            nf = 512 * passes
            out[f"{name}{'_saa' if saa else ''}"] = {
                "frames": nf, "wall_us_per_frame": round(wall / nf * 1e6, 3),
                "accumulate_us_per_frame": round((st["accumulate_ms"] - st0["accumulate_ms"]) / nf * 1e3, 3),
                "saa_us_per_frame": round((st["saa_ms"] - st0["saa_ms"]) / nf * 1e3, 3)}
            job.close()
        dev.close()
    return out


def phase_bins(m, d, u, fl) -> dict:
    """Median |phase difference| of the mixed path and of the one-ulp FP64 run
    against the FP64 path, over recovered points grouped by their FP64
    consistency |mean phasor|: well-determined phases (consistency near 1)
    against those the recursion barely constrains."""
    c = d.consistency
    out = {}
    for lo, hi in ((0.5, 1.01), (0.1, 0.5), (0.01, 0.1), (0.0, 0.01)):
        sel = fl & (c >= lo) & (c < hi)
        if not sel.any():
            continue
        out[f"[{lo},{min(hi, 1.0)})"] = {"points": int(sel.sum()),
                                         "mixed_median": float(np.median(np.abs(m.phases[sel] - d.phases[sel]))),
                                         "fp64_one_ulp_median": float(np.median(np.abs(u.phases[sel] - d.phases[sel])))}
    return out


def precision_leg(ctx, n: int, batch: int = 256, chunk: int = 512) -> dict:
    """Error of the mixed path against the FP64 path on the 512x512 float32
    extended-object stack (b = 64, p = 256) at n frames: FP32 sums per batch of
    `batch` frames folded in FP64, frames streamed from host memory in chunks of
    `chunk` through smip_job_push_frames_host, as a camera would deliver them.
    Reported: bispectrum magnitude and phase, phases, flags, shifts and the
    image.  The FP64 GPU path is the reference here (it matches the FP64 oracle
    to 1e-11 of max|B| on the sizes the oracle finishes, and its 3-frame result
    at this size is checked against the oracle's in the GPU tests).  A third job
    runs the FP64 path on the same frames moved by one float32 ulp: how far
    rounding alone moves the result (the recursion's conditioning)."""
    W = S.WORKLOADS["extended512"]
    N, b, pr = W["N"], W["b"], W["p"]
    prm = S.synth_params("extended512")
    jobs = {"mixed": S.Job(ctx, S.job_params(N, N, b, pr, np.float32, S.PRECISION_MIXED, frames_per_batch=batch)),
            "fp64": S.Job(ctx, S.job_params(N, N, b, pr, np.float32, S.PRECISION_FP64)),
            "fp64_ulp": S.Job(ctx, S.job_params(N, N, b, pr, np.float32, S.PRECISION_FP64))}
    threads = min(16, os.cpu_count() or 1)
    # the next chunk is generated while the GPU works on this one
    gen_s = [0.0]

    def gen(first):
        t = time.perf_counter()
        f = S.synth_frames(prm, first, min(chunk, n - first), threads=threads)
        gen_s[0] += time.perf_counter() - t
        return f

    import concurrent.futures as cf
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=1) as ex:
        nxt = ex.submit(gen, 0)
        for first in range(0, n, chunk):
            fr = nxt.result()
            if first + chunk < n:
                nxt = ex.submit(gen, first + chunk)
            jobs["mixed"].push_frames_host(fr)
            jobs["fp64"].push_frames_host(fr)
            jobs["fp64_ulp"].push_frames_host(np.nextafter(fr, np.float32(np.inf)))
            if (first // chunk) % 4 == 3:
                log(f"precision: {first + len(fr)}/{n} frames pushed")
    S.finalize_many(list(jobs.values()))
    wall = time.perf_counter() - t0
    log(f"precision: {n} frames through three jobs in {wall:.1f} s (generation {gen_s[0]:.1f} s)")
    st = {k: j.stats() for k, j in jobs.items()}
    res = {k: j.result(with_bispectrum=(k != "fp64_ulp")) for k, j in jobs.items()}
    for j in jobs.values():
        j.close()
    m, d, u = res["mixed"], res["fp64"], res["fp64_ulp"]
    bm, bd = m.bispectrum, d.bispectrum
    amax = float(np.abs(bd).max())
    mag_num = mag_den = 0.0
    ph_w = ph_n = 0.0
    ph_max = 0.0
    err_max = 0.0
    step = 1 << 24
    for k in range(0, bd.size, step):  # chunked: 4.5 GB per bispectrum
        a, c = bm[k:k + step], bd[k:k + step]
        aa, ab = np.abs(a), np.abs(c)
        mag_num += float(np.sum((aa - ab) ** 2))
        mag_den += float(np.sum(ab ** 2))
        err_max = max(err_max, float(np.max(np.abs(a - c))))
        sig = ab > 1e-6 * amax
        if sig.any():
            dph = np.abs(np.angle(a[sig] * np.conj(c[sig])))
            ph_w += float(np.sum(dph * ab[sig]))
            ph_n += float(np.sum(ab[sig]))
            ph_max = max(ph_max, float(dph.max()))
    fl = m.flags.astype(bool) & d.flags.astype(bool)
    cons = d.consistency[d.flags.astype(bool)]
    rel_img = lambda a, c: float(np.linalg.norm(np.abs(a) - np.abs(c)) / np.linalg.norm(np.abs(c)))
    return {"workload": f"extended512: {n} of {W['n']} frames 512x512 float32, b={b}, p={pr}",
            "frames_per_batch": batch, "ingest": f"smip_job_push_frames_host, chunks of {chunk} frames from "
                                                 "pageable memory, generation overlapped with the GPU work",
            "reference": "FP64 GPU path",
            "frames": int(st["mixed"]["frames"]), "fp64_folds": -(-n // batch),
            "bispectrum_max_abs_err_over_max": err_max / amax,
            "bispectrum_magnitude_rel_l2": mag_num ** 0.5 / mag_den ** 0.5,
            "bispectrum_phase_rad": {"weighted_mean": ph_w / ph_n, "max_where_|B|>1e-6max": ph_max},
            "phase_max_abs_diff": float(np.abs(m.phases[fl] - d.phases[fl]).max()),
            "flags_equal": bool(np.array_equal(m.flags, d.flags)),
            "shifts_equal": bool(np.array_equal(m.shifts, d.shifts)),
            "image_rel_l2": rel_img(m.image, d.image),
            "fp64_vs_fp64_one_ulp_input": {"image_rel_l2": rel_img(u.image, d.image),
                                           "phase_max_abs_diff": float(np.abs(u.phases[fl] - d.phases[fl]).max()),
                                           "flags_equal": bool(np.array_equal(u.flags, d.flags))},
            "phase_diff_by_consistency": phase_bins(m, d, u, fl),
            "recovered_points": int(cons.size), "consistency_below_1e-2": float(np.mean(cons < 1e-2)),
            "device_ms": {k: {q: round(v[q], 1) for q in ("accumulate_ms", "saa_ms", "reco_ms")}
                          for k, v in st.items()},
            "wall_s": round(wall, 1), "generation_s": round(gen_s[0], 1)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="triple256")
    ap.add_argument("--frames", type=int, default=0, help="override the workload's frame count")
    ap.add_argument("--batch", type=int, default=0, help="frames per FP64 fold (0 = engine default)")
    ap.add_argument("--vy-tile", type=int, default=0)
    ap.add_argument("--jobs", type=int, default=8, help="concurrent jobs of the latency measurement")
    ap.add_argument("--plain-frames", type=int, default=400)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-latency", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="arrival window of the serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson job arrivals per second")
    ap.add_argument("--sweep-rates", default="20,30,40,50,60,80",
                    help="Poisson rates of the serving capacity sweep (--only sweep)")
    ap.add_argument("--sweep-seconds", type=float, default=8.0, help="arrival window per rate of the sweep")
    ap.add_argument("--batch-seqs", type=int, default=64, help="sequences of the multi-sequence leg (0: skip)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="hardware queues per process for HIP (GPU_MAX_HW_QUEUES, at most 32; 0: HIP's default)")
    ap.add_argument("--c3-frames", type=int, default=512, help="frames of the 512x512 throughput sample")
    ap.add_argument("--no-configs", action="store_true", help="skip the 128x128 and 512x512 legs")
    ap.add_argument("--only", default="",
                    help="comma list of legs to run: latency,configs,batch,precision,serve,cpu,fft,sweep")
    ap.add_argument("--precision-frames", type=int, default=10000,
                    help="frames of the 512^2 precision comparison (0: skip)")
    args = ap.parse_args()
    if args.hw_queues:
        if not 1 <= args.hw_queues <= 32:
            print(json.dumps({"error": "--hw-queues must be in [1, 32]"}))
            return 2
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)  # read when HIP initialises (first context)
    args.other_configs = not args.no_configs
    if args.only:
        legs = set(args.only.split(","))
        args.no_latency = "latency" not in legs
        args.other_configs = "configs" in legs
        args.no_cpu = "cpu" not in legs
        args.batch_seqs = args.batch_seqs if "batch" in legs else 0
        args.precision_frames = args.precision_frames if "precision" in legs else 0
        args.serve_seconds = args.serve_seconds if "serve" in legs else 0
        args.fft_passes = "fft" in legs
        args.sweep = "sweep" in legs
    else:
        args.fft_passes = False
        args.sweep = False

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        return 2

    W = dict(S.WORKLOADS[args.workload])
    n = args.frames or W["n"]
    N, b, pr = W["N"], W["b"], W["p"]
    log(f"generating {n} frames of {N}x{N} {np.dtype(W['dtype']).name} ({args.workload})")
    t0 = time.perf_counter()
    frames = S.synth_frames(S.synth_params(args.workload), 0, n, threads=min(16, os.cpu_count() or 1))
    log(f"generated in {time.perf_counter() - t0:.1f} s")

    ctx = S.Context(0)
    dev = S.Frames(ctx, frames)

    # ---- batched mixed-precision accumulate: the metric ----------------------
    p = S.job_params(N, N, b, pr, W["dtype"], S.PRECISION_MIXED, saa=False, frames_per_batch=args.batch,
                     vy_tile=args.vy_tile)
    job = S.Job(ctx, p)
    for _ in range(args.warmup):
        job.add_frames(dev)
    ctx.synchronize()
    st0 = job.stats()
    log(f"timing {args.steps} steps of {n} frames (mixed)")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        job.add_frames(dev)
    ctx.synchronize()
    t1 = time.perf_counter()
    st1 = job.stats()
    job.close()
    elapsed = t1 - t0
    fps = args.steps * n / elapsed
    tile_ms = st1["tile_ms"] - st0["tile_ms"]
    launches = st1["tile_launches"] - st0["tile_launches"]
    acc_ms = st1["accumulate_ms"] - st0["accumulate_ms"]
    frames_per_launch = args.steps * n / launches
    triples = st1["triples_per_frame"]
    flops_per_launch = 14.0 * triples * frames_per_launch
    avg_launch_s = tile_ms / launches * 1e-3
    achieved = flops_per_launch / avg_launch_s / 1e12
    log(f"mixed: {fps:.0f} frames/s; triple-product kernel {achieved:.1f} TFLOP/s, "
        f"{tile_ms / (args.steps * n) * 1e3:.2f} us/frame of {acc_ms / (args.steps * n) * 1e3:.2f} us/frame")

    # ---- the plain FP64 GPU drop-in (one frame per accumulate launch) --------
    n64 = min(n, args.plain_frames)
    p64 = S.job_params(N, N, b, pr, W["dtype"], S.PRECISION_FP64, saa=False)
    j64 = S.Job(ctx, p64)
    j64.add_frames(dev, 0, min(16, n64))
    ctx.synchronize()
    t0 = time.perf_counter()
    j64.add_frames(dev, 0, n64)
    ctx.synchronize()
    fps64 = n64 / (time.perf_counter() - t0)
    j64.close()
    log(f"plain FP64: {fps64:.0f} frames/s over {n64} frames; speedup {fps / fps64:.1f}x")

    log("the same accumulate streamed from host memory")
    streamed = streamed_leg(ctx, frames, p)
    log(f"streamed from host: {streamed}")

    # one GPU's share of the stack split frame-wise over 8 GPUs: its partial
    # bispectrum (the full FP64 cell array) is what one reduce step would send
    share = max(1, n // 8)
    js = S.Job(ctx, p)
    js.add_frames(dev, 0, share)
    ctx.synchronize()
    t0 = time.perf_counter()
    js.add_frames(dev, share, share)
    ctx.synchronize()
    share_ms = (time.perf_counter() - t0) * 1e3
    js.close()
    sizes = [2 * (d // 2) + 1 for d in (N, N, b, b)]
    base = (sizes[0] - sizes[0] // 2) * sizes[1] * (sizes[2] - sizes[2] // 2) * sizes[3]
    gpu_share = {"frames": share, "accumulate_ms": round(share_ms, 2),
                 "partial_bispectrum_bytes": 16 * base,
                 "note": "frames of one GPU when the stack is split over 8; the partial FP64 bispectrum "
                         "is reduced to the GPU that recovers the phases"}
    log(f"one GPU's share of 8: {gpu_share}")

    out = {
        "metric": "bispectrum frames/s @256^2,b=48,p=128 vs plain FP64 GPU path; p50/p99 job latency",
        "value": round(fps, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n} frames {N}x{N} {np.dtype(W['dtype']).name}, "
                               f"triple star, b={b}, p={pr}, mixed precision (FP32 batch, FP64 fold)",
                   "frames": n, "ncols": N, "nrows": N, "bdepth": b, "reco_radius": pr,
                   "frames_per_batch": args.batch or 512, "triples_per_frame": triples,
                   "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "HIP default")},
        "plain_fp64_frames_per_s": round(fps64, 1),
        "speedup_vs_plain_fp64": round(fps / fps64, 2),
        "roofline": {"bound": "valu", "achieved": round(achieved, 2), "peak": VALU_FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / VALU_FP32_PEAK_TFLOPS, 4),
                     "traffic": traffic_from_profile(os.path.join(REPO, "profiles", "tile_triple256_pmc.txt")),
                     "kernel": "k_accum_tile", "flops_per_launch": flops_per_launch,
                     "avg_launch_ms": round(tile_ms / launches, 4)},
        "streamed_from_host": streamed,
        "gpu_share_of_8": gpu_share,
    }

    # ---- complete jobs: latency and accuracy ---------------------------------
    if not args.no_latency:
        log("running 5 complete jobs one after another (accumulate + SAA + phases + synthesis)")
        single = []
        for _ in range(5):
            t = time.perf_counter()
            lat1, st1 = run_latency_jobs(ctx, dev, W, 1, args.batch, args.vy_tile)
            single.append((time.perf_counter() - t) * 1e3)
        out["single_job_latency_ms"] = {"p50": round(float(np.percentile(single, 50)), 2),
                                        "p99": round(float(np.percentile(single, 99)), 2), "jobs": 5,
                                        "frames_per_job": n, "stages_ms": {k: round(st1[0][k], 2) for k in
                                        ("accumulate_ms", "saa_ms", "reco_ms", "synth_ms")}}
        log(f"single job latency {out['single_job_latency_ms']}")
        log(f"running {args.jobs} concurrent complete jobs (accumulate + SAA + phases + synthesis)")
        stagger = 40.0
        lat, stats = run_latency_jobs(ctx, dev, W, args.jobs, args.batch, args.vy_tile, grouped=True,
                                      stagger_ms=stagger)
        lat.sort()
        out["job_latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 2),
                                 "p99": round(float(np.percentile(lat, 99)), 2), "jobs": args.jobs,
                                 "each_ms": [round(x, 1) for x in lat],
                                 "frames_per_job": n,
                                 "arrival": f"one job every {stagger:.0f} ms, latency from its own arrival",
                                 "finalization": "grouped: the jobs whose frames are in recover their phases "
                                                 "together (smip_job_finalize_many), as the serving loop does"}
        s0 = stats[0]
        out["job_stage_ms"] = {k: round(s0[k], 2) for k in ("accumulate_ms", "saa_ms", "reco_ms", "synth_ms")}
        log(f"latency p50 {out['job_latency_ms']['p50']} ms p99 {out['job_latency_ms']['p99']} ms; "
            f"stages {out['job_stage_ms']}")
        lat_i, _ = run_latency_jobs(ctx, dev, W, args.jobs, args.batch, args.vy_tile, grouped=False,
                                    stagger_ms=stagger)
        out["job_latency_ms_individual_finalize"] = {"p50": round(float(np.percentile(lat_i, 50)), 2),
                                                     "p99": round(float(np.percentile(lat_i, 99)), 2)}
        log(f"each job finalising on its own: {out['job_latency_ms_individual_finalize']}")
        # accuracy of the mixed path against the FP64 GPU path on the whole stack
        rm = S.run_pipeline(ctx, frames, b, pr, precision=S.PRECISION_MIXED, with_bispectrum=True)
        rd = S.run_pipeline(ctx, frames, b, pr, precision=S.PRECISION_FP64, with_bispectrum=True)
        out["mixed_vs_fp64_gpu"] = {
            "frames": n, "fp64_folds": -(-n // (args.batch or 512)),
            "bispectrum_max_abs_err_over_max": float(np.max(np.abs(rm.bispectrum - rd.bispectrum))
                                                     / np.max(np.abs(rd.bispectrum))),
            "image_rel_l2": float(np.linalg.norm(np.abs(rm.image) - np.abs(rd.image)) / np.linalg.norm(np.abs(rd.image))),
            "phase_max_abs_diff": float(np.max(np.abs(rm.phases - rd.phases))),
            "flags_equal": bool(np.array_equal(rm.flags, rd.flags)),
            "shifts_equal": bool(np.array_equal(rm.shifts, rd.shifts)),
        }
        del rm, rd
        log(f"mixed vs FP64 GPU: {out['mixed_vs_fp64_gpu']}")

    if args.other_configs:
        log("the other single-sequence configurations")
        out["binary128"] = workload_leg(ctx, "binary128", 40, passes=50, latency_runs=5)
        log(f"binary128: {out['binary128']}")
        out["extended512"] = workload_leg(ctx, "extended512", args.c3_frames, passes=2, latency_runs=1)
        log(f"extended512: {out['extended512']}")

    if args.fft_passes:
        log("the transform passes alone (b = 2)")
        out["fft_passes"] = fft_passes_leg(ctx)
        log(f"fft passes: {out['fft_passes']}")

    if args.batch_seqs > 0:
        log(f"multi-sequence: {args.batch_seqs} sequences, 16 host threads")
        out["multi_sequence"] = batch_leg(ctx, args.batch_seqs, 16, plain_sample=32)
        log(f"multi-sequence: {out['multi_sequence']}")
        log(f"multi-sequence, grouped launches: {args.batch_seqs} sequences")
        for bpc in (1, 4):
            gg = batch_leg_grouped(ctx, args.batch_seqs, bpc)
            log(f"multi-sequence, grouped launches, {bpc} batches per call: {gg}")
        g = batch_leg_grouped(ctx, args.batch_seqs, 2)
        g["speedup_vs_plain_fp64_serial"] = round(out["multi_sequence"]["plain_fp64_serial_s_estimate"]
                                                  / g["makespan_s"], 2)
        out["multi_sequence_grouped"] = g
        log(f"multi-sequence, grouped launches: {g}")

    if args.precision_frames > 0:
        log(f"precision: mixed vs FP64 on {args.precision_frames} frames of 512x512, b=64, p=256")
        out["precision_512"] = precision_leg(ctx, args.precision_frames)
        log(f"precision: {out['precision_512']}")

    if args.serve_seconds > 0:
        log(f"serving: Poisson arrivals at {args.serve_rate} jobs/s for {args.serve_seconds} s")
        out["serving"] = serve_leg(ctx, args.serve_seconds, args.serve_rate, seed=0x5EC1E005)
        sv = out["serving"]
        log(f"serving: {sv['completed']}/{sv['jobs']} jobs, {sv['frames_per_s']} frames/s, latency {sv['latency_ms']}, "
            f"after last frame {sv['after_last_frame_ms']}")

    if args.sweep:
        rates = [float(x) for x in args.sweep_rates.split(",") if x]
        ms = out.get("multi_sequence")
        plain_fps = ms["frames"] / ms["plain_fp64_serial_s_estimate"] if ms else None
        log(f"serving capacity sweep: {rates} jobs/s, {args.sweep_seconds} s each")
        out["serving_capacity"] = serve_sweep(ctx, rates, args.sweep_seconds, plain_fps)
        log(f"serving capacity: knee {out['serving_capacity']['knee_rate_jobs_per_s']} jobs/s, "
            f"{out['serving_capacity']['capacity_frames_per_s']} frames/s")

    if not args.no_cpu:
        log("timing the CPU baseline")
        out["cpu_baseline"] = cpu_baseline(frames, b, pr, args.cpu_seconds)
        log(f"cpu baseline {out['cpu_baseline']['value']:.3f} frames/s")

    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
