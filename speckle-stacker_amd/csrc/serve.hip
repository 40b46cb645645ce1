// serve.hip -- continuous-batching request loop for reconstruction jobs.
//
// Jobs are submitted with their parameters, stream their frames in chunks, and
// are closed when the last chunk is in.  Admission is first-fit over the
// waiting jobs in arrival order: a job is admitted when a slot is free and its
// device footprint (smip_job_footprint: bispectrum, spectra batch, staging,
// phase state) fits the HBM budget beside the jobs already admitted.  Admitted
// jobs keep all their state on the device while frames stream in; worker
// threads move each job's pending chunks through smip_job_push_frames_host on
// the job's own stream, so the kernels of different jobs overlap on the GPU.
// A job whose frames are all in is finalised (phase recovery and synthesis) by
// the next free worker.  Batch boundaries depend only on a job's own frame
// sequence, so a job's result does not depend on what else is running.
// Device memory returns to the budget when the client releases the job.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "smip_internal.h"

using namespace smip_amd;
using Clock = std::chrono::steady_clock;

namespace {

struct ServedJob {
    uint64_t id = 0;
    smip_job_params p {};
    uint64_t footprint = 0;
    size_t frame_bytes = 0;
    int state = SMIP_JOB_WAITING;
    smip_job* job = nullptr;
    bool busy = false;     // a worker is driving it
    bool closed = false;   // no more frames will come
    bool released = false; // client gave it up while a worker held it
    std::deque<std::vector<uint8_t>> chunks;
    uint64_t frames_in = 0, frames_done = 0;
    Clock::time_point t_submit, t_admit, t_close, t_done;
    std::string error;
};

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

} // namespace

struct smip_server {
    CtxRef ctx_ref;
    smip_ctx* ctx = nullptr;
    uint64_t budget = 0;
    int max_active = 16;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    std::map<uint64_t, std::unique_ptr<ServedJob>> jobs;
    std::deque<uint64_t> waiting; // arrival order
    std::vector<uint64_t> active; // admitted, not released
    size_t rr = 0;                // round-robin position in `active`
    uint64_t next_id = 1, in_use = 0;
    uint64_t n_submitted = 0, n_admitted = 0, n_completed = 0, n_failed = 0, frames = 0;
    uint64_t n_groups = 0, n_grouped = 0; // finalisation groups and the jobs they held
    std::vector<std::thread> workers;
    // streams with hardware queues of their own for the phase recoveries of the
    // groups the workers finalise (worker w uses qstreams[w % size]): on a queue
    // shared with the jobs' streams, a recovery's chain of launches would hold up
    // their batches.  At most kQueueStreams of them: every such stream takes one
    // of the device's hardware queue slots for the server's lifetime, and beyond
    // the slots the command processor time-slices the queues.
    static constexpr int kQueueStreams = 8;
    std::vector<hipStream_t> qstreams;

    // first-fit admission in arrival order; caller holds m
    void admit_locked()
    {
        for (auto it = waiting.begin(); it != waiting.end() && (int)active.size() < max_active;) {
            ServedJob* j = jobs.at(*it).get();
            if (in_use + j->footprint <= budget) {
                in_use += j->footprint;
                j->state = SMIP_JOB_ACTIVE;
                j->t_admit = Clock::now();
                active.push_back(j->id);
                ++n_admitted;
                it = waiting.erase(it);
            } else {
                ++it;
            }
        }
    }

    // a job a worker can advance now: created-or-not, not busy, with a chunk or ready to finalise
    ServedJob* pick_locked()
    {
        const size_t n = active.size();
        for (size_t k = 0; k < n; ++k) {
            ServedJob* j = jobs.at(active[(rr + k) % n]).get();
            if (j->busy || j->state != SMIP_JOB_ACTIVE)
                continue;
            if (!j->job || !j->chunks.empty() || j->closed) {
                rr = (rr + k + 1) % n;
                return j;
            }
        }
        return nullptr;
    }

    void finish_locked(ServedJob* j, int state, const std::string& err)
    {
        j->state = state;
        j->error = err;
        j->t_done = Clock::now();
        if (state == SMIP_JOB_DONE)
            ++n_completed;
        else
            ++n_failed;
        cv_done.notify_all();
    }

    void destroy_locked(ServedJob* j)
    {
        if (j->job) {
            smip_job_destroy(j->job);
            j->job = nullptr;
        }
        if (j->state == SMIP_JOB_ACTIVE || j->state == SMIP_JOB_DONE || j->state == SMIP_JOB_FAILED) {
            auto a = std::find(active.begin(), active.end(), j->id);
            if (a != active.end()) {
                active.erase(a);
                in_use -= j->footprint;
            }
        }
        auto w = std::find(waiting.begin(), waiting.end(), j->id);
        if (w != waiting.end())
            waiting.erase(w);
        jobs.erase(j->id);
        admit_locked();
        cv_work.notify_all();
    }

    void worker_loop(int w)
    {
        std::unique_lock<std::mutex> lk(m);
        for (;;) {
            ServedJob* j = nullptr;
            cv_work.wait(lk, [&] { return stop || (j = pick_locked()) != nullptr; });
            if (stop)
                return;
            j->busy = true;
            std::vector<uint8_t> chunk;
            const bool create = j->job == nullptr;
            const bool finalize = !create && j->chunks.empty() && j->closed;
            // continuous batching of phase recovery: every job ready to finalise now
            // joins this one, and their recursions run as one group
            std::vector<ServedJob*> group;
            if (finalize) {
                group.push_back(j);
                for (uint64_t id : active) {
                    ServedJob* o = jobs.at(id).get();
                    if (o != j && !o->busy && o->state == SMIP_JOB_ACTIVE && o->job && o->closed && o->chunks.empty()) {
                        o->busy = true;
                        group.push_back(o);
                    }
                }
            }
            if (!create && !finalize) {
                chunk = std::move(j->chunks.front());
                j->chunks.pop_front();
            }
            lk.unlock();
            int rc = SMIP_OK;
            std::string err;
            uint64_t nf = 0;
            std::vector<int32_t> gst(group.size(), SMIP_OK);
            if (create) {
                smip_job* job = nullptr;
                rc = smip_job_create(ctx, &j->p, &job);
                if (rc == SMIP_OK)
                    j->job = job;
            } else if (finalize) {
                std::vector<smip_job*> handles;
                for (ServedJob* g : group)
                    handles.push_back(g->job);
                rc = smip_amd::job_finalize_many_on(handles.data(), handles.size(), gst.data(),
                    qstreams[(size_t)w % qstreams.size()]);
                if (rc != SMIP_OK && std::all_of(gst.begin(), gst.end(), [](int32_t c) { return c == SMIP_OK; }))
                    std::fill(gst.begin(), gst.end(), rc); // the group as a whole failed
            } else {
                nf = chunk.size() / j->frame_bytes;
                rc = smip_job_push_frames_host(j->job, chunk.data(), nf);
            }
            if (rc != SMIP_OK)
                err = smip_last_error();
            chunk.clear();
            lk.lock();
            if (finalize) {
                n_groups += 1;
                n_grouped += group.size();
                for (size_t g = 0; g < group.size(); ++g) {
                    ServedJob* o = group[g];
                    o->busy = false;
                    if (gst[g] == SMIP_OK)
                        finish_locked(o, SMIP_JOB_DONE, "");
                    else
                        finish_locked(o, SMIP_JOB_FAILED, err);
                    if (o->released)
                        destroy_locked(o);
                }
            } else {
                j->busy = false;
                if (rc != SMIP_OK) {
                    finish_locked(j, SMIP_JOB_FAILED, err);
                } else if (!create) {
                    j->frames_done += nf;
                    frames += nf;
                }
                if (j->released)
                    destroy_locked(j);
            }
            cv_work.notify_all();
        }
    }
};

namespace {

ServedJob* find_locked(smip_server* s, uint64_t id)
{
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || it->second->released)
        fail(SMIP_ERR_INVALID_ARGUMENT, "unknown job id " + std::to_string(id));
    return it->second.get();
}

} // namespace

extern "C" SMIP_API int smip_server_create(smip_ctx* ctx, const smip_server_params* p, smip_server** out)
{
    return guarded([&] {
        if (!ctx || !out)
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        smip_server_params q {};
        if (p)
            q = *p;
        auto s = std::make_unique<smip_server>();
        s->ctx_ref.hold(ctx);
        s->ctx = ctx;
        uint64_t free_b = 0, total_b = 0;
        if (smip_ctx_mem_info(ctx, &free_b, &total_b) != SMIP_OK)
            fail(SMIP_ERR_HIP, smip_last_error());
        s->budget = q.hbm_budget ? q.hbm_budget : free_b / 10 * 9;
        s->max_active = q.max_active > 0 ? q.max_active : 16;
        const int nw = q.workers > 0 ? q.workers : 8;
        if (nw > 64 || s->max_active > 4096)
            fail(SMIP_ERR_INVALID_ARGUMENT, "server parameters out of range");
        smip_amd::ensure_device(ctx);
        for (int w = 0; w < std::min(nw, smip_server::kQueueStreams); ++w)
            s->qstreams.push_back(smip_amd::make_queue_stream());
        smip_server* raw = s.get();
        for (int w = 0; w < nw; ++w)
            s->workers.emplace_back([raw, w] { raw->worker_loop(w); });
        *out = s.release();
    });
}

extern "C" SMIP_API int smip_server_destroy(smip_server* s)
{
    return guarded([&] {
        if (!s)
            return;
        {
            std::lock_guard<std::mutex> lk(s->m);
            s->stop = true;
            s->cv_work.notify_all();
        }
        for (auto& t : s->workers)
            t.join();
        for (auto& kv : s->jobs)
            if (kv.second->job)
                smip_job_destroy(kv.second->job);
        smip_amd::ensure_device(s->ctx);
        for (hipStream_t q : s->qstreams)
            smip_amd::destroy_stream(q);
        delete s;
    });
}

extern "C" SMIP_API int smip_server_submit(smip_server* s, const smip_job_params* p, uint64_t* id)
{
    return guarded([&] {
        if (!s || !p || !id)
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        // defaults resolved now, so the footprint admitted is the one created
        smip_job_params q {};
        if (smip_job_resolve_params(p, &q) != SMIP_OK)
            fail(SMIP_ERR_INVALID_ARGUMENT, smip_last_error());
        uint64_t fp = 0;
        if (smip_job_footprint(&q, &fp) != SMIP_OK)
            fail(SMIP_ERR_INVALID_ARGUMENT, smip_last_error());
        auto j = std::make_unique<ServedJob>();
        j->p = q;
        j->footprint = fp;
        j->frame_bytes = p->ncols * p->nrows * dtype_size(p->frame_dtype);
        j->t_submit = Clock::now();
        std::lock_guard<std::mutex> lk(s->m);
        if (fp > s->budget)
            fail(SMIP_ERR_NO_MEMORY, "job footprint " + std::to_string(fp) + " B exceeds the server's HBM budget");
        j->id = s->next_id++;
        *id = j->id;
        s->waiting.push_back(j->id);
        s->jobs[j->id] = std::move(j);
        ++s->n_submitted;
        s->admit_locked();
        s->cv_work.notify_all();
    });
}

extern "C" SMIP_API int smip_server_push(smip_server* s, uint64_t id, const void* frames, uint64_t count)
{
    return guarded([&] {
        if (!s || (!frames && count))
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        size_t fb = 0;
        {
            std::lock_guard<std::mutex> lk(s->m);
            ServedJob* j = find_locked(s, id);
            if (j->closed)
                fail(SMIP_ERR_LOGIC, "job already closed");
            fb = j->frame_bytes;
        }
        if (count == 0)
            return;
        // copy outside the lock; the caller's buffer is free on return
        std::vector<uint8_t> chunk(static_cast<const uint8_t*>(frames), static_cast<const uint8_t*>(frames) + count * fb);
        std::lock_guard<std::mutex> lk(s->m);
        ServedJob* j = find_locked(s, id);
        if (j->closed)
            fail(SMIP_ERR_LOGIC, "job already closed");
        j->chunks.push_back(std::move(chunk));
        j->frames_in += count;
        s->cv_work.notify_all();
    });
}

extern "C" SMIP_API int smip_server_close(smip_server* s, uint64_t id)
{
    return guarded([&] {
        if (!s)
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        std::lock_guard<std::mutex> lk(s->m);
        ServedJob* j = find_locked(s, id);
        if (j->closed)
            fail(SMIP_ERR_LOGIC, "job already closed");
        if (j->frames_in == 0)
            fail(SMIP_ERR_RUNTIME, "Array_base<T>::operator/=(T) : division by zero");
        j->closed = true;
        j->t_close = Clock::now();
        s->cv_work.notify_all();
    });
}

extern "C" SMIP_API int smip_server_wait(smip_server* s, uint64_t id, double timeout_s, int32_t* state)
{
    return guarded([&] {
        if (!s)
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        std::unique_lock<std::mutex> lk(s->m);
        find_locked(s, id);
        // the job is looked up on every wake: another thread may release it meanwhile
        auto done = [&] {
            auto it = s->jobs.find(id);
            return it == s->jobs.end() || it->second->released || it->second->state == SMIP_JOB_DONE
                || it->second->state == SMIP_JOB_FAILED;
        };
        if (timeout_s < 0)
            s->cv_done.wait(lk, done);
        else
            s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
        ServedJob* j = find_locked(s, id);
        if (state)
            *state = j->state;
        if (j->state == SMIP_JOB_FAILED)
            fail(SMIP_ERR_RUNTIME, "job failed: " + j->error);
    });
}

extern "C" SMIP_API int smip_server_job(smip_server* s, uint64_t id, smip_job** out)
{
    return guarded([&] {
        if (!s || !out)
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        std::lock_guard<std::mutex> lk(s->m);
        ServedJob* j = find_locked(s, id);
        if (j->state != SMIP_JOB_DONE)
            fail(SMIP_ERR_LOGIC, "job has not completed");
        *out = j->job;
    });
}

extern "C" SMIP_API int smip_server_get_job_info(smip_server* s, uint64_t id, smip_server_job_info* out)
{
    return guarded([&] {
        if (!s || !out)
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        std::lock_guard<std::mutex> lk(s->m);
        ServedJob* j = find_locked(s, id);
        out->state = j->state;
        out->footprint = j->footprint;
        out->frames_in = j->frames_in;
        out->frames_done = j->frames_done;
        const auto now = Clock::now();
        const bool admitted = j->state != SMIP_JOB_WAITING;
        const bool done = j->state == SMIP_JOB_DONE || j->state == SMIP_JOB_FAILED;
        out->queue_ms = ms_between(j->t_submit, admitted ? j->t_admit : now);
        out->latency_ms = ms_between(j->t_submit, done ? j->t_done : now);
        out->tail_ms = j->closed ? ms_between(j->t_close, done ? j->t_done : now) : 0.0;
        std::strncpy(out->error, j->error.c_str(), sizeof(out->error) - 1);
        out->error[sizeof(out->error) - 1] = 0;
    });
}

extern "C" SMIP_API int smip_server_release(smip_server* s, uint64_t id)
{
    return guarded([&] {
        if (!s)
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        std::lock_guard<std::mutex> lk(s->m);
        ServedJob* j = find_locked(s, id);
        j->released = true;
        s->cv_done.notify_all();
        if (!j->busy) // otherwise the worker holding it destroys it
            s->destroy_locked(j);
    });
}

extern "C" SMIP_API int smip_server_get_stats(smip_server* s, smip_server_stats* out)
{
    return guarded([&] {
        if (!s || !out)
            fail(SMIP_ERR_INVALID_ARGUMENT, "null pointer");
        std::lock_guard<std::mutex> lk(s->m);
        out->submitted = s->n_submitted;
        out->admitted = s->n_admitted;
        out->completed = s->n_completed;
        out->failed = s->n_failed;
        out->waiting = s->waiting.size();
        out->active = s->active.size();
        out->bytes_in_use = s->in_use;
        out->budget = s->budget;
        out->frames = s->frames;
        out->finalize_groups = s->n_groups;
        out->finalized_in_groups = s->n_grouped;
    });
}
