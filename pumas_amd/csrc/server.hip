/*
 * Request server: continuous batching of transport requests on one GPU.
 *
 * Requests of uneven sizes arrive at any time (pumas_gpu_server_submit).
 * A request is admitted when a slot is free (one HIP stream per slot) and
 * its device memory fits the HBM budget beside the requests in flight; it
 * then runs on its slot's stream while others run on theirs, so that
 * arriving work overlaps with work already on the device instead of
 * waiting for a batch boundary. Admission is first come, first served.
 *
 * A request's results depend only on its own inputs (states, seed, history
 * ids, context, geometry): histories draw from Philox streams keyed by
 * their ids, and the kernels have no cross-request state. The same request
 * alone or in a mix gives identical results.
 *
 * No counterpart in the reference library, which transports one history per
 * call on the calling thread (pumas_context_transport, src/pumas.c:3114).
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pumas_gpu.h"
#include "status.h"

namespace {

using Clock = std::chrono::steady_clock;

/* Device bytes of a request in flight: the scheduler's history record and
 * queue slots per history, plus the state and output copies of a
 * host-memory request. */
size_t request_bytes(size_t n, int device)
{
        const size_t per = pg_bytes_per_history() + (device ? 0 : 104 + 24);
        return n * per + (size_t(1) << 20);
}

struct Job {
        uint64_t ticket;
        struct pumas_gpu_request req;
        size_t bytes;
        Clock::time_point t_submit, t_admit, t_done;
        struct pumas_gpu_request_status status;
        bool done;
};

} /* namespace */

struct pumas_gpu_server {
        int device;
        size_t budget, in_use;
        int running; /* requests admitted and not done */
        int slots;
        std::mutex mu;
        std::condition_variable cv_work, cv_done;
        std::deque<Job *> pending;
        std::unordered_map<uint64_t, Job *> jobs;
        uint64_t next_ticket;
        bool stop;
        std::vector<std::thread> workers;
};

namespace {

/* Transport of a host-memory request on `stream`: copies in, transport,
 * copies out. */
enum pumas_return run_host_request(const struct pumas_gpu_request & r,
    const struct pumas_gpu_options * opts, struct pumas_gpu_stats * stats,
    hipStream_t stream)
{
        const size_t n = r.n;
        const size_t ns = n > 0 ? n : 1;
        const size_t b_states = ns * sizeof(struct pumas_state);
        const size_t b_out = ns * (sizeof(int32_t) * 4 + sizeof(uint32_t) * 2);
        void * mem = nullptr;
        HIP_TRY(hipMallocAsync(&mem, b_states + b_out + 256, stream));
        /* From here on, a failed copy frees the buffer before returning. */
#define HIP_TRY_FREE(expr)                                                     \
        do {                                                                   \
                const hipError_t e_ = (expr);                                  \
                if (e_ != hipSuccess) {                                        \
                        (void)hipFreeAsync(mem, stream);                       \
                        return pg_fail(PUMAS_RETURN_INTERNAL_ERROR,            \
                            "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),    \
                            __FILE__, __LINE__);                               \
                }                                                              \
        } while (0)
        char * p = static_cast<char *>(mem);
        struct pumas_state * d_states = reinterpret_cast<struct pumas_state *>(p);
        char * q = p + ((b_states + 255) & ~size_t(255));
        struct pumas_gpu_outputs d_out;
        memset(&d_out, 0x0, sizeof(d_out));
        const struct pumas_gpu_outputs & o = r.outputs;
        if (o.events) d_out.events = reinterpret_cast<enum pumas_event *>(q);
        q += ns * sizeof(int32_t);
        if (o.media) d_out.media = reinterpret_cast<int32_t *>(q);
        q += 2 * ns * sizeof(int32_t);
        if (o.steps) d_out.steps = reinterpret_cast<uint32_t *>(q);
        q += ns * sizeof(uint32_t);
        if (o.draws) d_out.draws = reinterpret_cast<uint32_t *>(q);
        q += ns * sizeof(uint32_t);
        if (o.rc) d_out.rc = reinterpret_cast<int32_t *>(q);

        /* Outputs start at zero: a transport that fails before its first
         * kernel (a context error) leaves them so, and they are copied back
         * all the same. */
        HIP_TRY_FREE(hipMemsetAsync(p + ((b_states + 255) & ~size_t(255)), 0,
            b_out, stream));
        HIP_TRY_FREE(hipMemcpyAsync(d_states, r.states, n * sizeof(*r.states),
            hipMemcpyHostToDevice, stream));
        const enum pumas_return rc = pumas_gpu_transport(r.physics,
            &r.context, r.geometry, d_states, n, r.seed, r.first_history,
            &d_out, opts, stats, stream);
        {
                /* Results are copied back also when a history failed. */
                HIP_TRY_FREE(hipMemcpyAsync(r.states, d_states,
                    n * sizeof(*r.states), hipMemcpyDeviceToHost, stream));
                if (o.events)
                        HIP_TRY_FREE(hipMemcpyAsync(o.events, d_out.events,
                            n * sizeof(int32_t), hipMemcpyDeviceToHost,
                            stream));
                if (o.media)
                        HIP_TRY_FREE(hipMemcpyAsync(o.media, d_out.media,
                            2 * n * sizeof(int32_t), hipMemcpyDeviceToHost,
                            stream));
                if (o.steps)
                        HIP_TRY_FREE(hipMemcpyAsync(o.steps, d_out.steps,
                            n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            stream));
                if (o.draws)
                        HIP_TRY_FREE(hipMemcpyAsync(o.draws, d_out.draws,
                            n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                            stream));
                if (o.rc)
                        HIP_TRY_FREE(hipMemcpyAsync(o.rc, d_out.rc,
                            n * sizeof(int32_t), hipMemcpyDeviceToHost,
                            stream));
        }
#undef HIP_TRY_FREE
        HIP_TRY(hipFreeAsync(mem, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        return rc;
}

void worker_loop(pumas_gpu_server * s)
{
        (void)hipSetDevice(s->device);
        hipStream_t stream = nullptr;
        (void)hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        for (;;) {
                Job * job = nullptr;
                int job_lanes = 64;
                {
                        std::unique_lock<std::mutex> lock(s->mu);
                        s->cv_work.wait(lock, [s] {
                                if (s->pending.empty()) return s->stop;
                                const Job * head = s->pending.front();
                                return s->in_use + head->bytes <= s->budget;
                        });
                        if (s->pending.empty()) break;
                        job = s->pending.front();
                        s->pending.pop_front();
                        s->in_use += job->bytes;
                        s->running += 1;
                        /* Alone on the device, a small request is spread
                         * over every CU (latency); with others in flight or
                         * waiting it takes full wavefronts, which leaves
                         * the rest of the device to them (throughput). */
                        job_lanes = ((s->running > 1) || !s->pending.empty()) ?
                            64 :
                            0;
                        job->t_admit = Clock::now();
                }
                /* Admitting one may let the next fit as well. */
                s->cv_work.notify_all();

                struct pumas_gpu_options opts;
                memset(&opts, 0x0, sizeof(opts));
                opts.precision = job->req.precision;
                opts.lanes_per_wave = job_lanes;
                struct pumas_gpu_stats stats;
                memset(&stats, 0x0, sizeof(stats));
                enum pumas_return rc;
                if (job->req.device)
                        rc = pumas_gpu_transport(job->req.physics,
                            &job->req.context, job->req.geometry,
                            job->req.states, job->req.n, job->req.seed,
                            job->req.first_history, &job->req.outputs, &opts,
                            &stats, stream);
                else
                        rc = run_host_request(job->req, &opts, &stats, stream);

                {
                        std::lock_guard<std::mutex> lock(s->mu);
                        job->t_done = Clock::now();
                        job->status.rc = rc;
                        job->status.steps = stats.steps;
                        job->status.kernel_ms = stats.kernel_ms;
                        const double q = std::chrono::duration<double,
                            std::milli>(job->t_admit - job->t_submit).count();
                        const double l = std::chrono::duration<double,
                            std::milli>(job->t_done - job->t_submit).count();
                        job->status.queue_ms = q;
                        job->status.latency_ms = l;
                        if (rc != PUMAS_RETURN_SUCCESS)
                                strncpy(job->status.message,
                                    pumas_gpu_error_message(),
                                    sizeof(job->status.message) - 1);
                        job->done = true;
                        s->in_use -= job->bytes;
                        s->running -= 1;
                }
                s->cv_done.notify_all();
                s->cv_work.notify_all();
        }
        if (stream) (void)hipStreamDestroy(stream);
}

} /* namespace */

extern "C" enum pumas_return pumas_gpu_server_create(
    const struct pumas_gpu_server_options * options,
    pumas_gpu_server ** server)
{
        pg_clear_error();
        if (server == NULL)
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL argument");
        *server = nullptr;
        int device = 0;
        HIP_TRY(hipGetDevice(&device));
        size_t free_b = 0, total_b = 0;
        HIP_TRY(hipMemGetInfo(&free_b, &total_b));
        auto * s = new pumas_gpu_server();
        s->device = device;
        s->budget = (options && options->hbm_budget > 0) ?
            options->hbm_budget :
            size_t(200) << 30;
        if (s->budget > free_b) s->budget = free_b;
        s->slots = (options && options->slots > 0) ? options->slots : 4;
        if (s->slots > PUMAS_GPU_SERVER_MAX_SLOTS) {
                delete s;
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "at most %d server slots", PUMAS_GPU_SERVER_MAX_SLOTS);
        }
        s->in_use = 0;
        s->running = 0;
        s->next_ticket = 1;
        s->stop = false;
        for (int i = 0; i < s->slots; i++)
                s->workers.emplace_back(worker_loop, s);
        *server = s;
        return PUMAS_RETURN_SUCCESS;
}

extern "C" void pumas_gpu_server_destroy(pumas_gpu_server ** server)
{
        if ((server == NULL) || (*server == NULL)) return;
        pumas_gpu_server * s = *server;
        {
                std::lock_guard<std::mutex> lock(s->mu);
                s->stop = true;
        }
        s->cv_work.notify_all();
        for (auto & t : s->workers) t.join();
        for (auto & kv : s->jobs) delete kv.second;
        delete s;
        *server = nullptr;
}

extern "C" enum pumas_return pumas_gpu_server_submit(pumas_gpu_server * s,
    const struct pumas_gpu_request * request, uint64_t * ticket)
{
        pg_clear_error();
        if ((s == NULL) || (request == NULL) || (ticket == NULL))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL argument");
        if ((request->physics == NULL) || (request->geometry == NULL) ||
            ((request->states == NULL) && (request->n > 0)))
                return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                    "incomplete request (NULL physics, geometry or states)");
        auto * job = new Job();
        job->req = *request;
        job->bytes = request_bytes(request->n, request->device);
        if (job->bytes > s->budget) {
                const size_t b = job->bytes;
                delete job;
                return pg_fail(PUMAS_RETURN_MEMORY_ERROR,
                    "request needs %zu bytes of device memory, above the "
                    "server budget of %zu", b, s->budget);
        }
        memset(&job->status, 0x0, sizeof(job->status));
        job->done = false;
        job->t_submit = Clock::now();
        {
                std::lock_guard<std::mutex> lock(s->mu);
                if (s->stop) {
                        delete job;
                        return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                            "the server is stopping");
                }
                job->ticket = s->next_ticket++;
                s->jobs[job->ticket] = job;
                s->pending.push_back(job);
                *ticket = job->ticket;
        }
        s->cv_work.notify_all();
        return PUMAS_RETURN_SUCCESS;
}

extern "C" size_t pumas_gpu_server_request_bytes(size_t n, int device)
{
        return request_bytes(n, device);
}

extern "C" int pumas_gpu_server_poll(pumas_gpu_server * s, uint64_t ticket)
{
        if (s == NULL) return -1;
        std::lock_guard<std::mutex> lock(s->mu);
        auto it = s->jobs.find(ticket);
        if (it == s->jobs.end()) return -1;
        return it->second->done ? 1 : 0;
}

extern "C" enum pumas_return pumas_gpu_server_wait(pumas_gpu_server * s,
    uint64_t ticket, struct pumas_gpu_request_status * status)
{
        pg_clear_error();
        if (s == NULL) return pg_fail(PUMAS_RETURN_VALUE_ERROR, "NULL server");
        Job * job = nullptr;
        {
                std::unique_lock<std::mutex> lock(s->mu);
                auto it = s->jobs.find(ticket);
                if (it == s->jobs.end())
                        return pg_fail(PUMAS_RETURN_VALUE_ERROR,
                            "unknown ticket %llu",
                            (unsigned long long)ticket);
                job = it->second;
                s->cv_done.wait(lock, [job] { return job->done; });
                /* By key: submissions made while waiting may have rehashed
                 * the map and invalidated `it`. */
                s->jobs.erase(ticket);
        }
        if (status) *status = job->status;
        const enum pumas_return rc = job->status.rc;
        if (rc != PUMAS_RETURN_SUCCESS)
                pg_fail(rc, "%s", job->status.message);
        delete job;
        return rc;
}
