// Serving: 'simulate sounding S for T seconds' jobs on one GPU
// (include/skt.h, skt_server_*).
//
// A server owns one ensemble (skt_batch). Jobs wait in arrival order until
// their HBM footprint fits the budget and a slot is free (first fit: a job
// that fits may pass one that does not yet), run as ensemble members while
// active, and are retired when they have run their T seconds. Each server
// step is: admission, one ensemble step of the active jobs (a HIP graph
// replay), the report diagnostics of the jobs at a report boundary (one
// kernel each on the ensemble's stream, values copied to pinned memory, one
// wait for all), retirement. Between steps the host does no device waits
// unless a report or a retirement is due, so the device stays fed. A queued
// job's initial state is generated on the host by a worker thread while it
// waits (the first few jobs of the queue), so admitting it only uploads it.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <deque>
#include <future>
#include <memory>
#include <string>
#include <vector>
#include "skt_internal.h"

using namespace skt;

namespace {

using Clock = std::chrono::steady_clock;

// A job's initial state on the host (skt_sounding_generate), generated by a
// worker thread. `done` is declared last, so destroying a Prep first waits
// for its thread and only then frees the arrays the thread writes.
struct Prep {
    std::vector<float> th, vel, qv, gT, gP, p, pS;
    std::future<int> done;
};
// Queued jobs whose initial states are generated ahead of admission (host
// memory: 7 floats per cell each).
constexpr int kPrefetch = 4;

std::shared_ptr<Prep> startPrep(const skt_job_request& r) {
    auto pp = std::make_shared<Prep>();
    Prep* q = pp.get();
    q->done = std::async(std::launch::async, [q, r]() -> int {
        try {
            const size_t N = (size_t)r.nx * r.ny * r.nz, G = (size_t)r.nx * r.nz;
            q->th.resize(N);
            q->vel.resize(3 * N);
            q->qv.resize(N);
            q->gT.resize(G);
            q->gP.resize(G);
            q->p.resize(N);
            q->pS.resize((size_t)r.ny);
        } catch (...) {
            return SKT_ERR_NOMEM;
        }
        return skt_sounding_generate(&r.sounding, r.nx, r.ny, r.nz, r.dx, q->th.data(), q->vel.data(), q->qv.data(),
                                     q->gT.data(), q->gP.data(), q->p.data(), q->pS.data());
    });
    return pp;
}

struct Job {
    std::shared_ptr<Prep> prep;  // the initial state being generated, if started
    int64_t id = 0;
    skt_job_request req{};
    skt_config cfg{};
    int64_t bytes = 0;
    skt_env* env = nullptr;
    int32_t member = -1;
    int steps = 0, total = 0, report_steps = 0;
    double submitted = 0.0;  // server clock at submission
};

// The diagnostics fields of a member's device scalars, as copied to the host.
struct DgPin {
    float w;
    int level;
    double rain;
};
static_assert(offsetof(Scalars, dg_level) == offsetof(Scalars, dg_w) + sizeof(float),
              "max updraft and cloud-top level are adjacent in the device scalars");

}  // namespace

struct skt_server {
    skt_server_options opt{};
    skt_batch* batch = nullptr;
    std::deque<Job> queue;
    std::vector<Job> active;
    std::deque<skt_job_event> events;
    int64_t next_id = 1, bytes = 0, budget = 0, steps = 0, completed = 0;
    double cell_steps = 0.0, busy = 0.0, admit_s = 0.0, replan_s = 0.0;
    bool changed = true;  // membership changed since the last ensemble step
    Clock::time_point t0;
    DgPin* pin = nullptr;  // pinned, one per slot
    std::string err;
};

namespace {

int serr(skt_server* s, int rc, const std::string& m) {
    s->err = m;
    return rc;
}
double now(const skt_server* s) { return std::chrono::duration<double>(Clock::now() - s->t0).count(); }

skt_config jobConfig(const skt_server* s, const skt_job_request& r) {
    skt_config c = s->opt.member_config;
    c.nx = r.nx;
    c.ny = r.ny;
    c.nz = r.nz;
    c.dx = r.dx;
    c.device = s->opt.device;
    return c;
}

void pushEvent(skt_server* s, const Job& j, int kind, const DgPin* d) {
    skt_job_event ev{};
    ev.job = j.id;
    ev.kind = kind;
    ev.steps = j.steps;
    ev.sim_time = (float)((double)j.steps * s->opt.dt);
    ev.wall_s = now(s);
    ev.submitted_s = j.submitted;
    if (d) {
        ev.diag.max_updraft = d->w;
        ev.diag.cloud_top_level = d->level;
        ev.diag.cloud_top_height = d->level >= 0 ? (float)d->level * j.req.dx : 0.0f;
        ev.diag.surface_rain = d->rain;
    }
    s->events.push_back(ev);
}

// Builds the job's environment (its initial state, generated on the host by
// a worker thread or here, then init) and adds it to the ensemble. The host
// arrays are freed once uploaded.
int admit(skt_server* s, Job& j) {
    int rc;
    try {
        if (!j.prep) j.prep = startPrep(j.req);
        rc = j.prep->done.get();
    } catch (...) {  // no thread or no memory for one
        rc = SKT_ERR_NOMEM;
    }
    if (rc) {
        j.prep.reset();
        return serr(s, rc, "generating a job's initial state");
    }
    rc = skt_env_create(&j.cfg, &j.env);
    if (rc) {
        j.prep.reset();
        return serr(s, rc, "creating a job's environment");
    }
    Prep& q = *j.prep;
    rc = skt_env_init(j.env, q.th.data(), q.vel.data(), q.qv.data(), q.gT.data(), q.gP.data(), q.p.data(),
                      q.pS.data());
    j.prep.reset();
    if (!rc) rc = skt_batch_add(s->batch, j.env, &j.member);
    if (rc) {
        const char* m = skt_env_last_error(j.env);
        std::string what = std::string("initialising a job: ") + (m ? m : "") + " " +
                           (skt_batch_last_error(s->batch) ? skt_batch_last_error(s->batch) : "");
        skt_env_destroy(j.env);
        j.env = nullptr;
        return serr(s, rc, what);
    }
    return SKT_OK;
}

// Starts the initial states of the first kPrefetch queued jobs. A job that
// first fit admits from further back is generated at admission.
void prefetch(skt_server* s) {
    int k = 0;
    for (Job& j : s->queue) {
        if (k++ >= kPrefetch) break;
        if (j.prep) continue;
        try {
            j.prep = startPrep(j.req);
        } catch (...) {  // no thread now: admission generates it
            return;
        }
    }
}

}  // namespace

extern "C" {

void skt_server_options_default(skt_server_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->dt = 0.5f;
    o->hbm_budget_bytes = 0;
    o->max_active = 64;
    o->cloud_threshold = 1e-5f;
    skt_config_default(&o->member_config, 64, 64, 64, 100.0f);
    // the mixed-precision form: BF16 hydrometeors, multigrid-preconditioned
    // CG to 1e-6 with two FP64 refinement passes
    o->member_config.pcg_mode = SKT_PCG_STANDARD;
    o->member_config.pcg_precon = SKT_PRECON_MG;
    o->member_config.pcg_tol = 1e-6f;
    o->member_config.pcg_max_iter = 20;
    o->member_config.tracer_format = SKT_FMT_BF16;
    o->member_config.pcg_refine = 2;
    o->member_config.pcg_refine_tol = 1e-3f;
    skt_batch_options_default(&o->batch);
}

int skt_server_create(const skt_server_options* o, skt_server** out) {
    if (!o || !out || !(o->dt > 0.0f) || o->max_active < 1 || o->hbm_budget_bytes < 0) return SKT_ERR_ARG;
    *out = nullptr;
    skt_server* s = new skt_server();
    s->opt = *o;
    s->t0 = Clock::now();
    auto bail = [&](int rc) {
        if (s->pin) (void)hipHostFree(s->pin);
        if (s->batch) skt_batch_destroy(s->batch);
        delete s;
        return rc;
    };
    if (hipSetDevice(o->device) != hipSuccess) return bail(SKT_ERR_HIP);
    s->budget = o->hbm_budget_bytes;
    if (s->budget == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return bail(SKT_ERR_HIP);
        s->budget = (int64_t)(0.9 * (double)fr);
    }
    int rc = skt_batch_create(o->device, &s->batch);
    if (rc) return bail(rc);
    rc = skt_batch_set_options(s->batch, &o->batch);
    if (rc) return bail(rc);
    if (hipHostMalloc((void**)&s->pin, sizeof(DgPin) * (size_t)o->max_active, hipHostMallocDefault) != hipSuccess)
        return bail(SKT_ERR_NOMEM);
    *out = s;
    return SKT_OK;
}

int skt_server_destroy(skt_server* s) {
    if (!s) return SKT_ERR_ARG;
    if (s->batch) {
        (void)skt_batch_synchronize(s->batch);
        for (Job& j : s->active) (void)skt_batch_remove(s->batch, j.member);
        skt_batch_destroy(s->batch);
    }
    for (Job& j : s->active)
        if (j.env) skt_env_destroy(j.env);
    if (s->pin) (void)hipHostFree(s->pin);
    delete s;
    return SKT_OK;
}

const char* skt_server_last_error(const skt_server* s) { return s ? s->err.c_str() : "null server"; }

int skt_server_submit(skt_server* s, const skt_job_request* r, int64_t* job_id) {
    if (!s || !r) return SKT_ERR_ARG;
    if (!(r->sim_seconds > 0.0f) || !(r->report_every > 0.0f))
        return serr(s, SKT_ERR_ARG, "a job needs a positive duration and report interval");
    Job j;
    j.req = *r;
    j.cfg = jobConfig(s, *r);
    int rc = skt_env_footprint(&j.cfg, &j.bytes);
    if (rc) return serr(s, rc, "invalid job grid or configuration");
    if (j.bytes > s->budget) return serr(s, SKT_ERR_ARG, "the job's footprint exceeds the server's HBM budget");
    j.total = std::max(1, (int)std::lround((double)r->sim_seconds / s->opt.dt));
    j.report_steps = std::max(1, (int)std::lround((double)r->report_every / s->opt.dt));
    j.id = s->next_id++;
    j.submitted = now(s);
    if (job_id) *job_id = j.id;
    s->queue.push_back(j);
    prefetch(s);
    return SKT_OK;
}

int skt_server_step(skt_server* s) {
    if (!s) return SKT_ERR_ARG;
    const Clock::time_point tb = Clock::now();
    if (hipSetDevice(s->opt.device) != hipSuccess) return serr(s, SKT_ERR_HIP, "hipSetDevice");
    // admission, first fit in arrival order
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        if ((int)s->active.size() >= s->opt.max_active) break;
        if (s->bytes + it->bytes > s->budget) {
            ++it;
            continue;
        }
        Job j = *it;
        const Clock::time_point ta = Clock::now();
        int rc = admit(s, j);
        s->admit_s += std::chrono::duration<double>(Clock::now() - ta).count();
        if (rc) return rc;
        s->changed = true;
        s->bytes += j.bytes;
        pushEvent(s, j, SKT_EVENT_ADMITTED, nullptr);
        s->active.push_back(j);
        it = s->queue.erase(it);
    }
    prefetch(s);
    if (s->active.empty()) {
        s->busy += std::chrono::duration<double>(Clock::now() - tb).count();
        return SKT_OK;
    }
    const Clock::time_point tp = Clock::now();
    int rc = skt_batch_step(s->batch, s->opt.dt, 1.0f);
    if (rc) return serr(s, rc, std::string("ensemble step: ") + skt_batch_last_error(s->batch));
    if (s->changed) s->replan_s += std::chrono::duration<double>(Clock::now() - tp).count();
    s->changed = false;
    s->steps++;
    // reports: diagnostics of every job at a boundary, then one wait
    std::vector<int> rep;
    bool retire = false;
    for (size_t k = 0; k < s->active.size(); k++) {
        Job& j = s->active[k];
        j.steps++;
        s->cell_steps += (double)j.req.nx * j.req.ny * j.req.nz;
        if (j.steps % j.report_steps == 0 || j.steps >= j.total) rep.push_back((int)k);
        retire |= j.steps >= j.total;
    }
    for (int k : rep) {
        EnvState& e = s->active[k].env->s;
        if (launch_diagnostics(e, s->opt.cloud_threshold) != hipSuccess ||
            hipMemcpyAsync(&s->pin[k].w, &e.sc->dg_w, sizeof(float) + sizeof(int), hipMemcpyDeviceToHost,
                           e.stream) != hipSuccess ||
            hipMemcpyAsync(&s->pin[k].rain, &e.sc->dg_rain, sizeof(double), hipMemcpyDeviceToHost, e.stream) !=
                hipSuccess)
            return serr(s, SKT_ERR_HIP, "report diagnostics");
    }
    if (!rep.empty() || retire) {
        rc = skt_batch_synchronize(s->batch);
        if (rc) return serr(s, rc, std::string("ensemble: ") + skt_batch_last_error(s->batch));
    }
    for (int k : rep) {
        const Job& j = s->active[k];
        pushEvent(s, j, j.steps >= j.total ? SKT_EVENT_DONE : SKT_EVENT_REPORT, &s->pin[k]);
    }
    // retirement (the reports of this step are already in the events)
    for (size_t k = 0; k < s->active.size();) {
        Job& j = s->active[k];
        if (j.steps < j.total) {
            k++;
            continue;
        }
        rc = skt_batch_remove(s->batch, j.member);
        if (rc) return serr(s, rc, "retiring a job");
        s->changed = true;
        skt_env_destroy(j.env);
        s->bytes -= j.bytes;
        s->completed++;
        s->active.erase(s->active.begin() + (long)k);
    }
    s->busy += std::chrono::duration<double>(Clock::now() - tb).count();
    return SKT_OK;
}

int skt_server_poll(skt_server* s, skt_job_event* out, int32_t capacity, int32_t* count) {
    if (!s || !count || capacity < 0 || (capacity > 0 && !out)) return SKT_ERR_ARG;
    int n = 0;
    while (n < capacity && !s->events.empty()) {
        out[n++] = s->events.front();
        s->events.pop_front();
    }
    *count = n;
    return SKT_OK;
}

int skt_server_status_get(skt_server* s, skt_server_status* o) {
    if (!s || !o) return SKT_ERR_ARG;
    o->active = (int32_t)s->active.size();
    o->queued = (int32_t)s->queue.size();
    o->bytes_in_use = s->bytes;
    o->budget_bytes = s->budget;
    o->steps = s->steps;
    o->completed = s->completed;
    o->cell_steps = s->cell_steps;
    o->busy_s = s->busy;
    o->admit_s = s->admit_s;
    o->replan_s = s->replan_s;
    return SKT_OK;
}

}  // extern "C"
