// Ensembles: several environments (members, each its own grid) advanced by
// one launch per phase (include/skt.h, skt_batch_*).
//
// The plan of an ensemble step is recorded once: each member's updateGPU runs
// with a recorder attached (launches are recorded, not issued, and the values
// the plain step reads back from the device stay on the device), the members'
// launch lists are checked to be the same sequence of bodies, and launch i of
// the plan becomes one ensemble launch of body i over the tiles of every
// member (k_batch / k_batch1 in skt_launch.h). Each member keeps its own tile
// decomposition, arguments and reduction partials, so its results are those
// of skt_env_update bit for bit. The plan is replayed every step, as a HIP
// graph by default; the per-step inputs it reads (the irradiance of each
// member's time of day) are copied to the device before each replay.
//
// Launch width: each body runs either one block per work item or a
// persistent grid of one or two rounds of resident blocks striding over the
// work list. skt_batch_autotune times the three on the ensemble and keeps the
// fastest per launch of the plan, keyed by the body, its index in the plan
// and the members' grid shapes.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <string>
#include <vector>
#include "skt_internal.h"

using namespace skt;

namespace {

constexpr int kModes = 3;  // 0: one block per work item; 1, 2: persistent, 1 or 2 rounds

struct PlanLaunch {
    RecLaunch proto;  // member 0's record (body, block, argument size)
    size_t args_off = 0, args_stride = 0;
    size_t wm_off = 0;  // unsigned index of start[] in the work-map arena
    int M = 0;
    unsigned total = 0;
    int mode = 0;
    unsigned grid = 0;
};

}  // namespace

struct skt_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    struct Member {
        skt_env* env;
        int32_t id;
        hipStream_t prev_stream;
    };
    std::vector<Member> members;
    int32_t next_id = 1;
    skt_batch_options opt{};
    int cus = 0;
    // plan
    bool dirty = true;
    float plan_dt = 0.0f, plan_speed = 0.0f;
    std::vector<PlanLaunch> plan;
    unsigned char* dargs = nullptr;
    size_t dargs_cap = 0;
    unsigned* dwm = nullptr;
    size_t dwm_cap = 0;
    float* dirr = nullptr;
    int irr_cap = 0;
    float* hirr[2] = {nullptr, nullptr};  // pinned staging, alternating per step
    hipEvent_t irr_ev[2] = {nullptr, nullptr};
    int irr_slot = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // launch width per (body, member shapes)
    // launch width per plan launch: (body, index in the plan, member shapes)
    std::map<std::tuple<const void*, size_t, std::string>, int> tuned;
    long long steps = 0;
};

namespace {

int berr(skt_batch* b, int code, const std::string& msg) {
    b->err = msg;
    return code;
}

#define BCHK(b, call)                                                                                 \
    do {                                                                                              \
        hipError_t r_ = (call);                                                                       \
        if (r_ != hipSuccess) return berr((b), SKT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(r_)); \
    } while (0)

std::string shapesKey(const skt_batch* b) {
    std::string k;
    for (const auto& m : b->members) {
        const EnvState& e = m.env->s;
        k += std::to_string(e.X) + "x" + std::to_string(e.Y) + "x" + std::to_string(e.Z) + ";";
    }
    return k;
}

void dropGraph(skt_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

unsigned gridFor(const skt_batch* b, const PlanLaunch& L, int mode) {
    if (mode == 0) return L.total;
    const int occ = L.proto.occ ? L.proto.occ((int)(L.proto.block.x * L.proto.block.y * L.proto.block.z),
                                             L.proto.shm)
                                : 0;
    const long resident = (long)std::max(occ, 1) * std::max(b->cus, 1);
    return (unsigned)std::min<long>(L.total, resident * mode);
}

// Records every member's step and builds the ensemble plan.
int buildPlan(skt_batch* b, float dt, float speed) {
    dropGraph(b);
    b->plan.clear();
    const int M = (int)b->members.size();
    if (M == 0) return SKT_OK;
    // the device slots of the per-step irradiance
    if (b->irr_cap < M) {
        if (b->dirr) (void)hipFree(b->dirr);
        for (int k = 0; k < 2; k++)
            if (b->hirr[k]) (void)hipHostFree(b->hirr[k]);
        b->dirr = nullptr;
        b->hirr[0] = b->hirr[1] = nullptr;
        b->irr_cap = 0;
        BCHK(b, hipMalloc((void**)&b->dirr, sizeof(float) * M));
        for (int k = 0; k < 2; k++) BCHK(b, hipHostMalloc((void**)&b->hirr[k], sizeof(float) * M, hipHostMallocDefault));
        b->irr_cap = M;
    }
    // precipitation substeps of the member with the most (fall speeds 5, 1, 8
    // m/s as advectPrecip guesses them)
    int precipCap = 1;
    for (const auto& m : b->members) {
        const EnvState& e = m.env->s;
        for (float mv : {5.0f, 1.0f, 8.0f}) {
            const float C = mv * (dt * speed) / e.dx;
            precipCap = std::max(precipCap, int(fminf(ceilf(fabsf(C)), (float)e.cfg.max_substeps)));
        }
    }
    std::vector<Recorder> rec((size_t)M);
    for (int m = 0; m < M; m++) {
        EnvState& e = b->members[m].env->s;
        rec[m].adv_cap = b->opt.advection_substep_cap;
        rec[m].precip_cap = precipCap;
        rec[m].irr = b->dirr + m;
        e.rec = &rec[m];
        const int rc = update(e, dt, speed);
        e.rec = nullptr;
        if (rc) return berr(b, rc, "recording member " + std::to_string(b->members[m].id) + ": " + e.err);
    }
    const size_t L = rec[0].launches.size();
    for (int m = 1; m < M; m++)
        if (rec[m].launches.size() != L)
            return berr(b, SKT_ERR_STATE, "members record different step sequences (solver settings differ)");
    // argument arena and work maps
    std::vector<unsigned char> args;
    std::vector<unsigned> wm;
    std::map<std::vector<unsigned>, size_t> wmIndex;
    const std::string key = shapesKey(b);
    for (size_t i = 0; i < L; i++) {
        // the members whose launch i runs the same body (a body can differ in a
        // variant chosen by the grid, e.g. the vector width of the PCG update
        // for rows whose length is not a multiple of 4): one launch each
        std::vector<std::vector<int>> groups;
        for (int m = 0; m < M; m++) {
            bool placed = false;
            for (auto& g : groups)
                if (rec[g[0]].launches[i].fn == rec[m].launches[i].fn) {
                    g.push_back(m);
                    placed = true;
                    break;
                }
            if (!placed) groups.push_back({m});
        }
        for (const auto& g : groups) {
            const RecLaunch& a = rec[g[0]].launches[i];
            for (int m : g) {
                const RecLaunch& c = rec[m].launches[i];
                if (a.args.size() != c.args.size() || a.block.x != c.block.x || a.block.y != c.block.y ||
                    a.block.z != c.block.z || a.shm != c.shm)
                    return berr(b, SKT_ERR_STATE, "members record different launches at " + std::to_string(i));
            }
            PlanLaunch P;
            P.proto = a;
            P.proto.args.clear();
            const int G = (int)g.size();
            P.M = G;
            // the device indexes the table as an array of argument records, so
            // the stride is the record's size (a multiple of its alignment)
            const size_t sz = a.args.size();
            P.args_stride = sz;
            P.args_off = (args.size() + 255) / 256 * 256;
            args.resize(P.args_off + P.args_stride * G);
            for (int k = 0; k < G; k++)
                std::memcpy(args.data() + P.args_off + P.args_stride * k, rec[g[k]].launches[i].args.data(), sz);
            // each member's tiles start at a multiple of 8: a tile's work index
            // and its index in the member's own launch agree modulo the 8 XCDs
            std::vector<unsigned> st((size_t)(2 * G + 1));
            unsigned pos = 0;
            for (int k = 0; k < G; k++) {
                st[k] = pos;
                st[G + 1 + k] = rec[g[k]].launches[i].nb;
                pos += (rec[g[k]].launches[i].nb + 7u) / 8u * 8u;
            }
            st[G] = pos;
            P.total = pos;
            if (pos == 0) continue;  // no member has tiles in this launch
            auto it = wmIndex.find(st);
            if (it == wmIndex.end()) {
                it = wmIndex.emplace(st, wm.size()).first;
                wm.insert(wm.end(), st.begin(), st.end());
            }
            P.wm_off = it->second;
            auto t = b->tuned.find({(const void*)P.proto.fn, b->plan.size(), key});
            P.mode = t != b->tuned.end() ? t->second : b->opt.launch_mode;
            P.grid = gridFor(b, P, P.mode);
            b->plan.push_back(std::move(P));
        }
    }
    if (args.size() > b->dargs_cap) {
        if (b->dargs) (void)hipFree(b->dargs);
        b->dargs = nullptr;
        b->dargs_cap = 0;
        BCHK(b, hipMalloc((void**)&b->dargs, args.size()));
        b->dargs_cap = args.size();
    }
    if (wm.size() * sizeof(unsigned) > b->dwm_cap) {
        if (b->dwm) (void)hipFree(b->dwm);
        b->dwm = nullptr;
        b->dwm_cap = 0;
        BCHK(b, hipMalloc((void**)&b->dwm, wm.size() * sizeof(unsigned)));
        b->dwm_cap = wm.size() * sizeof(unsigned);
    }
    BCHK(b, hipMemcpyAsync(b->dargs, args.data(), args.size(), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipMemcpyAsync(b->dwm, wm.data(), wm.size() * sizeof(unsigned), hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    b->plan_dt = dt;
    b->plan_speed = speed;
    b->dirty = false;
    return SKT_OK;
}

// Issues launch i of the plan (width 'mode' unless given).
hipError_t issue(skt_batch* b, const PlanLaunch& P, int mode = -1) {
    const unsigned* st = b->dwm + P.wm_off;
    const WorkMap wm{st, st + P.M + 1, P.M, P.total};
    const unsigned grid = mode < 0 ? P.grid : gridFor(b, P, mode);
    return P.proto.fn(b->dargs + P.args_off, wm, grid, P.proto.block, P.proto.shm, b->stream);
}

int issuePlan(skt_batch* b) {
    for (const auto& P : b->plan) BCHK(b, issue(b, P));
    return SKT_OK;
}

int captureGraph(skt_batch* b) {
    dropGraph(b);
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = issuePlan(b);
    hipGraph_t g = nullptr;
    const hipError_t r = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (r != hipSuccess) return berr(b, SKT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(r));
    b->graph = g;
    BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return SKT_OK;
}

// Per-step inputs and the ground-change path, before the plan runs.
int beforeStep(skt_batch* b) {
    const int M = (int)b->members.size();
    for (auto& m : b->members) {
        EnvState& e = m.env->s;
        if (e.params.ground_changed) {
            const int rc = editor_data(e);
            if (rc) return berr(b, rc, e.err);
        }
    }
    const int k = b->irr_slot;
    BCHK(b, hipEventSynchronize(b->irr_ev[k]));  // the copy that last read this staging slot
    for (int m = 0; m < M; m++) b->hirr[k][m] = irradianceOf(b->members[m].env->s.params);
    BCHK(b, hipMemcpyAsync(b->dirr, b->hirr[k], sizeof(float) * M, hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipEventRecord(b->irr_ev[k], b->stream));
    b->irr_slot ^= 1;
    return SKT_OK;
}

void afterBatchStep(skt_batch* b, float dt, float speed) {
    for (auto& m : b->members) {
        EnvState& e = m.env->s;
        afterStep(e, dt, speed);
        e.stats_on_device = true;
        e.pcg_stats_pending = false;
    }
    b->steps++;
}

}  // namespace

extern "C" {

void skt_batch_options_default(skt_batch_options* o) {
    if (!o) return;
    o->use_graph = 1;
    o->advection_substep_cap = 4;
    o->launch_mode = 0;
}

int skt_batch_create(int32_t device, skt_batch** out) {
    if (!out) return SKT_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return SKT_ERR_HIP;
    skt_batch* b = new skt_batch();
    b->device = device;
    skt_batch_options_default(&b->opt);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return SKT_ERR_HIP;
    }
    for (int k = 0; k < 2; k++)
        if (hipEventCreateWithFlags(&b->irr_ev[k], hipEventDisableTiming) != hipSuccess) {
            skt_batch_destroy(b);
            return SKT_ERR_HIP;
        }
    if (hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) b->cus = 0;
    *out = b;
    return SKT_OK;
}

int skt_batch_destroy(skt_batch* b) {
    if (!b) return SKT_ERR_ARG;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto& m : b->members) m.env->s.stream = m.prev_stream;
    dropGraph(b);
    if (b->dargs) (void)hipFree(b->dargs);
    if (b->dwm) (void)hipFree(b->dwm);
    if (b->dirr) (void)hipFree(b->dirr);
    for (int k = 0; k < 2; k++) {
        if (b->hirr[k]) (void)hipHostFree(b->hirr[k]);
        if (b->irr_ev[k]) (void)hipEventDestroy(b->irr_ev[k]);
    }
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SKT_OK;
}

const char* skt_batch_last_error(const skt_batch* b) { return b ? b->err.c_str() : "null batch"; }

int skt_batch_set_options(skt_batch* b, const skt_batch_options* o) {
    if (!b || !o) return SKT_ERR_ARG;
    if (o->advection_substep_cap < 1 || o->launch_mode < 0 || o->launch_mode >= kModes)
        return berr(b, SKT_ERR_ARG, "options out of range");
    b->opt = *o;
    b->dirty = true;
    return SKT_OK;
}

int skt_batch_add(skt_batch* b, skt_env* env, int32_t* member_id) {
    if (!b || !env) return SKT_ERR_ARG;
    EnvState& e = env->s;
    if (!e.initialized) return berr(b, SKT_ERR_STATE, "member not initialised");
    if (e.device != b->device) return berr(b, SKT_ERR_ARG, "member on another device");
    for (const auto& m : b->members) {
        if (m.env == env) return berr(b, SKT_ERR_ARG, "environment already in the ensemble");
        const skt_config &a = m.env->s.cfg, &c = e.cfg;
        if (a.pcg_mode != c.pcg_mode || a.pcg_max_iter != c.pcg_max_iter || a.diffusion_sweeps != c.diffusion_sweeps ||
            a.pcg_refine != c.pcg_refine || a.pcg_precon != c.pcg_precon)
            return berr(b, SKT_ERR_ARG,
                        "members must share the solver mode, iteration limit, refinement passes, preconditioner and diffusion sweeps");
    }
    // the member's work joins the ensemble's stream (after its own pending work)
    hipEvent_t ev = nullptr;
    BCHK(b, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t r = hipEventRecord(ev, e.stream);
    if (r == hipSuccess) r = hipStreamWaitEvent(b->stream, ev, 0);
    (void)hipEventDestroy(ev);
    if (r != hipSuccess) return berr(b, SKT_ERR_HIP, std::string("joining the ensemble stream: ") + hipGetErrorString(r));
    b->members.push_back({env, b->next_id, e.stream});
    e.stream = b->stream;
    if (member_id) *member_id = b->next_id;
    b->next_id++;
    b->dirty = true;
    return SKT_OK;
}

int skt_batch_remove(skt_batch* b, int32_t member_id) {
    if (!b) return SKT_ERR_ARG;
    for (size_t k = 0; k < b->members.size(); k++)
        if (b->members[k].id == member_id) {
            BCHK(b, hipStreamSynchronize(b->stream));
            b->members[k].env->s.stream = b->members[k].prev_stream;
            b->members.erase(b->members.begin() + (long)k);
            b->dirty = true;
            return SKT_OK;
        }
    return berr(b, SKT_ERR_ARG, "no such member");
}

int skt_batch_size(const skt_batch* b, int32_t* n) {
    if (!b || !n) return SKT_ERR_ARG;
    *n = (int32_t)b->members.size();
    return SKT_OK;
}

int skt_batch_step(skt_batch* b, float dt, float speed) {
    if (!b) return SKT_ERR_ARG;
    if (b->members.empty()) return SKT_OK;
    BCHK(b, hipSetDevice(b->device));
    if (b->dirty || dt != b->plan_dt || speed != b->plan_speed) {
        const int rc = buildPlan(b, dt, speed);
        if (rc) return rc;
    }
    int rc = beforeStep(b);
    if (rc) return rc;
    if (b->opt.use_graph) {
        if (!b->exec) {
            rc = captureGraph(b);
            if (rc) return rc;
        }
        BCHK(b, hipGraphLaunch(b->exec, b->stream));
    } else {
        rc = issuePlan(b);
        if (rc) return rc;
    }
    afterBatchStep(b, dt, speed);
    return SKT_OK;
}

int skt_batch_synchronize(skt_batch* b) {
    if (!b) return SKT_ERR_ARG;
    BCHK(b, hipStreamSynchronize(b->stream));
    // a member whose advection needed more substep passes than the plan holds
    for (auto& m : b->members) {
        EnvState& e = m.env->s;
        int over = 0;
        BCHK(b, hipMemcpy(&over, &e.sc->sub_overflow, sizeof(int), hipMemcpyDeviceToHost));
        if (over) {
            const int zero = 0;
            BCHK(b, hipMemcpy(&e.sc->sub_overflow, &zero, sizeof(int), hipMemcpyHostToDevice));
            return berr(b, SKT_ERR_STATE,
                        "member " + std::to_string(m.id) + ": an advection needed more substeps than "
                        "advection_substep_cap; raise it with skt_batch_set_options and repeat the steps");
        }
    }
    return SKT_OK;
}

int skt_batch_autotune(skt_batch* b, float dt, float speed, skt_batch_tuning* out) {
    if (!b) return SKT_ERR_ARG;
    if (b->members.empty()) return SKT_OK;
    BCHK(b, hipSetDevice(b->device));
    if (b->dirty || dt != b->plan_dt || speed != b->plan_speed) {
        const int rc = buildPlan(b, dt, speed);
        if (rc) return rc;
    }
    // one step per launch width, every launch bracketed by events; each is an
    // ordinary step of the ensemble (the width changes no result)
    const size_t L = b->plan.size();
    std::vector<hipEvent_t> ev(L + 1);
    for (auto& x : ev) BCHK(b, hipEventCreate(&x));
    std::vector<double> ms[kModes];
    for (auto& m : ms) m.assign(L, 0.0);
    int rc = SKT_OK;
    for (int mode = 0; mode < kModes && !rc; mode++) {
        rc = beforeStep(b);
        if (rc) break;
        for (size_t i = 0; i < L && !rc; i++) {
            if (hipEventRecord(ev[i], b->stream) != hipSuccess) rc = berr(b, SKT_ERR_HIP, "event record");
            else if (issue(b, b->plan[i], mode) != hipSuccess) rc = berr(b, SKT_ERR_HIP, "tuning launch failed");
        }
        if (rc) break;
        if (hipEventRecord(ev[L], b->stream) != hipSuccess || hipEventSynchronize(ev[L]) != hipSuccess) {
            rc = berr(b, SKT_ERR_HIP, "event synchronize");
            break;
        }
        afterBatchStep(b, dt, speed);
        for (size_t i = 0; i < L; i++) {
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, ev[i], ev[i + 1]);
            ms[mode][i] = t;
        }
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (rc) return rc;
    const std::string key = shapesKey(b);
    // each launch of the plan keeps its own fastest width: a body's launches
    // differ in size (multigrid levels) and in how many members are idle
    double tot[kModes] = {0, 0, 0}, best = 0.0;
    int persistent = 0;
    std::set<const void*> bodies;
    for (size_t i = 0; i < L; i++) {
        int bm = 0;
        for (int mode = 0; mode < kModes; mode++) {
            tot[mode] += ms[mode][i];
            if (ms[mode][i] < ms[bm][i]) bm = mode;
        }
        best += ms[bm][i];
        const void* fn = (const void*)b->plan[i].proto.fn;
        b->tuned[{fn, i, key}] = bm;
        bodies.insert(fn);
        persistent += bm != 0;
    }
    if (out) {
        for (int mode = 0; mode < kModes; mode++) out->step_ms[mode] = tot[mode];
        out->tuned_step_ms = best;
        out->bodies = (int32_t)bodies.size();
        out->persistent_bodies = persistent;
    }
    b->dirty = true;
    return SKT_OK;
}


int skt_batch_profile(skt_batch* b, float dt, float speed, char* out, int64_t cap) {
    if (!b || (cap > 0 && !out)) return SKT_ERR_ARG;
    if (b->members.empty()) return SKT_OK;
    BCHK(b, hipSetDevice(b->device));
    if (b->dirty || dt != b->plan_dt || speed != b->plan_speed) {
        const int rc = buildPlan(b, dt, speed);
        if (rc) return rc;
    }
    const size_t L = b->plan.size();
    std::vector<hipEvent_t> ev(L + 1);
    for (auto& x : ev) BCHK(b, hipEventCreate(&x));
    int rc = beforeStep(b);
    for (size_t i = 0; i < L && !rc; i++) {
        if (hipEventRecord(ev[i], b->stream) != hipSuccess) rc = berr(b, SKT_ERR_HIP, "event record");
        else if (issue(b, b->plan[i]) != hipSuccess) rc = berr(b, SKT_ERR_HIP, "profiled launch failed");
    }
    if (!rc && (hipEventRecord(ev[L], b->stream) != hipSuccess || hipEventSynchronize(ev[L]) != hipSuccess))
        rc = berr(b, SKT_ERR_HIP, "event synchronize");
    std::map<std::string, std::pair<double, int>> per;
    if (!rc) {
        afterBatchStep(b, dt, speed);
        for (size_t i = 0; i < L; i++) {
            float t = 0.0f;
            (void)hipEventElapsedTime(&t, ev[i], ev[i + 1]);
            std::string n = b->plan[i].proto.name ? b->plan[i].proto.name : "?";
            const size_t k = n.find("K = ");
            if (k != std::string::npos) n = n.substr(k + 4, n.size() - k - 5);
            auto& p = per[n];
            p.first += t;
            p.second++;
        }
    }
    for (auto& x : ev) (void)hipEventDestroy(x);
    if (rc) return rc;
    std::string text;
    for (const auto& kv : per)
        text += kv.first + "\t" + std::to_string(kv.second.first) + "\t" + std::to_string(kv.second.second) + "\n";
    if (cap > 0) {
        const size_t n = std::min<size_t>(text.size(), (size_t)cap - 1);
        std::memcpy(out, text.data(), n);
        out[n] = 0;
    }
    return SKT_OK;
}

}  // extern "C"
