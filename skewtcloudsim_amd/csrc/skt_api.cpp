// Host orchestration and C ABI of the skewT cloud step (include/skt.h).
//
// The step sequence follows environmentGPU::updateGPU
// (Game/environment.cu:391-555). Differences in mechanism, not in results:
// one stream per environment instead of a device-wide sync after every
// launch; the PCG scalars (alpha, beta, stop test) live on the device and the
// host only polls the stop flag every few iterations; the advection substep
// count is reduced on the device and read back once per advected field.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "skt_ingest.h"
#include "skt_internal.h"

using namespace skt;

namespace {

int setErr(EnvState& e, int code, const std::string& msg) {
    e.err = msg;
    return code;
}

#define HIPCHK(e, call)                                                                              \
    do {                                                                                             \
        hipError_t r_ = (call);                                                                      \
        if (r_ != hipSuccess) return setErr((e), SKT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(r_)); \
    } while (0)

hipEvent_t evGet(EnvState& e) {
    if (!e.evpool.empty()) {
        hipEvent_t ev = e.evpool.back();
        e.evpool.pop_back();
        return ev;
    }
    hipEvent_t ev = nullptr;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    return ev;
}

// Launch timed with HIP events on the env's stream when its class is
// enabled (every t_stride-th launch of the class): attached to the kernel's
// dispatch for the single-kernel classes (kTimedOnDispatch), else recorded
// around the call.
#define TL(e, cls, call)                                                          \
    do {                                                                          \
        hipEvent_t a_ = nullptr, b_ = nullptr;                                    \
        const bool on_ = !(e).rec && ((e).timing & (1 << (cls))) != 0 &&         \
                         (e).t_seen[cls]++ % (e).t_stride == 0;                    \
        if (on_) {                                                                \
            a_ = evGet(e);                                                        \
            b_ = evGet(e);                                                        \
        }                                                                         \
        const bool ok_ = a_ && b_;                                                \
        if (ok_ && kTimedOnDispatch[cls]) {                                       \
            (e).kevA = a_;                                                        \
            (e).kevB = b_;                                                        \
        } else if (ok_) {                                                         \
            (void)hipEventRecord(a_, (e).stream);                                 \
        }                                                                         \
        const hipError_t r_ = (call);                                             \
        const bool armedLeft_ = (e).kevA != nullptr;                              \
        (e).kevA = (e).kevB = nullptr;                                            \
        if (r_ != hipSuccess) return setErr((e), SKT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(r_)); \
        if (ok_ && !kTimedOnDispatch[cls]) (void)hipEventRecord(b_, (e).stream); \
        if (ok_ && armedLeft_) return setErr((e), SKT_ERR_STATE, "timed launch issued no kernel: " #call); \
        if (ok_) (e).pending.push_back(EnvState::TimedLaunch{cls, a_, b_});        \
        if ((e).pending.size() >= 4096) {                                         \
            const int rt_ = resolveTiming(e);                                     \
            if (rt_) return rt_;                                                  \
        }                                                                         \
    } while (0)

int resolveTiming(EnvState& e) {
    if (e.pending.empty()) return SKT_OK;
    HIPCHK(e, hipStreamSynchronize(e.stream));
    for (auto& t : e.pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            e.t_ms[t.cls] += ms;
            e.t_count[t.cls]++;
        }
        e.evpool.push_back(t.a);
        e.evpool.push_back(t.b);
    }
    e.pending.clear();
    return SKT_OK;
}

long fieldLen(const EnvState& e, int f) {
    if (f <= SKT_PRESSURE_PROJ) return e.N;
    if (f <= SKT_CLOUD_COVER) return e.G;
    return e.Y;
}
bool validField(int f) { return f >= 0 && f < SKT_FIELD_COUNT; }
bool is16(const EnvState& e, int f) { return e.fst[f].fmt != FMT_F32; }
size_t fieldBytes(const EnvState& e, int f) { return (is16(e, f) ? 2 : 4) * (size_t)fieldLen(e, f); }

}  // namespace

// irridianceGPU, Game/environment.cu:1021-1034 (host float arithmetic).
float skt::irradianceOf(const skt_env_params& p) {
    const float Gs = 1361.0f * p.sun_strength;
    const float rd = 1 + 0.034f * cosf(2 * 3.14159265359f * p.day / 365);
    const float sd = 0.409f * sinf(2 * 3.14159265359f / 365 * (p.day - 81));
    const float th = p.time / 3600;
    const float lr = p.longitude * 0.01745329251994329576923690768489f;
    const float sr = (th - (6.0f + 14.0f / 2.0f)) * (3.14159265359f / 12.0f);
    return std::max(0.0f, Gs * rd * (sinf(lr) * sinf(sd) + cosf(lr) * cosf(sd) * cosf(sr)));
}

namespace {

int readScalars(EnvState& e) {
    HIPCHK(e, hipMemcpyAsync(e.sc_host, e.sc, sizeof(Scalars), hipMemcpyDeviceToHost, e.stream));
    HIPCHK(e, hipStreamSynchronize(e.stream));
    return SKT_OK;
}

// diffuseGPU, Game/environment.cu:597-673 (boundary table :602-639).
void diffusionBounds(EnvState& e, int type, const float** prof, Bounds* b) {
    *b = Bounds{BC_NEUMANN, BC_DIRICHLET, BC_DIRICHLET};
    switch (type) {
        case 0: *prof = e.f[SKT_ISEN_TEMP]; *b = Bounds{BC_CUSTOM, BC_CUSTOM, BC_CUSTOM}; break;
        case 1: *prof = e.f[SKT_ISEN_VAPOR]; *b = Bounds{BC_CUSTOM, BC_CUSTOM, BC_DIRICHLET}; break;
        case 2: *prof = e.f[SKT_DEFAULT_VELX]; *b = Bounds{BC_CUSTOM, BC_CUSTOM, BC_DIRICHLET}; break;
        case 3: *prof = nullptr; break;
        case 4: *prof = e.f[SKT_DEFAULT_VELZ]; *b = Bounds{BC_CUSTOM, BC_CUSTOM, BC_DIRICHLET}; break;
        default: *prof = e.zeros; break;
    }
}

// Diffusion of n independent fields (each reads only itself, the ground
// temperature and pressure), swept together: field j ping-pongs with work
// buffer j. Identical to diffusing them one after the other. A field stored
// in 16 bits is read by the first sweep and written by the last; the sweeps
// in between run in FP32 between work buffers j and j + SKT_BATCH, so the
// field is rounded to its storage once per diffusion, not once per sweep.
static_assert(SKT_NWORK >= 2 * SKT_BATCH, "two float buffers per diffused field");
int diffuseFields(EnvState& e, const int* flds, const int* types, int n, float dt) {
    const float k = 0.005f * dt / (e.dx * e.dx);
    const int sweeps = e.cfg.diffusion_sweeps;
    DiffBatch db{};
    db.n = n;
    for (int j = 0; j < n; j++) {
        diffusionBounds(e, types[j], &db.f[j].prof, &db.f[j].b);
        db.f[j].type = types[j];
    }
    const Store f32 = storeF32();
    for (int s = 0; s < sweeps; s++) {
        for (int j = 0; j < n; j++) {
            DiffField& d = db.f[j];
            float* fld = e.f[flds[j]];
            const Store sf = e.fst[flds[j]];
            if (sf.fmt == FMT_F32) {
                float* bufs[2] = {fld, e.work[j]};
                d.in = bufs[s & 1];
                d.out = bufs[(s + 1) & 1];
                d.sin = d.sout = f32;
            } else {
                float *A = e.work[j], *B = e.work[j + SKT_BATCH];
                d.in = s == 0 ? fld : ((s & 1) ? A : B);
                d.sin = s == 0 ? sf : f32;
                const bool toField = sweeps > 1 && s == sweeps - 1;
                d.out = toField ? fld : ((s & 1) ? B : A);
                d.sout = toField ? sf : f32;
            }
        }
        TL(e, TC_DIFFUSE, launch_diffuse_batch(e, db, k));
    }
    for (int j = 0; j < n; j++) {
        const Store sf = e.fst[flds[j]];
        if (sf.fmt == FMT_F32 ? (sweeps & 1) : sweeps == 1)
            TL(e, TC_DIFFUSE, launch_copy_sky(e, e.work[j], e.f[flds[j]], f32, sf));
    }
    return SKT_OK;
}
int diffuse(EnvState& e, int fld, int type, float dt) { return diffuseFields(e, &fld, &type, 1, dt); }

// advectPPMWGPU, Game/environment.cu:692-824. The five Strang passes rotate
// through three work buffers per field so that no pass writes what it reads;
// a velocity component being advected is read in its current (partially
// advected) state, as in the reference (quirk A11), and the result lands in
// place. Fields that are not velocity components are independent given the
// velocities and move together in one launch per pass; the substep count
// comes from the shared velocities, so it is the same for all of them.
struct AdvField {
    int fld;
    const float* prof;
    Bounds b;
};
// The substep count from the current velocities, launched ahead of the
// advection that uses it: the count is copied to pinned host memory behind an
// event, so the host reads it after queueing the work in between (the
// diffusion of the advected fields) instead of draining the device. Valid
// until the velocities change (advectFields takes it while mv_ready holds).
int startMaxvel(EnvState& e, float dt) {
    TL(e, TC_MAXVEL, launch_maxvel(e, dt));
    if (e.rec) {  // the passes read the count on the device
        e.mv_ready = true;
        e.mv_dt = dt;
        return SKT_OK;
    }
    HIPCHK(e, hipMemcpyAsync(e.mv_snap, &e.sc->maxvel, sizeof(MaxvelSnap), hipMemcpyDeviceToHost, e.stream));
    HIPCHK(e, hipEventRecord(e.mv_ev, e.stream));
    e.mv_ready = true;
    e.mv_dt = dt;
    return SKT_OK;
}
int advectFields(EnvState& e, const AdvField* fs, int n, float dt) {
    int nsub;
    float dts;
    if (e.rec) {
        // recorded ensemble step: rec->adv_cap substep passes, each reading
        // the substep length on the device and skipped past the member's count
        if (!(e.mv_ready && e.mv_dt == dt)) TL(e, TC_MAXVEL, launch_maxvel(e, dt));
        nsub = e.rec->adv_cap;
        dts = 0.0f;
    } else if (e.mv_ready && e.mv_dt == dt) {
        HIPCHK(e, hipEventSynchronize(e.mv_ev));
        nsub = e.mv_snap->nsub;
        dts = e.mv_snap->dtsub;
    } else {
        TL(e, TC_MAXVEL, launch_maxvel(e, dt));
        int rc = readScalars(e);
        if (rc) return rc;
        nsub = e.sc_host->nsub;
        dts = e.sc_host->dtsub;
    }
    for (int j = 0; j < n; j++) {
        if (e.stats.n_advect < 16) e.stats.substeps[e.stats.n_advect] = nsub;
        e.stats.n_advect++;
    }
    const float h = dts * 0.5f;
    float *U = e.f[SKT_VELX], *V = e.f[SKT_VELY], *W = e.f[SKT_VELZ];
    PpmBatch pb[5]{};
    const Store f32 = storeF32();
    for (int j = 0; j < n; j++) {
        const int fld = fs[j].fld;
        float* A = e.f[fld];
        const Store sa = e.fst[fld];  // the work buffers are float
        float *T1 = e.work[3 * j], *T2 = e.work[3 * j + 1], *T3 = e.work[3 * j + 2];
        // At ground positions each pass sees what the reference's buffers hold
        // there: the field's own ground values in passes 1, 3 and 5 (they read
        // advectArray, whose ground cells no pass writes) and 0 in passes 2 and
        // 4 (they read m_outputArray, written only at sky cells).
        const PpmField f0{A, T1, fld == SKT_VELX ? A : U, fs[j].prof, fs[j].b, A, sa, f32, sa};
        const PpmField f1{T1, T2, fld == SKT_VELY ? A : V, fs[j].prof, fs[j].b, nullptr, f32, f32, f32};
        const PpmField f2{T2, T1, fld == SKT_VELZ ? T2 : W, fs[j].prof, fs[j].b, A, f32, f32, sa};
        const PpmField f3{T1, T3, fld == SKT_VELY ? T2 : V, fs[j].prof, fs[j].b, nullptr, f32, f32, f32};
        const PpmField f4{T3, A, fld == SKT_VELX ? T3 : U, fs[j].prof, fs[j].b, A, f32, sa, sa};
        pb[0].f[j] = f0;
        pb[1].f[j] = f1;
        pb[2].f[j] = f2;
        pb[3].f[j] = f3;
        pb[4].f[j] = f4;
    }
    for (int q = 0; q < 5; q++) pb[q].n = n;
    const int dir[5] = {0, 1, 2, 1, 0};
    const float pdt[5] = {h, h, dts, h, h};
    for (int s = 0; s < nsub; s++)
        for (int q = 0; q < 5; q++) {
            if (e.rec) {
                pb[q].dsc = e.sc;
                pb[q].sub = s;
                pb[q].half = dir[q] != 2;
            }
            TL(e, TC_PPM, launch_ppm_batch(e, pb[q], dir[q], pdt[q]));
        }
    return SKT_OK;
}
int advect(EnvState& e, int fld, const float* prof, Bounds b, float dt) {
    const AdvField f{fld, prof, b};
    return advectFields(e, &f, 1, dt);
}

// advectPrecip, Game/environment.cu:826-849 (guessed fall speeds 5/1/8 m/s).
// Each field falls on its own substeps; the fields only read themselves and
// the (unchanged) temperature, pressure and vapour, so they share launches.
int precipFields(EnvState& e, const int* flds, const int* types, int n, float dt) {
    int sub[3], maxsub = 0;
    for (int j = 0; j < n; j++) {
        const float maxVel = types[j] == 0 ? 5.0f : (types[j] == 1 ? 1.0f : 8.0f);
        const float C = maxVel * dt / e.dx;
        sub[j] = int(fminf(ceilf(fabsf(C)), (float)e.cfg.max_substeps));
        maxsub = std::max(maxsub, sub[j]);
    }
    // a recorded ensemble step issues the ensemble's largest count; the extra
    // launches carry step 0 for every field and return at once
    if (e.rec) maxsub = std::max(maxsub, e.rec->precip_cap);
    for (int s = 0; s < maxsub; s++) {
        PrecipBatch pbt{};
        pbt.n = n;
        for (int j = 0; j < n; j++) {
            pbt.q[j] = e.f[flds[j]];
            pbt.st[j] = e.fst[flds[j]];
            pbt.type[j] = types[j];
            pbt.dts[j] = s < sub[j] ? dt / sub[j] : 0.0f;
        }
        TL(e, TC_PRECIP, launch_precip_batch(e, pbt));
    }
    return SKT_OK;
}
int precip(EnvState& e, int fld, int type, float dt) { return precipFields(e, &fld, &type, 1, dt); }

// calculatePressureProject, Game/environment.cu:898-987. The scalar state
// lives on the device (two slots) and every kernel of a finished solve
// returns at once. After each chunk of 25 iterations (4 with the multigrid
// preconditioner) the state is copied to
// pinned host memory behind an event, and the host waits for the copy of the
// chunk before: the device always has the current chunk queued while the
// host decides whether to enqueue the next. A solve that converges in chunk
// c stops after chunk c + 1 has been queued (its kernels exit at once). The
// final state is copied the same way and read by skt_env_get_stats.
//
// With FP64 refinement (pcg_refine passes) the divergence b is kept, the
// first solve's p' starts the FP64 sum x64, and each pass solves for the
// correction from r = b - A x64 formed in FP64; p' ends as float(x64), and the
// final FP64 residual is reduced for the statistics. Every pass is queued
// whole (a pass whose residual already meets the shared stop value runs no
// iteration), so the device holds the control flow and a recorded step
// replays it.
int solveOnce(EnvState& e, int pass, int* slot_out) {
    const bool mg = e.mg_depth > 0;
    TL(e, TC_OTHER, launch_pcg_start(e, pass));
    if (mg) TL(e, TC_MG, launch_mg_vcycle(e, 1));
    const int maxIt = e.cfg.pcg_max_iter;
    // a multigrid iteration costs a few Jacobi ones and a solve takes tens,
    // not hundreds: poll in chunks of 4
    const int chunk = mg ? 4 : 25;
    int it = 0, c = 0;
    for (; it < maxIt; it++) {
        TL(e, TC_PCG_A, launch_pcg_A(e, it));
        TL(e, TC_PCG_B, launch_pcg_B(e, it));
        if (mg) TL(e, TC_MG, launch_mg_vcycle(e, 0));
        // a recorded ensemble step holds every iteration (those after the
        // member's solve has finished return at once)
        if (!e.rec && (it + 1) % chunk == 0 && it + 1 < maxIt) {
            const int w = c & 1;
            HIPCHK(e, hipMemcpyAsync(&e.pcg_snap[w], &e.sc->st[1], sizeof(PcgState), hipMemcpyDeviceToHost,
                                     e.stream));
            HIPCHK(e, hipEventRecord(e.snap_ev[w], e.stream));
            if (c > 0) {
                HIPCHK(e, hipEventSynchronize(e.snap_ev[w ^ 1]));
                if (e.pcg_snap[w ^ 1].done) {
                    it++;
                    break;
                }
            }
            c++;
        }
    }
    int slot = 1;
    if (it > 0) {
        TL(e, TC_OTHER, launch_pcg_final(e, it));
        slot = 0;
    }
    *slot_out = slot;
    return SKT_OK;
}
int solvePressure(EnvState& e, float dt) {
    // the divergence pass also writes 1/density and the Jacobi preconditioner
    // (initAMatrix + initPrecon), which the multigrid levels are built from
    TL(e, TC_OTHER, launch_divergence(e, dt));
    if (e.mg_depth > 0) TL(e, TC_MG, launch_mg_setup(e));
    const int R = e.cfg.pcg_refine;
    if (R > 0) TL(e, TC_OTHER, launch_refine_keep_rhs(e));
    int slot = 1;
    for (int pass = 0; pass <= R; pass++) {
        if (pass > 0) {
            TL(e, TC_OTHER, launch_refine_accum(e, pass == 1, 0));
            TL(e, TC_OTHER, launch_refine_residual(e, 1));
        }
        const int rc = solveOnce(e, R > 0 ? pass : -1, &slot);
        if (rc) return rc;
    }
    if (R > 0) {
        TL(e, TC_OTHER, launch_refine_accum(e, 0, 1));
        TL(e, TC_OTHER, launch_refine_residual(e, 0));
        TL(e, TC_OTHER, launch_refine_report(e));
    }
    if (e.rec) return SKT_OK;  // the ensemble reads the final state from the device
    HIPCHK(e, hipMemcpyAsync(&e.pcg_snap[2], &e.sc->st[slot], sizeof(PcgState), hipMemcpyDeviceToHost, e.stream));
    if (R > 0)
        HIPCHK(e, hipMemcpyAsync(e.ref_snap, &e.sc->ref, sizeof(RefineStats), hipMemcpyDeviceToHost, e.stream));
    HIPCHK(e, hipEventRecord(e.snap_ev[2], e.stream));
    e.pcg_stats_pending = true;
    return SKT_OK;
}

// pressureProject, Game/environment.cu:851-896; with new_pressure the
// calculateNewPressure that follows it in updateGPU (:501) runs per cell in
// the same pass.
int pressureProject(EnvState& e, float dt, int new_pressure) {
    TL(e, TC_OTHER, launch_reset_vel(e));
    int rc = solvePressure(e, dt);
    if (rc) return rc;
    TL(e, TC_OTHER, launch_update_pressure_density(e));
    TL(e, TC_OTHER, launch_apply_projection(e, new_pressure));
    return SKT_OK;
}

}  // namespace
int skt::editor_data(EnvState& e) {
    if (e.params.ground_changed) {
        TL(e, TC_OTHER, launch_editor_reset(e, 9));
        TL(e, TC_OTHER, launch_isen_temp_ground(e));
    }
    return SKT_OK;
}

namespace {

int microSky(EnvState& e, float dt, float speed) {
    TL(e, TC_MICRO, launch_microphysics(e, dt, speed));
    if (e.region_active && !e.rec) {
        TL(e, TC_OTHER, launch_region_finalize(e));
        const unsigned nb = (unsigned)blocks3d(e);
        HIPCHK(e, hipMemcpyAsync(e.region_host, (float*)(e.region + 33L * nb), 33 * sizeof(float),
                                 hipMemcpyDeviceToHost, e.stream));
        HIPCHK(e, hipStreamSynchronize(e.stream));
    }
    return SKT_OK;
}

#define RC(x)                 \
    do {                      \
        int rc_ = (x);        \
        if (rc_) return rc_;  \
    } while (0)

// Clears the launched-ahead substep count when update() leaves, also on errors.
struct MaxvelScope {
    EnvState& e;
    explicit MaxvelScope(EnvState& s) : e(s) { e.mv_ready = false; }
    ~MaxvelScope() { e.mv_ready = false; }
};

}  // namespace

// environmentGPU::updateGPU (Game/environment.cu:391-555). While the env
// holds a recorder (an ensemble step being recorded) the ground-change path
// and the host bookkeeping after the step are left to the ensemble, which
// runs them per member and step.
int skt::update(EnvState& e, float dt, float speed) {
    const float ds = dt * speed;
    MaxvelScope mvs(e);
    e.stats.n_advect = 0;
    if (!e.rec) RC(editor_data(e));
    const float irr = irradianceOf(e.params);
    // ground (Game/environment.cu:420-456) and setTempsAtGround (:469)
    TL(e, TC_OTHER, launch_ground_water_snapshot(e, dt, speed));
    TL(e, TC_OTHER, launch_ground_columns(e, 1 | 2 | 4 | 8 | 16, dt, speed, irr));
    const Bounds bT{BC_CUSTOM, BC_CUSTOM, BC_CUSTOM}, bU{BC_CUSTOM, BC_CUSTOM, BC_DIRICHLET},
        bV{BC_DIRICHLET, BC_DIRICHLET, BC_DIRICHLET}, bW{BC_CUSTOM, BC_CUSTOM, BC_DIRICHLET},
        bQ{BC_DIRICHLET, BC_DIRICHLET, BC_DIRICHLET}, bQv{BC_CUSTOM, BC_CUSTOM, BC_CUSTOM};
    // the velocities are unchanged until buoyancy: theta's substep count is
    // launched ahead of its diffusion
    RC(startMaxvel(e, ds));
    RC(diffuse(e, SKT_POTTEMP, 0, ds));
    RC(advect(e, SKT_POTTEMP, e.f[SKT_ISEN_TEMP], bT, ds));
    e.mv_ready = false;
    TL(e, TC_OTHER, launch_buoyancy(e, ds));
    {
        const int f[3] = {SKT_VELX, SKT_VELY, SKT_VELZ}, t[3] = {2, 3, 4};
        RC(diffuseFields(e, f, t, 3, ds));
    }
    RC(advect(e, SKT_VELX, e.f[SKT_DEFAULT_VELX], bU, ds));
    RC(advect(e, SKT_VELY, e.zeros, bV, ds));
    RC(advect(e, SKT_VELZ, e.f[SKT_DEFAULT_VELZ], bW, ds));
    TL(e, TC_OTHER, launch_init_density(e, 1));
    RC(pressureProject(e, ds, 1));
    // moisture (Game/environment.cu:508-533): each field is diffused, rained
    // out and advected on its own; only the precipitation fall speed reads
    // the vapour field, which is final by then. The velocities are final for
    // the step, so both moisture advections take one substep count, launched
    // ahead of the moisture diffusion.
    RC(startMaxvel(e, ds));
    {
        const int f[6] = {SKT_QV, SKT_QW, SKT_QC, SKT_QR, SKT_QS, SKT_QI}, t[6] = {1, 5, 5, 5, 5, 5};
        RC(diffuseFields(e, f, t, 6, ds));
    }
    RC(advect(e, SKT_QV, e.f[SKT_ISEN_VAPOR], bQv, ds));
    {
        const int f[3] = {SKT_QR, SKT_QS, SKT_QI}, t[3] = {0, 1, 2};
        RC(precipFields(e, f, t, 3, ds));
    }
    {
        const AdvField a[5] = {{SKT_QW, e.zeros, bQ}, {SKT_QC, e.zeros, bQ}, {SKT_QR, e.zeros, bQ},
                               {SKT_QS, e.zeros, bQ}, {SKT_QI, e.zeros, bQ}};
        RC(advectFields(e, a, 5, ds));
    }
    RC(microSky(e, dt, speed));
    if (!e.rec) afterStep(e, dt, speed);
    HIPCHK(e, hipGetLastError());
    return SKT_OK;
}

// Host bookkeeping at the end of updateGPU (Game/environment.cu:535-552).
void skt::afterStep(EnvState& e, float dt, float speed) {
    e.params.time += speed * dt * float(!e.params.pause_diurnal);
    if (e.params.time > 86400.0f) e.params.time = 0.0f;
    e.params.ground_changed = 0;
}

namespace {

void freeAll(EnvState& e) {
    auto fr = [](void* p) {
        if (p) (void)hipFree(p);
    };
    for (int k = 0; k < SKT_FIELD_COUNT; k++) fr(e.f[k]);
    for (int k = 0; k < SKT_NWORK; k++) fr(e.work[k]);
    for (int k = 0; k < 6; k++) fr(e.pcg[k]);
    fr(e.gh);
    fr(e.ghm1);
    fr(e.ghm2);
    fr(e.ghz);
    fr(e.ghEdit);
    fr(e.irho);
    fr(e.rhs);
    fr(e.x64);
    for (int l = 0; l < e.mg_depth; l++) {
        fr(e.mg[l].C);
        fr(e.mg[l].xa);
        if (l > 0) {
            fr(e.mg[l].r);
            fr(e.mg[l].xb);
        }
    }
    fr(e.zeros);
    fr(e.ones);
    fr(e.partials);
    fr(e.sc);
    fr(e.phase);
    fr(e.region);
    for (auto& t : e.pending) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto ev : e.evpool) (void)hipEventDestroy(ev);
    if (e.sc_host) (void)hipHostFree(e.sc_host);
    if (e.mv_snap) (void)hipHostFree(e.mv_snap);
    if (e.mv_ev) (void)hipEventDestroy(e.mv_ev);
    if (e.pcg_snap) (void)hipHostFree(e.pcg_snap);
    if (e.ref_snap) (void)hipHostFree(e.ref_snap);
    for (int k = 0; k < 3; k++)
        if (e.snap_ev[k]) (void)hipEventDestroy(e.snap_ev[k]);
    if (e.own_stream) (void)hipStreamDestroy(e.own_stream);
}

// Device allocation of an env's array; in a dry run (footprint sizing) only
// the bytes are counted.
template <class T>
int dalloc(EnvState& e, T** p, size_t bytes) {
    e.dev_bytes += (long long)bytes;
    if (e.dry) {
        *p = nullptr;
        return SKT_OK;
    }
    hipError_t r = hipMalloc((void**)p, bytes);
    if (r != hipSuccess) return setErr(e, SKT_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(r));
    return SKT_OK;
}

}  // namespace

// ------------------------------------------------------------------ C ABI ---
extern "C" {

void skt_config_default(skt_config* c, int32_t nx, int32_t ny, int32_t nz, float dx) {
    if (!c) return;
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->dx = dx;
    c->device = 0;
    c->pcg_mode = SKT_PCG_REFERENCE;
    c->pcg_tol = 1e-5f;
    c->pcg_max_iter = 200;
    c->diffusion_sweeps = 20;
    c->max_substeps = 100;
    c->maxabs_mode = 0;
    c->noise_seed = 100u;
    c->init_noise = 1;
    c->max_blocks = 0;
    c->tracer_format = SKT_FMT_F32;
    for (int k = 0; k < 5; k++) c->tracer_scale_log2[k] = 16;
    c->pcg_refine = 0;
    c->pcg_refine_tol = 1e-3f;
    c->pcg_precon = SKT_PRECON_JACOBI;
}

// Checks of a configuration (skt_env_create's SKT_ERR_ARG cases).
static bool configOk(const skt_config* cfg) {
    if (cfg->nx < 1 || cfg->ny < 2 || cfg->nz < 1 || !(cfg->dx > 0.0f)) return false;
    const long N = (long)cfg->nx * cfg->ny * cfg->nz;
    if (N > (1L << 31) - 1 || cfg->diffusion_sweeps < 0 || cfg->pcg_max_iter < 0 || cfg->max_substeps < 0 ||
        cfg->max_blocks < 0)
        return false;
    if (cfg->tracer_format < SKT_FMT_F32 || cfg->tracer_format > SKT_FMT_F16) return false;
    for (int k = 0; k < 5; k++)
        if (cfg->tracer_scale_log2[k] < -60 || cfg->tracer_scale_log2[k] > 60) return false;
    // refinement needs the symmetric operator, a solve to refine and a
    // positive relative tolerance
    if (cfg->pcg_refine < 0 || cfg->pcg_refine > 8 ||
        (cfg->pcg_refine > 0 && (cfg->pcg_mode != SKT_PCG_STANDARD || cfg->pcg_max_iter < 1 ||
                                 !(cfg->pcg_refine_tol > 0.0f && cfg->pcg_refine_tol < 1.0f))))
        return false;
    // the multigrid cycle preconditions the symmetric operator only
    if (cfg->pcg_precon != SKT_PRECON_JACOBI &&
        (cfg->pcg_precon != SKT_PRECON_MG || cfg->pcg_mode != SKT_PCG_STANDARD))
        return false;
    return true;
}
// Multigrid levels of a grid: halve each dimension (rounding up) until a
// level holds at most kMgCoarsest cells; false if kMgLevels do not reach it.
static bool mgLevels(EnvState& e) {
    MgLevel* L = e.mg;
    L[0].X = e.X;
    L[0].Y = e.Y;
    L[0].Z = e.Z;
    L[0].N = e.N;
    int d = 1;
    while (L[d - 1].N > kMgCoarsest) {
        if (d == kMgLevels) return false;
        L[d].X = (L[d - 1].X + 1) / 2;
        L[d].Y = (L[d - 1].Y + 1) / 2;
        L[d].Z = (L[d - 1].Z + 1) / 2;
        L[d].N = (long)L[d].X * L[d].Y * L[d].Z;
        d++;
    }
    e.mg_depth = d;
    return true;
}
static void initState(EnvState& e, const skt_config* cfg) {
    e.cfg = *cfg;
    e.X = cfg->nx;
    e.Y = cfg->ny;
    e.Z = cfg->nz;
    e.N = (long)cfg->nx * cfg->ny * cfg->nz;
    e.G = (long)cfg->nx * cfg->nz;
    e.dx = cfg->dx;
    e.device = cfg->device;
    for (int k = 0; k < SKT_FIELD_COUNT; k++) e.fst[k] = storeF32();
    if (cfg->tracer_format != SKT_FMT_F32) {
        for (int k = 0; k < 5; k++) {
            const float sc = std::ldexp(1.0f, cfg->tracer_scale_log2[k]);
            e.fst[SKT_QW + k] = Store{cfg->tracer_format == SKT_FMT_BF16 ? FMT_BF16 : FMT_F16, sc, 1.0f / sc};
        }
        e.mix = true;
    }
}
// Every device array of an env (its whole HBM footprint).
static int allocDevice(EnvState& e) {
    for (int k = 0; k < SKT_FIELD_COUNT; k++)
        if (dalloc(e, &e.f[k], fieldBytes(e, k))) return SKT_ERR_NOMEM;
    for (int k = 0; k < SKT_NWORK; k++)
        if (dalloc(e, &e.work[k], sizeof(float) * std::max(e.N, 2 * e.G))) return SKT_ERR_NOMEM;
    for (int k = 0; k < 6; k++)
        if (dalloc(e, &e.pcg[k], sizeof(float) * e.N)) return SKT_ERR_NOMEM;
    if (e.cfg.pcg_refine > 0 &&
        (dalloc(e, &e.rhs, sizeof(float) * e.N) || dalloc(e, &e.x64, sizeof(double) * e.N)))
        return SKT_ERR_NOMEM;
    if (e.cfg.pcg_precon == SKT_PRECON_MG) {
        if (!mgLevels(e)) return setErr(e, SKT_ERR_ARG, "grid too large for the multigrid levels");
        for (int l = 0; l < e.mg_depth; l++) {
            MgLevel& L = e.mg[l];
            if (dalloc(e, &L.C, sizeof(MgC) * L.N) || dalloc(e, &L.xa, sizeof(float) * L.N)) return SKT_ERR_NOMEM;
            if (l > 0 && (dalloc(e, &L.r, sizeof(float) * L.N) || dalloc(e, &L.xb, sizeof(float) * L.N)))
                return SKT_ERR_NOMEM;
        }
        // level 0 works on the solver's residual and preconditioned residual
        e.mg[0].r = e.pcg[0];
        e.mg[0].xb = e.pcg[1];
    }
    if (dalloc(e, &e.gh, sizeof(int) * e.G) || dalloc(e, &e.ghm1, sizeof(int) * e.G) ||
        dalloc(e, &e.ghm2, sizeof(int) * e.G) || dalloc(e, &e.ghz, sizeof(int) * e.G) ||
        dalloc(e, &e.irho, sizeof(float) * e.N) ||
        dalloc(e, &e.ghEdit, sizeof(int) * e.G) ||
        dalloc(e, &e.zeros, sizeof(float) * e.Y) || dalloc(e, &e.ones, sizeof(float) * e.Y) ||
        dalloc(e, &e.phase, (size_t)e.N))
        return SKT_ERR_NOMEM;
    setup_grids(e, resident_blocks(e.device));
    e.maxBlocks = std::max(blocks3d(e), (int)e.gB.nblocks);
    if (dalloc(e, &e.partials, sizeof(double) * PART_COUNT * (size_t)e.maxBlocks) ||
        dalloc(e, &e.region, sizeof(double) * 33 * (size_t)e.maxBlocks + 64 * sizeof(float)) ||
        dalloc(e, &e.sc, sizeof(Scalars)))
        return SKT_ERR_NOMEM;
    return SKT_OK;
}

int skt_env_footprint(const skt_config* cfg, int64_t* bytes) {
    if (!cfg || !bytes || !configOk(cfg)) return SKT_ERR_ARG;
    EnvState e;
    initState(e, cfg);
    e.dry = true;
    if (hipSetDevice(e.device) != hipSuccess) return SKT_ERR_HIP;
    if (allocDevice(e)) return SKT_ERR_ARG;
    *bytes = e.dev_bytes;
    return SKT_OK;
}
int skt_env_device_bytes(skt_env* env, int64_t* bytes) {
    if (!env || !bytes) return SKT_ERR_ARG;
    *bytes = env->s.dev_bytes;
    return SKT_OK;
}

int skt_env_create(const skt_config* cfg, skt_env** out) {
    if (!cfg || !out) return SKT_ERR_ARG;
    *out = nullptr;
    if (!configOk(cfg)) return SKT_ERR_ARG;
    skt_env* env = new skt_env();
    EnvState& e = env->s;
    initState(e, cfg);
    auto bail = [&](int rc) {
        freeAll(e);
        delete env;
        return rc;
    };
    if (hipSetDevice(e.device) != hipSuccess) return bail(SKT_ERR_HIP);
    if (hipStreamCreateWithFlags(&e.own_stream, hipStreamNonBlocking) != hipSuccess) return bail(SKT_ERR_HIP);
    e.stream = e.own_stream;
    if (const int rc = allocDevice(e)) return bail(rc);
    for (int k = 0; k < SKT_NWORK; k++)
        if (hipMemsetAsync(e.work[k], 0, sizeof(float) * std::max(e.N, 2 * e.G), e.stream) != hipSuccess)
            return bail(SKT_ERR_HIP);
    if (hipHostMalloc((void**)&e.sc_host, sizeof(Scalars), hipHostMallocDefault) != hipSuccess) return bail(SKT_ERR_NOMEM);
    if (hipHostMalloc((void**)&e.pcg_snap, 3 * sizeof(PcgState), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&e.ref_snap, sizeof(RefineStats), hipHostMallocDefault) != hipSuccess)
        return bail(SKT_ERR_NOMEM);
    for (int k = 0; k < 3; k++)
        if (hipEventCreateWithFlags(&e.snap_ev[k], hipEventDisableTiming) != hipSuccess) return bail(SKT_ERR_HIP);
    if (hipHostMalloc((void**)&e.mv_snap, sizeof(MaxvelSnap), hipHostMallocDefault) != hipSuccess)
        return bail(SKT_ERR_NOMEM);
    if (hipEventCreateWithFlags(&e.mv_ev, hipEventDisableTiming) != hipSuccess) return bail(SKT_ERR_HIP);
    // environmentGPU ctor defaults (Game/environment.cu:151-168)
    for (int k = 0; k < SKT_FIELD_COUNT; k++)
        if (hipMemsetAsync(e.f[k], 0, fieldBytes(e, k), e.stream) != hipSuccess) return bail(SKT_ERR_HIP);
    if (hipMemsetAsync(e.gh, 0, sizeof(int) * e.G, e.stream) != hipSuccess ||
        hipMemsetAsync(e.ghm1, 0, sizeof(int) * e.G, e.stream) != hipSuccess ||
        hipMemsetAsync(e.ghm2, 0, sizeof(int) * e.G, e.stream) != hipSuccess ||
        hipMemsetAsync(e.ghz, 0, sizeof(int) * e.G, e.stream) != hipSuccess ||
        hipMemsetAsync(e.sc, 0, sizeof(Scalars), e.stream) != hipSuccess ||
        hipMemsetAsync(e.phase, 0, (size_t)e.N, e.stream) != hipSuccess ||
        hipMemsetAsync(e.irho, 0, sizeof(float) * e.N, e.stream) != hipSuccess)
        return bail(SKT_ERR_HIP);
    for (int k = 0; k < 6; k++)
        if (hipMemsetAsync(e.pcg[k], 0, sizeof(float) * e.N, e.stream) != hipSuccess) return bail(SKT_ERR_HIP);
    if (launch_fill(e, e.f[SKT_QRS], 0.001f, e.G) || launch_fill(e, e.f[SKT_GROUND_P], 1000.0f, e.G) ||
        launch_fill(e, e.f[SKT_GROUND_T], 315.15f, e.G) || launch_fill(e, e.zeros, 0.0f, e.Y) ||
        launch_fill(e, e.ones, 1.0f, e.Y))
        return bail(SKT_ERR_HIP);
    if (hipStreamSynchronize(e.stream) != hipSuccess) return bail(SKT_ERR_HIP);
    // Gamma constants (Game/kernelSky.cu:30-37, Game/microPhysics.cu:31-61)
    // float arguments from b = 0.8f, d = 0.25f; FP64 tgamma rounded once
    const float b = 0.8f, d = 0.25f;
    auto G = [](float x) { return (float)std::tgamma((double)x); };
    e.gm.R48 = G(4.0f + b);
    e.gm.S425 = G(4.0f + d);
    e.gm.gR = G(3 + b);
    e.gm.gER = G((b + 5.0f) / 2.0f);
    e.gm.gS = G(3 + d);
    e.gm.gRC = G(6 + b);
    e.gm.gSS = G((d + 5.0f) / 2.0f);
    e.gm.gI = G(3.5f);
    e.gm.gRI = G(2.75f);
    e.params = skt_env_params{43200.0f, 130, 52.37f, 1.0f, 0, 1};
    std::memset(&e.stats, 0, sizeof(e.stats));
    *out = env;
    return SKT_OK;
}

int skt_env_destroy(skt_env* env) {
    if (!env) return SKT_ERR_ARG;
    (void)hipSetDevice(env->s.device);
    (void)hipStreamSynchronize(env->s.stream);
    freeAll(env->s);
    delete env;
    return SKT_OK;
}

const char* skt_env_last_error(const skt_env* env) { return env ? env->s.err.c_str() : "null environment"; }

int skt_env_set_stream(skt_env* env, void* stream) {
    if (!env) return SKT_ERR_ARG;
    env->s.stream = stream ? (hipStream_t)stream : env->s.own_stream;
    return SKT_OK;
}

int skt_env_synchronize(skt_env* env) {
    if (!env) return SKT_ERR_ARG;
    HIPCHK(env->s, hipStreamSynchronize(env->s.stream));
    return SKT_OK;
}

int skt_env_set_ground_height(skt_env* env, const int32_t* gh) {
    if (!env || !gh) return SKT_ERR_ARG;
    EnvState& e = env->s;
    for (long g = 0; g < e.G; g++)
        if (gh[g] < 0 || gh[g] >= e.Y) return setErr(e, SKT_ERR_ARG, "ground height outside [0, ny)");
    // column maxima over the 3x3 and 5x5 neighbourhoods (clipped at the
    // domain edge) select the stencil kernels' boundary-free fast paths
    std::vector<int32_t> m1((size_t)e.G), m2((size_t)e.G), mz((size_t)e.G);
    for (int z = 0; z < e.Z; z++)
        for (int x = 0; x < e.X; x++) {
            int a = 0, b = 0;
            for (int dz = -2; dz <= 2; dz++)
                for (int dx = -2; dx <= 2; dx++) {
                    const int xx = x + dx, zz = z + dz;
                    if (xx < 0 || xx >= e.X || zz < 0 || zz >= e.Z) continue;
                    const int h = gh[xx + (long)e.X * zz];
                    b = std::max(b, h);
                    if (dx >= -1 && dx <= 1 && dz >= -1 && dz <= 1) a = std::max(a, h);
                }
            m1[x + (size_t)e.X * z] = a;
            m2[x + (size_t)e.X * z] = b;
            // running maximum along z: a column (x, y) of the z walk in the
            // velocity reset is still open at plane z iff y > mz
            const int h = gh[x + (long)e.X * z];
            mz[x + (size_t)e.X * z] = z == 0 ? h : std::max(h, mz[x + (size_t)e.X * (z - 1)]);
        }
    HIPCHK(e, hipMemcpyAsync(e.gh, gh, sizeof(int) * e.G, hipMemcpyHostToDevice, e.stream));
    HIPCHK(e, hipMemcpyAsync(e.ghm1, m1.data(), sizeof(int) * e.G, hipMemcpyHostToDevice, e.stream));
    HIPCHK(e, hipMemcpyAsync(e.ghm2, m2.data(), sizeof(int) * e.G, hipMemcpyHostToDevice, e.stream));
    HIPCHK(e, hipMemcpyAsync(e.ghz, mz.data(), sizeof(int) * e.G, hipMemcpyHostToDevice, e.stream));
    HIPCHK(e, hipStreamSynchronize(e.stream));
    e.params.ground_changed = 1;
    return SKT_OK;
}

int skt_env_get_ground_height(skt_env* env, int32_t* gh) {
    if (!env || !gh) return SKT_ERR_ARG;
    EnvState& e = env->s;
    HIPCHK(e, hipMemcpyAsync(gh, e.gh, sizeof(int) * e.G, hipMemcpyDeviceToHost, e.stream));
    HIPCHK(e, hipStreamSynchronize(e.stream));
    return SKT_OK;
}

// Steps of environmentGPU::init (Game/environment.cu:249-389) after the
// initial fields are on the device.
static int finishInit(EnvState& e) {
    HIPCHK(e, launch_init_density(e, 1));
    HIPCHK(e, launch_profiles_from_column(e));
    if (e.cfg.init_noise) {
        const uint32_t s = e.cfg.noise_seed;
        HIPCHK(e, launch_random(e, e.f[SKT_POTTEMP], -0.1f, 0.1f, s + 0));
        HIPCHK(e, launch_random(e, e.f[SKT_QV], 0.0f, 0.0001f, s + 1));
        HIPCHK(e, launch_random(e, e.f[SKT_VELX], 0.5f, 0.5f, s + 2));
        HIPCHK(e, launch_random(e, e.f[SKT_VELY], -0.2f, 0.2f, s + 3));
        HIPCHK(e, launch_random(e, e.f[SKT_VELZ], 0.5f, 0.5f, s + 4));
    }
    HIPCHK(e, launch_editor_reset(e, 10));
    HIPCHK(e, launch_isen_temp_ground(e));
    HIPCHK(e, hipStreamSynchronize(e.stream));
    e.params = skt_env_params{43200.0f, 130, 52.37f, 1.0f, 0, 1};
    e.initialized = true;
    return SKT_OK;
}

int skt_env_init(skt_env* env, const float* th, const float* vel, const float* qv, float* gT, const float* gP,
                 const float* p, const float* pS) {
    if (!env || !th || !vel || !qv || !gT || !gP || !p || !pS) return SKT_ERR_ARG;
    EnvState& e = env->s;
    HIPCHK(e, hipSetDevice(e.device));
    std::vector<float> comp((size_t)e.N);
    auto upload = [&](int f, const float* src) -> int {
        HIPCHK(e, hipMemcpyAsync(e.f[f], src, sizeof(float) * fieldLen(e, f), hipMemcpyHostToDevice, e.stream));
        return SKT_OK;
    };
    RC(upload(SKT_POTTEMP, th));
    for (int c = 0; c < 3; c++) {
        for (long i = 0; i < e.N; i++) comp[i] = vel[3 * i + c];
        RC(upload(SKT_VELX + c, comp.data()));
        HIPCHK(e, hipStreamSynchronize(e.stream));
    }
    RC(upload(SKT_QV, qv));
    RC(upload(SKT_GROUND_T, gT));
    RC(upload(SKT_GROUND_P, gP));
    RC(upload(SKT_DEFAULT_PRESSURE, pS));
    RC(upload(SKT_PRESSURE, p));
    return finishInit(e);
}

int skt_env_init_sounding(skt_env* env, const skt_sounding_table* table) {
    if (!env) return SKT_ERR_ARG;
    EnvState& e = env->s;
    skt::SoundingLevels L;
    if (int rc = skt::soundingLevels(table, e.Y, e.dx, L)) return setErr(e, rc, skt_input_last_error());
    HIPCHK(e, hipSetDevice(e.device));
    // theta, u, w, qv, p per level, staged in a work array (5*ny <= nx*ny*nz
    // floats for every grid with nx*nz >= 5).
    const size_t Y = (size_t)e.Y;
    if (5 * Y > (size_t)std::max(e.N, 2 * e.G)) return setErr(e, SKT_ERR_ARG, "grid too small for the sounding broadcast");
    std::vector<float> lv(5 * Y);
    for (size_t y = 0; y < Y; y++) {
        lv[y] = L.theta[y];
        lv[Y + y] = L.velX[y];
        lv[2 * Y + y] = L.velZ[y];
        lv[3 * Y + y] = L.qv[y];
        lv[4 * Y + y] = L.p[y];
    }
    HIPCHK(e, hipMemcpyAsync(e.work[0], lv.data(), sizeof(float) * lv.size(), hipMemcpyHostToDevice, e.stream));
    HIPCHK(e, hipMemcpyAsync(e.f[SKT_DEFAULT_PRESSURE], L.pSmall.data(), sizeof(float) * Y, hipMemcpyHostToDevice,
                             e.stream));
    HIPCHK(e, launch_sounding_fill(e, e.work[0], L.groundT, L.groundP));
    HIPCHK(e, hipStreamSynchronize(e.stream));
    return finishInit(e);
}

int skt_env_update(skt_env* env, float dt, float speed) {
    if (!env) return SKT_ERR_ARG;
    if (!env->s.initialized) return setErr(env->s, SKT_ERR_STATE, "update before init");
    env->s.stats_on_device = false;
    return update(env->s, dt, speed);
}

int skt_env_micro_physics_ground(skt_env* env, float dt, float speed, float irr) {
    if (!env) return SKT_ERR_ARG;
    HIPCHK(env->s, launch_ground_columns(env->s, 4 | 8, dt, speed, irr));
    return SKT_OK;
}
int skt_env_micro_physics_sky(skt_env* env, float dt, float speed) {
    if (!env) return SKT_ERR_ARG;
    return microSky(env->s, dt, speed);
}
int skt_env_diffuse(skt_env* env, int32_t field, int32_t type, float dt) {
    if (!env || !validField(field) || field > SKT_PRESSURE || type < 0 || type > 5) return SKT_ERR_ARG;
    return diffuse(env->s, field, type, dt);
}
int skt_env_advect_ground_water(skt_env* env, float dt, float speed) {
    if (!env) return SKT_ERR_ARG;
    HIPCHK(env->s, launch_ground_water_snapshot(env->s, dt, speed));
    return SKT_OK;
}
int skt_env_set_temps_at_ground(skt_env* env, float dt, float speed) {
    if (!env) return SKT_ERR_ARG;
    HIPCHK(env->s, launch_ground_columns(env->s, 16, dt, speed, 0.0f));
    return SKT_OK;
}
int skt_env_advect_ppm(skt_env* env, int32_t field, int32_t profile, skt_bounds bounds, float dt) {
    if (!env || !validField(field) || field > SKT_PRESSURE) return SKT_ERR_ARG;
    if (profile >= 0 && (profile < SKT_ISEN_TEMP || profile >= SKT_FIELD_COUNT)) return SKT_ERR_ARG;
    EnvState& e = env->s;
    e.stats.n_advect = 0;
    const float* prof = profile < 0 ? e.zeros : e.f[profile];
    return advect(e, field, prof, Bounds{bounds.sides, bounds.up, bounds.ground}, dt);
}
int skt_env_advect_precip(skt_env* env, int32_t field, int32_t fall_type, float dt) {
    if (!env || !validField(field) || field > SKT_PRESSURE || fall_type < 0 || fall_type > 2) return SKT_ERR_ARG;
    return precip(env->s, field, fall_type, dt);
}
int skt_env_pressure_project(skt_env* env, float dt) {
    if (!env) return SKT_ERR_ARG;
    return pressureProject(env->s, dt, 0);
}
int skt_env_calculate_pressure_project(skt_env* env, float dt) {
    if (!env) return SKT_ERR_ARG;
    EnvState& e = env->s;
    HIPCHK(e, launch_reset_vel(e));
    return solvePressure(e, dt);
}
int skt_env_editor_data(skt_env* env) {
    if (!env) return SKT_ERR_ARG;
    return editor_data(env->s);
}
int skt_env_irradiance(skt_env* env, float* out) {
    if (!env || !out) return SKT_ERR_ARG;
    *out = irradianceOf(env->s.params);
    return SKT_OK;
}
int skt_env_ground_coverage_factor(skt_env* env) {
    if (!env) return SKT_ERR_ARG;
    HIPCHK(env->s, launch_ground_columns(env->s, 1, 0.0f, 0.0f, 0.0f));
    return SKT_OK;
}
int skt_env_update_ground_temps(skt_env* env, float dt, float speed, float irr) {
    if (!env) return SKT_ERR_ARG;
    HIPCHK(env->s, launch_ground_columns(env->s, 2, dt, speed, irr));
    return SKT_OK;
}
int skt_env_calculate_buoyancy(skt_env* env, float dt) {
    if (!env) return SKT_ERR_ARG;
    HIPCHK(env->s, launch_buoyancy(env->s, dt));
    return SKT_OK;
}
int skt_env_init_density(skt_env* env) {
    if (!env) return SKT_ERR_ARG;
    HIPCHK(env->s, launch_init_density(env->s, 1));
    return SKT_OK;
}
int skt_env_calculate_new_pressure(skt_env* env) {
    if (!env) return SKT_ERR_ARG;
    HIPCHK(env->s, launch_new_pressure(env->s));
    return SKT_OK;
}

int skt_env_get_param_array(skt_env* env, int32_t parameter, int32_t direction, float** dev_ptr) {
    if (!env || !dev_ptr) return SKT_ERR_ARG;
    EnvState& e = env->s;
    *dev_ptr = nullptr;
    switch (parameter) {
        case SKT_PARAM_POTTEMP: *dev_ptr = e.f[SKT_POTTEMP]; break;
        case SKT_PARAM_QV: *dev_ptr = e.f[SKT_QV]; break;
        case SKT_PARAM_QW: *dev_ptr = e.f[SKT_QW]; break;
        case SKT_PARAM_QC: *dev_ptr = e.f[SKT_QC]; break;
        case SKT_PARAM_QR: *dev_ptr = e.f[SKT_QR]; break;
        case SKT_PARAM_QS: *dev_ptr = e.f[SKT_QS]; break;
        case SKT_PARAM_QI: *dev_ptr = e.f[SKT_QI]; break;
        case SKT_PARAM_WIND:
            if (direction == SKT_DIR_LEFT || direction == SKT_DIR_RIGHT) *dev_ptr = e.f[SKT_VELX];
            else if (direction == SKT_DIR_UP || direction == SKT_DIR_DOWN) *dev_ptr = e.f[SKT_VELY];
            else if (direction == SKT_DIR_FORWARD || direction == SKT_DIR_BACKWARD) *dev_ptr = e.f[SKT_VELZ];
            break;
        case SKT_PARAM_PRESSURE: *dev_ptr = e.f[SKT_PRESSURE]; break;
        default: break;
    }
    // PGROUND and unknown ids: the reference prints an error and returns nullptr
    return *dev_ptr ? SKT_OK : setErr(e, SKT_ERR_ARG, "getParamArray: parameter not supported");
}

int skt_env_get_field_ptr(skt_env* env, int32_t field, void** dev_ptr, int64_t* count) {
    if (!env || !validField(field) || !dev_ptr) return SKT_ERR_ARG;
    *dev_ptr = env->s.f[field];
    if (count) *count = fieldLen(env->s, field);
    return SKT_OK;
}
// A 16-bit field crosses as its words and is converted on the host with the
// device's conversions (fromH16 / toH16).
int skt_env_copy_field_to_host(skt_env* env, int32_t field, float* host) {
    if (!env || !validField(field) || !host) return SKT_ERR_ARG;
    EnvState& e = env->s;
    const long n = fieldLen(e, field);
    if (is16(e, field)) {
        std::vector<unsigned short> h((size_t)n);
        HIPCHK(e, hipMemcpyAsync(h.data(), e.f[field], fieldBytes(e, field), hipMemcpyDeviceToHost, e.stream));
        HIPCHK(e, hipStreamSynchronize(e.stream));
        for (long i = 0; i < n; i++) host[i] = fromH16(h[i], e.fst[field]);
        return SKT_OK;
    }
    HIPCHK(e, hipMemcpyAsync(host, e.f[field], sizeof(float) * n, hipMemcpyDeviceToHost, e.stream));
    HIPCHK(e, hipStreamSynchronize(e.stream));
    return SKT_OK;
}
int skt_env_copy_field_from_host(skt_env* env, int32_t field, const float* host) {
    if (!env || !validField(field) || !host) return SKT_ERR_ARG;
    EnvState& e = env->s;
    const long n = fieldLen(e, field);
    if (is16(e, field)) {
        std::vector<unsigned short> h((size_t)n);
        for (long i = 0; i < n; i++) h[i] = toH16(host[i], e.fst[field]);
        HIPCHK(e, hipMemcpyAsync(e.f[field], h.data(), fieldBytes(e, field), hipMemcpyHostToDevice, e.stream));
        HIPCHK(e, hipStreamSynchronize(e.stream));
        return SKT_OK;
    }
    HIPCHK(e, hipMemcpyAsync(e.f[field], host, sizeof(float) * n, hipMemcpyHostToDevice, e.stream));
    HIPCHK(e, hipStreamSynchronize(e.stream));
    return SKT_OK;
}
int skt_env_get_field_format(skt_env* env, int32_t field, int32_t* format, float* scale) {
    if (!env || !validField(field)) return SKT_ERR_ARG;
    const Store& st = env->s.fst[field];
    if (format) *format = st.fmt == FMT_F32 ? SKT_FMT_F32 : (st.fmt == FMT_BF16 ? SKT_FMT_BF16 : SKT_FMT_F16);
    if (scale) *scale = st.fmt == FMT_F32 ? 1.0f : st.scale;
    return SKT_OK;
}
int skt_env_set_params(skt_env* env, const skt_env_params* p) {
    if (!env || !p) return SKT_ERR_ARG;
    env->s.params = *p;
    return SKT_OK;
}
int skt_env_get_params(skt_env* env, skt_env_params* p) {
    if (!env || !p) return SKT_ERR_ARG;
    *p = env->s.params;
    return SKT_OK;
}
static void setRefineStats(EnvState& e, const PcgState& last) {
    if (e.cfg.pcg_refine > 0) {
        e.stats.pcg_passes = e.ref_snap->passes;
        e.stats.pcg_iterations_total = e.ref_snap->iters_total;
        e.stats.pcg_residual64 = e.ref_snap->res64;
        e.stats.pcg_rhs64 = e.ref_snap->rhs64;
    } else {
        e.stats.pcg_passes = 1;
        e.stats.pcg_iterations_total = last.iters;
        e.stats.pcg_residual64 = 0.0;
        e.stats.pcg_rhs64 = 0.0;
    }
}
int skt_env_get_stats(skt_env* env, skt_step_stats* st) {
    if (!env || !st) return SKT_ERR_ARG;
    EnvState& e = env->s;
    if (e.stats_on_device) {
        // after an ensemble step: the final solver state (slot 0 after the
        // last iteration's stop test) and the last advection's substep count
        PcgState ps{};
        MaxvelSnap mv{};
        HIPCHK(e, hipMemcpyAsync(&ps, &e.sc->st[0], sizeof(PcgState), hipMemcpyDeviceToHost, e.stream));
        HIPCHK(e, hipMemcpyAsync(&mv, &e.sc->maxvel, sizeof(MaxvelSnap), hipMemcpyDeviceToHost, e.stream));
        if (e.cfg.pcg_refine > 0)
            HIPCHK(e, hipMemcpyAsync(e.ref_snap, &e.sc->ref, sizeof(RefineStats), hipMemcpyDeviceToHost, e.stream));
        HIPCHK(e, hipStreamSynchronize(e.stream));
        e.stats.pcg_iterations = ps.iters;
        e.stats.pcg_residual = ps.maxr;
        e.stats.pcg_residual0 = ps.maxr0;
        setRefineStats(e, ps);
        for (int k = 0; k < std::min(e.stats.n_advect, 16); k++) e.stats.substeps[k] = mv.nsub;
        e.pcg_stats_pending = false;
        e.stats_on_device = false;
    }
    if (e.pcg_stats_pending) {
        HIPCHK(e, hipEventSynchronize(e.snap_ev[2]));
        e.stats.pcg_iterations = e.pcg_snap[2].iters;
        e.stats.pcg_residual = e.pcg_snap[2].maxr;
        e.stats.pcg_residual0 = e.pcg_snap[2].maxr0;
        setRefineStats(e, e.pcg_snap[2]);
        e.pcg_stats_pending = false;
    }
    *st = e.stats;
    return SKT_OK;
}
// environmentGPU::prepareBrushGPU / prepareSelectionGPU
// (Game/environment.cu:1127-1260): the edit runs on a copy of the ground
// heights; when it touched them, an erase first resets the cells it uncovered
// (compareAndResetValuesOutGround), then the copy becomes the ground (with its
// derived maps) and the ground-change path of editorDataGPU runs.
static int applyEdit(skt_env* env, EditArgs& a, bool brush) {
    EnvState& e = env->s;
    a.a[0] = a.a[1] = a.a[2] = nullptr;
    switch (a.param) {
        case SKT_PARAM_POTTEMP: a.a[0] = e.f[SKT_POTTEMP]; break;
        case SKT_PARAM_QV: a.a[0] = e.f[SKT_QV]; break;
        case SKT_PARAM_QW: a.a[0] = e.f[SKT_QW]; break;
        case SKT_PARAM_QC: a.a[0] = e.f[SKT_QC]; break;
        case SKT_PARAM_QR: a.a[0] = e.f[SKT_QR]; break;
        case SKT_PARAM_QS: a.a[0] = e.f[SKT_QS]; break;
        case SKT_PARAM_QI: a.a[0] = e.f[SKT_QI]; break;
        case SKT_PARAM_WIND:
            a.a[0] = e.f[SKT_VELX];
            a.a[1] = e.f[SKT_VELY];
            a.a[2] = e.f[SKT_VELZ];
            break;
        case SKT_PARAM_PGROUND: break;
        case SKT_PARAM_PRESSURE: a.a[0] = e.f[SKT_PRESSURE]; break;
        default: return setErr(e, SKT_ERR_ARG, "edit: unknown parameter");
    }
    HIPCHK(e, hipSetDevice(e.device));
    HIPCHK(e, hipMemcpyAsync(e.ghEdit, e.gh, sizeof(int) * e.G, hipMemcpyDeviceToDevice, e.stream));
    HIPCHK(e, hipMemsetAsync(&e.sc->edit_ground, 0, sizeof(int), e.stream));
    HIPCHK(e, brush ? launch_apply_brush(e, a) : launch_apply_selection(e, a));
    int rc = readScalars(e);
    if (rc) return rc;
    if (e.sc_host->edit_ground) {
        if (a.erase) HIPCHK(e, launch_reset_uncovered(e));
        std::vector<int32_t> h((size_t)e.G);
        HIPCHK(e, hipMemcpyAsync(h.data(), e.ghEdit, sizeof(int) * e.G, hipMemcpyDeviceToHost, e.stream));
        HIPCHK(e, hipStreamSynchronize(e.stream));
        rc = skt_env_set_ground_height(env, h.data());
        if (rc) return rc;
    }
    rc = editor_data(e);
    if (rc) return rc;
    HIPCHK(e, hipStreamSynchronize(e.stream));
    return SKT_OK;
}
int skt_env_apply_brush(skt_env* env, int32_t param, float brush_size, const int32_t pos[3], float smoothness,
                        float dt, float intensity, float value, const float dir[3], int32_t ground_erase) {
    if (!env || !pos || !dir) return SKT_ERR_ARG;
    EditArgs a{};
    a.param = param;
    a.size = brush_size;
    a.smooth = smoothness;
    a.dt = dt;
    a.intensity = intensity;
    a.value = value;
    a.erase = ground_erase != 0;
    for (int k = 0; k < 3; k++) {
        a.pos[k] = pos[k];
        a.dir[k] = dir[k];
    }
    return applyEdit(env, a, true);
}
int skt_env_apply_selection(skt_env* env, int32_t param, const int32_t min_pos[3], const int32_t max_pos[3],
                            float value, const float dir[3], int32_t ground_erase) {
    if (!env || !min_pos || !max_pos || !dir) return SKT_ERR_ARG;
    EditArgs a{};
    a.param = param;
    a.value = value;
    a.erase = ground_erase != 0;
    for (int k = 0; k < 3; k++) {
        a.mn[k] = min_pos[k];
        a.mx[k] = max_pos[k];
        a.dir[k] = dir[k];
    }
    return applyEdit(env, a, false);
}
// environmentGPU::resetParameterGPU (Game/environment.cu:1274-1295).
int skt_env_reset_parameter(skt_env* env, int32_t param) {
    if (!env) return SKT_ERR_ARG;
    EnvState& e = env->s;
    HIPCHK(e, hipSetDevice(e.device));
    switch (param) {
        case SKT_PARAM_POTTEMP: HIPCHK(e, launch_reset_param(e, SKT_POTTEMP, e.f[SKT_ISEN_TEMP])); break;
        case SKT_PARAM_QV: HIPCHK(e, launch_reset_param(e, SKT_QV, e.f[SKT_ISEN_VAPOR])); break;
        case SKT_PARAM_QW: HIPCHK(e, launch_reset_param(e, SKT_QW, nullptr)); break;
        case SKT_PARAM_QC: HIPCHK(e, launch_reset_param(e, SKT_QC, nullptr)); break;
        case SKT_PARAM_QR: HIPCHK(e, launch_reset_param(e, SKT_QR, nullptr)); break;
        case SKT_PARAM_QS: HIPCHK(e, launch_reset_param(e, SKT_QS, nullptr)); break;
        case SKT_PARAM_QI: HIPCHK(e, launch_reset_param(e, SKT_QI, nullptr)); break;
        case SKT_PARAM_WIND:
            HIPCHK(e, launch_reset_param(e, SKT_VELX, e.f[SKT_DEFAULT_VELX]));
            HIPCHK(e, launch_reset_param(e, SKT_VELY, nullptr));
            HIPCHK(e, launch_reset_param(e, SKT_VELZ, e.f[SKT_DEFAULT_VELZ]));
            break;
        case SKT_PARAM_PRESSURE: HIPCHK(e, launch_reset_param(e, SKT_PRESSURE, e.f[SKT_DEFAULT_PRESSURE])); break;
        case SKT_PARAM_PGROUND: break;  // the reference has no reset for the ground
        default: return setErr(e, SKT_ERR_ARG, "resetParameterGPU: unknown parameter");
    }
    return SKT_OK;
}
int skt_env_diagnostics(skt_env* env, float thr, skt_diagnostics* out) {
    if (!env || !out) return SKT_ERR_ARG;
    EnvState& e = env->s;
    HIPCHK(e, hipSetDevice(e.device));
    HIPCHK(e, launch_diagnostics(e, thr));
    const int rc = readScalars(e);
    if (rc) return rc;
    out->max_updraft = e.sc_host->dg_w;
    out->cloud_top_level = e.sc_host->dg_level;
    out->cloud_top_height = e.sc_host->dg_level >= 0 ? (float)e.sc_host->dg_level * e.dx : 0.0f;
    out->pad = 0;
    out->surface_rain = e.sc_host->dg_rain;
    return SKT_OK;
}
int skt_env_compute_masks(skt_env* env, uint8_t* host) {
    if (!env || !host) return SKT_ERR_ARG;
    EnvState& e = env->s;
    uint8_t* d = nullptr;
    if (dalloc(e, &d, 12 * (size_t)e.N)) return SKT_ERR_NOMEM;
    hipError_t r = launch_masks(e, d);
    if (r == hipSuccess) r = hipMemcpyAsync(host, d, 12 * (size_t)e.N, hipMemcpyDeviceToHost, e.stream);
    if (r == hipSuccess) r = hipStreamSynchronize(e.stream);
    (void)hipFree(d);
    if (r != hipSuccess) return setErr(e, SKT_ERR_HIP, std::string("masks: ") + hipGetErrorString(r));
    return SKT_OK;
}
int skt_env_phase_mask(skt_env* env, uint8_t* host) {
    if (!env || !host) return SKT_ERR_ARG;
    EnvState& e = env->s;
    HIPCHK(e, hipMemcpyAsync(host, e.phase, (size_t)e.N, hipMemcpyDeviceToHost, e.stream));
    HIPCHK(e, hipStreamSynchronize(e.stream));
    return SKT_OK;
}
int skt_env_set_timing(skt_env* env, int32_t class_mask) {
    if (!env) return SKT_ERR_ARG;
    EnvState& e = env->s;
    RC(resolveTiming(e));
    e.timing = class_mask & ((1 << SKT_TIMING_CLASSES) - 1);
    for (int k = 0; k < SKT_TIMING_CLASSES; k++) {
        e.t_ms[k] = 0.0;
        e.t_count[k] = 0;
        e.t_seen[k] = 0;
    }
    return SKT_OK;
}
int skt_env_set_timing_stride(skt_env* env, int32_t stride) {
    if (!env || stride < 1) return SKT_ERR_ARG;
    env->s.t_stride = stride;
    return SKT_OK;
}
int skt_env_get_timing(skt_env* env, skt_timing* out) {
    if (!env || !out) return SKT_ERR_ARG;
    EnvState& e = env->s;
    RC(resolveTiming(e));
    for (int k = 0; k < SKT_TIMING_CLASSES; k++) {
        out->launches[k] = e.t_count[k];
        out->total_ms[k] = e.t_ms[k];
    }
    return SKT_OK;
}
int skt_env_copy_probe(skt_env* env, int32_t repeats, float* ms_per_copy, double* bytes_per_copy) {
    if (!env || repeats < 1 || !ms_per_copy) return SKT_ERR_ARG;
    EnvState& e = env->s;
    hipEvent_t a = evGet(e), b = evGet(e);
    if (!a || !b) return setErr(e, SKT_ERR_HIP, "hipEventCreate failed");
    HIPCHK(e, launch_stream_copy(e, e.work[0], e.work[1], e.N));  // warm
    HIPCHK(e, hipEventRecord(a, e.stream));
    for (int r = 0; r < repeats; r++) HIPCHK(e, launch_stream_copy(e, e.work[(r & 1)], e.work[1 - (r & 1)], e.N));
    HIPCHK(e, hipEventRecord(b, e.stream));
    HIPCHK(e, hipEventSynchronize(b));
    float ms = 0.0f;
    HIPCHK(e, hipEventElapsedTime(&ms, a, b));
    e.evpool.push_back(a);
    e.evpool.push_back(b);
    *ms_per_copy = ms / repeats;
    if (bytes_per_copy) *bytes_per_copy = 8.0 * (double)(e.N / 4 * 4);
    return SKT_OK;
}
int skt_terrain_generate(int32_t device, int32_t seed, int32_t nx, int32_t nz, int32_t octaves, float max_height,
                         float* noise_out, int32_t* heights_out) {
    if (nx < 1 || nz < 1 || octaves < 0 || (!noise_out && !heights_out)) return SKT_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return SKT_ERR_HIP;
    const size_t n = (size_t)nx * nz;
    float* dn = nullptr;
    int* dh = nullptr;
    if (hipMalloc((void**)&dn, n * sizeof(float)) != hipSuccess) return SKT_ERR_NOMEM;
    if (hipMalloc((void**)&dh, n * sizeof(int)) != hipSuccess) {
        (void)hipFree(dn);
        return SKT_ERR_NOMEM;
    }
    hipError_t r = launch_terrain(seed, nx, nz, octaves, max_height, dn, dh, nullptr);
    if (r == hipSuccess && noise_out) r = hipMemcpy(noise_out, dn, n * sizeof(float), hipMemcpyDeviceToHost);
    if (r == hipSuccess && heights_out) r = hipMemcpy(heights_out, dh, n * sizeof(int), hipMemcpyDeviceToHost);
    (void)hipFree(dn);
    (void)hipFree(dh);
    return r == hipSuccess ? SKT_OK : SKT_ERR_HIP;
}

int skt_terrain_generate_noalloc(int32_t device, int32_t seed, int32_t nx, int32_t nz, int32_t octaves,
                                 float max_height, float* noise_out, int32_t* heights_out) {
    if (nx < 1 || nz < 1 || octaves < 0 || (!noise_out && !heights_out)) return SKT_ERR_ARG;
    if ((long)nx * nz > SKT_TERRAIN_NOALLOC_MAX_COLUMNS) return SKT_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return SKT_ERR_HIP;
    float* dn = nullptr;
    int* dh = nullptr;
    if (terrain_scratch(device, &dn, &dh) != hipSuccess) return SKT_ERR_HIP;
    const size_t n = (size_t)nx * nz;
    hipError_t r = launch_terrain(seed, nx, nz, octaves, max_height, dn, dh, nullptr);
    if (r == hipSuccess && noise_out) r = hipMemcpy(noise_out, dn, n * sizeof(float), hipMemcpyDeviceToHost);
    if (r == hipSuccess && heights_out) r = hipMemcpy(heights_out, dh, n * sizeof(int), hipMemcpyDeviceToHost);
    return r == hipSuccess ? SKT_OK : SKT_ERR_HIP;
}

int skt_device_synchronize(int32_t device) {
    if (hipSetDevice(device) != hipSuccess) return SKT_ERR_HIP;
    return hipDeviceSynchronize() == hipSuccess ? SKT_OK : SKT_ERR_HIP;
}

int skt_mphys_region_set(skt_env* env, const int32_t mn[3], const int32_t mx[3]) {
    if (!env || !mn || !mx) return SKT_ERR_ARG;
    EnvState& e = env->s;
    for (int k = 0; k < 3; k++) {
        e.rmin[k] = mn[k];
        e.rmax[k] = mx[k];
    }
    e.region_active = 1;
    std::memset(e.region_host, 0, sizeof(e.region_host));
    return SKT_OK;
}
int skt_mphys_region_clear(skt_env* env) {
    if (!env) return SKT_ERR_ARG;
    env->s.region_active = 0;
    return SKT_OK;
}
int skt_mphys_region_get(skt_env* env, float out[33]) {
    if (!env || !out) return SKT_ERR_ARG;
    std::memcpy(out, env->s.region_host, sizeof(env->s.region_host));
    return SKT_OK;
}

}  // extern "C"
