// Internal state of one environment and the launch interface between the
// host orchestration (skt_api.cpp) and the kernels (skt_kernels.hip).
#pragma once
#include <cstddef>
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "../../include/skt.h"
#include "skt_device.h"

namespace skt {

// State of the pressure solve's scalar recurrence (alpha, beta, sigma and
// the stop test). Two slots: a kernel reads one and its block 0 writes the
// updated copy into the other, so every block of one launch sees the same input.
struct PcgState {
    double sigma0, sigma1;  // sigma values (stored rounded to float, kept in double slots)
    float alpha, beta;
    float maxr, maxr0;      // stop-test values
    float tol;              // stop tolerance in effect
    int done;               // solve finished
    int iters;              // iterations performed
    int pad[5];
};

// One level of the multigrid preconditioner: dims, the operator as
// {+x face, +y face, +z face, diagonal} per cell (faces as positive
// couplings, diagonal 0 marks a ground cell; level l holds its operator
// divided by 2^l, so every level's values stay in the 16-bit range), the
// right-hand side r (likewise divided by 2^l) and the two iterate buffers
// (the level's result ends in xb). Level 0's r is the PCG residual and its
// xb the preconditioned residual z.
constexpr int kMgLevels = 8;
constexpr long kMgCoarsest = 512;  // cells of a coarsest level at most
// A cell's operator row on a multigrid level in 16-bit floats: the +x, +y,
// +z couplings and the diagonal.
struct alignas(8) MgC {
    _Float16 x, y, z, w;
};
struct MgLevel {
    int X = 0, Y = 0, Z = 0;
    long N = 0;
    MgC* C = nullptr;
    float *r = nullptr, *xa = nullptr, *xb = nullptr;
};

// Totals of an FP64-refined pressure solve (written on the device).
struct RefineStats {
    int iters_total;  // iterations over the projection's solves
    int passes;       // solves that ran iterations
    double res64;     // max |b - A x| in FP64 for the final x
    double rhs64;     // max |b|
    float target;     // pcg_tol * max |b|: the stop value shared by the refinement passes
    int pad;
};

// Device scalars shared by the reduction kernels.
struct Scalars {
    PcgState st[2];
    float maxvel;  // advection max face velocity
    int nsub;      // advection substeps
    float dtsub;   // advection substep length
    float dg_w;    // serving diagnostics: max vertical velocity over sky cells
    int dg_level;  //   highest cloudy level (-1: none)
    int edit_ground;  // an edit touched the ground heights
    double dg_rain;  //   ground rain water summed over the columns
    int face_exact;  // a sky 1/density of this solve lies outside the short face form's range
    int sub_overflow;  // an ensemble advection needed more substeps than its recorded passes
    int pad[2];
    RefineStats ref;
};

// Host copy of the advection substep fields of Scalars (maxvel, nsub, dtsub).
struct MaxvelSnap {
    float maxvel;
    int nsub;
    float dtsub;
};
static_assert(offsetof(Scalars, nsub) == offsetof(Scalars, maxvel) + offsetof(MaxvelSnap, nsub) &&
                  offsetof(Scalars, dtsub) == offsetof(Scalars, maxvel) + offsetof(MaxvelSnap, dtsub),
              "MaxvelSnap mirrors the substep fields of Scalars");

// Regions of the per-block reduction partials (each maxBlocks doubles).
enum { PART_DIV = 0, PART_START, PART_A, PART_BMAX, PART_BSUM, PART_DG_W, PART_DG_L, PART_DG_R, PART_COUNT };

// N-sized work buffers: three per field of a batched advection.
constexpr int SKT_NWORK = 16;

// z-marching launch shape: 64x4-cell tiles, each block walks zc planes; about
// 2048 blocks for the large domains (one reduction partial per block). The
// grid is one-dimensional; kernels map the block index to a tile.
struct ZGrid {
    dim3 grid;
    int zc;
    unsigned nblocks;
};
// Flat launch shape of the elementwise solver kernels: nblocks blocks of rpb
// whole rows (x fastest) each.
struct FGrid {
    unsigned nblocks, rpb;
};
// Tile index of a block in its launch, and the launch's tile count.
struct VB {
    unsigned id, n;
};
// Work list of an ensemble launch: member m owns the work indices
// [start[m], start[m] + nb[m]) (start[M] = total).
struct WorkMap {
    const unsigned* __restrict__ start;
    const unsigned* __restrict__ nb;
    int M;
    unsigned total;
};
// Every kernel is compiled with __launch_bounds__(256).
constexpr unsigned kMaxBlockThreads = 256;
inline bool blockFits(unsigned bx, unsigned by, unsigned bz) {
    return bx > 0 && by > 0 && bz > 0 && (unsigned long long)bx * by * bz <= kMaxBlockThreads;
}
// One recorded launch of an ensemble step: the ensemble launcher of its body,
// its tile count and block shape in this member's own launch, and the bytes
// of its argument record.
struct RecLaunch {
    hipError_t (*fn)(const void* dargs, const WorkMap& wm, unsigned grid, dim3 block, uint32_t shm,
                     hipStream_t s) = nullptr;
    int (*occ)(int threads, uint32_t shm) = nullptr;
    const char* name = nullptr;  // the body's name (measurement reports)
    unsigned nb = 0;
    dim3 block;
    uint32_t shm = 0;
    std::vector<unsigned char> args;
};
// Recording of one member's step. While an env holds a recorder, launches go
// to it instead of the device, and the step takes the values it would read
// back from the device (advection substeps, solver progress) on the device:
// each advection runs adv_cap substep passes that the kernels skip past the
// member's own count, each sedimentation precip_cap launches, and the
// pressure solve all its iterations (a finished solve's kernels return at
// once).
struct Recorder {
    std::vector<RecLaunch> launches;
    int adv_cap = 1;
    int precip_cap = 1;
    const float* irr = nullptr;  // device slot of this member's irradiance
};

struct EnvState {
    skt_config cfg;
    int X, Y, Z;
    long N, G;
    float dx;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int device = 0;
    std::string err;
    bool initialized = false;

    float* f[SKT_FIELD_COUNT] = {nullptr};  // field arrays (sizes N, G or Y)
    Store fst[SKT_FIELD_COUNT];             // their storage (16-bit for the tracers of the mixed form)
    bool mix = false;                       // some field is stored in 16 bits
    int* gh = nullptr;                      // ground height [G]
    int* ghm1 = nullptr;                    // max ground height over the 3x3 column neighbourhood [G]
    int* ghm2 = nullptr;                    // max ground height over the 5x5 column neighbourhood [G]
    int* ghz = nullptr;                     // max ground height over planes 0..z of each x column [G]
    int* ghEdit = nullptr;                  // ground heights being edited by a brush or selection [G]
    float* work[SKT_NWORK] = {nullptr};     // N-sized work buffers
    float* pcg[6] = {nullptr};              // r, z, s_a, s_b, As, pre
    float* irho = nullptr;                  // 1/density of the last pressure solve (0 in ground)
    float* rhs = nullptr;                   // FP64 refinement: the solve's right-hand side b [N]
    double* x64 = nullptr;                  //   and the FP64 sum of its solves [N]
    // multigrid preconditioner: levels 0..mg_depth-1 in use (0: Jacobi); the
    // launch sequence always covers kMgLevels levels (a level past a grid's
    // depth has no tiles), so grids of any size share it in an ensemble
    int mg_depth = 0;
    MgLevel mg[kMgLevels];
    unsigned mg_dot_n = 0;                  // z.r partials the last V-cycle wrote
    float* zeros = nullptr;                 // [Y] zero profile
    float* ones = nullptr;                  // [Y] unit profile
    double* partials = nullptr;             // per-block reduction partials (PART_COUNT x maxBlocks)
    int maxBlocks = 0;
    ZGrid gZ{};                             // z-marching grid of the stencil kernels
    ZGrid gA{};                             // grid of the pressure operator kernel
    ZGrid gD{};                             // grid of a diffusion sweep (per field)
    FGrid gB{};                             // grid of the elementwise solver kernels
    Scalars* sc = nullptr;                  // device scalars
    Scalars* sc_host = nullptr;             // pinned mirror for readbacks
    uint8_t* phase = nullptr;               // [N] microphysics phase mask
    double* region = nullptr;               // region diagnostic partials
    int region_active = 0;
    int rmin[3] = {0, 0, 0}, rmax[3] = {0, 0, 0};
    float region_host[33] = {0};

    Gammas gm;
    skt_env_params params;
    skt_step_stats stats;

    // per-class kernel timing (skt_env_set_timing)
    struct TimedLaunch {
        int cls;
        hipEvent_t a, b;
    };
    int timing = 0;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> evpool;
    double t_ms[SKT_TIMING_CLASSES] = {0};
    long long t_count[SKT_TIMING_CLASSES] = {0};
    long long t_seen[SKT_TIMING_CLASSES] = {0};  // launches of each class since skt_env_set_timing
    int t_stride = 1;           // time every t_stride-th launch of a class
    // pressure-solve state snapshots in pinned host memory: two alternating
    // chunk snapshots and the final state, each behind its event
    PcgState* pcg_snap = nullptr;
    RefineStats* ref_snap = nullptr;  // refinement totals, copied with the final state
    hipEvent_t snap_ev[3] = {nullptr, nullptr, nullptr};
    bool pcg_stats_pending = false;
    // substep count launched ahead of an advection (startMaxvel)
    MaxvelSnap* mv_snap = nullptr;
    hipEvent_t mv_ev = nullptr;
    bool mv_ready = false;
    float mv_dt = 0.0f;
    // events armed for the next klaunch (a timed single-kernel launch)
    hipEvent_t kevA = nullptr, kevB = nullptr;
    // set while an ensemble step of this member is being recorded
    Recorder* rec = nullptr;
    // the last step ran in an ensemble: its solver state is read from the device
    bool stats_on_device = false;
    // device bytes requested by the env's allocations; dry: count them
    // without allocating (footprint sizing)
    long long dev_bytes = 0;
    bool dry = false;
};

}  // namespace skt

// The ABI's environment handle.
struct skt_env {
    skt::EnvState s;
};

namespace skt {

enum { TC_DIFFUSE = 0, TC_PPM, TC_PCG_A, TC_PCG_B, TC_MICRO, TC_MAXVEL, TC_PRECIP, TC_OTHER, TC_MG };
// Classes whose launch functions issue one kernel through klaunch, so their
// timing events ride on the kernel's own dispatch.
constexpr bool kTimedOnDispatch[SKT_TIMING_CLASSES] = {true, true, true, true, true, false, true, false, false};


// --------------------------------------------------- step (skt_api.cpp) ---
// environmentGPU::updateGPU on one env, or its recording when e.rec is set.
int update(EnvState& e, float dt, float speed);
// The ground-change path of editorDataGPU (Game/environment.cu:989-1019).
int editor_data(EnvState& e);
// Host bookkeeping after a step: the time of day and the ground flag.
void afterStep(EnvState& e, float dt, float speed);
// irridianceGPU (Game/environment.cu:1021-1034).
float irradianceOf(const skt_env_params& p);

// ------------------------------------------------------------- launches ---
Geom geom(const EnvState& e);
ZGrid zgrid(const EnvState& e);
// Blocks the device holds at once of the pressure operator kernel and of a
// diffusion sweep (0 if unknown).
struct ResidentBlocks {
    long pcgA, diffuse;  // blocks the device holds at once (0: unknown)
    int cus;             // compute units (0: unknown)
};
ResidentBlocks resident_blocks(int device);
// Fills gZ, gA and gD (at most one round of resident blocks each) and gB.
void setup_grids(EnvState& e, ResidentBlocks rb);
hipError_t launch_fill(EnvState& e, float* a, float v, long n);
hipError_t launch_copy_sky(EnvState& e, const float* src, float* dst, Store ss, Store ds);
hipError_t launch_ground_water_snapshot(EnvState& e, float dt, float speed);
// stages: 1 cloud cover, 2 ground temp, 4 precip hitting ground, 8 ground microphysics, 16 setTempsAtGround
hipError_t launch_ground_columns(EnvState& e, int stages, float dt, float speed, float irr);
// Independent fields swept by one diffusion launch.
constexpr int SKT_BATCH = 6;
struct DiffField {
    const float* in;
    float* out;
    const float* prof;
    Bounds b;
    int type;
    Store sin, sout;  // storage of in and out
};
struct DiffBatch {
    DiffField f[SKT_BATCH];
    int n;
};
hipError_t launch_diffuse_batch(EnvState& e, const DiffBatch& db, float k);
hipError_t launch_maxvel(EnvState& e, float dt);
hipError_t launch_diagnostics(EnvState& e, float thr);
// Brush / selection edits (environmentGPU::prepareBrushGPU / prepareSelectionGPU).
struct EditArgs {
    int param;               // parameter id (POTTEMP 0 .. QI 6, WIND 7, PGROUND 8, PRESSURE 9)
    float* a[3];             // the parameter's arrays (wind: x, y, z components)
    Store st[3];             // their storage
    float value, dir[3];     // applyValue, valueDir
    int erase;               // groundErase
    // brush
    float size, smooth, intensity, dt;
    int pos[3];
    // selection
    int mn[3], mx[3];
};
hipError_t launch_apply_brush(EnvState& e, const EditArgs& a);
hipError_t launch_apply_selection(EnvState& e, const EditArgs& a);
hipError_t launch_reset_uncovered(EnvState& e);
hipError_t launch_terrain(int seed, int width, int depth, int octaves, float maxHeight, float* noise, int* heights,
                          hipStream_t stream);
// The device scratch of the allocation-free terrain entry point on 'device'
// (which must be current): room for SKT_TERRAIN_NOALLOC_MAX_COLUMNS columns of
// noise and of heights. The address is looked up once per device.
hipError_t terrain_scratch(int device, float** noise, int** heights);
// Independent fields moved by one PPM pass launch.
struct PpmField {
    const float* in;
    float* out;
    const float* vel;
    const float* prof;
    Bounds b;
    const float* gsrc;  // values at ground positions (nullptr: 0), used under a NEUMANN ground condition
    Store sin, sout, sg;  // storage of in, out and gsrc
};
struct PpmBatch {
    PpmField f[SKT_BATCH];
    int n;
    // set in a recorded ensemble step: the pass takes its time step from the
    // member's device scalars (substep sub, half = 1 for the x and y passes)
    // and is skipped when sub is not below the member's substep count
    const Scalars* dsc;
    int sub, half;
};
hipError_t launch_ppm_batch(EnvState& e, const PpmBatch& pb, int dir, float dt);
// Precipitation fields falling in one launch (dts 0: field skipped).
struct PrecipBatch {
    float* q[3];
    int type[3];
    float dts[3];
    int n;
    Store st[3];
};
hipError_t launch_precip_batch(EnvState& e, const PrecipBatch& pbt);
hipError_t launch_buoyancy(EnvState& e, float dt);
hipError_t launch_init_density(EnvState& e, int copy_old);
hipError_t launch_new_pressure(EnvState& e);
hipError_t launch_reset_vel(EnvState& e);
// setToDefault (prof) or setToValue 0 (prof null) of one field
hipError_t launch_reset_param(EnvState& e, int field, const float* prof);
hipError_t launch_divergence(EnvState& e, float dt);  // r = div, partial max |r|; irho and the preconditioner
hipError_t launch_pcg_start(EnvState& e, int pass);    // stop tolerance; z = pre r, partial z.r (pass: refinement
                                                       // solve index, -1 without refinement)
hipError_t launch_pcg_A(EnvState& e, int it);          // stop test / beta; s update, As, partial s.As
hipError_t launch_pcg_B(EnvState& e, int it);          // alpha; p', r, z updates, partial max |r| and z.r
hipError_t launch_pcg_final(EnvState& e, int it);      // stop test after the last iteration
hipError_t launch_refine_keep_rhs(EnvState& e);        // rhs = r (the divergence) before the first solve
hipError_t launch_refine_accum(EnvState& e, int first, int last);  // x64 += p'
hipError_t launch_refine_residual(EnvState& e, int write_r);      // r = b - A x64 in FP64 (+ max partials)
hipError_t launch_refine_report(EnvState& e);          // final max |b - A x64|
hipError_t launch_mg_setup(EnvState& e);               // the levels' operators from 1/density
// z = V-cycle(r) and the z.r partials for the next search direction (into
// PART_START when start, else PART_BSUM; e.mg_dot_n of them)
hipError_t launch_mg_vcycle(EnvState& e, int start);
hipError_t launch_update_pressure_density(EnvState& e);
hipError_t launch_apply_projection(EnvState& e, int new_pressure);
hipError_t launch_microphysics(EnvState& e, float dt, float speed);
hipError_t launch_region_finalize(EnvState& e);
hipError_t launch_masks(EnvState& e, uint8_t* dev_out);
hipError_t launch_editor_reset(EnvState& e, int nfields);
hipError_t launch_isen_temp_ground(EnvState& e);
hipError_t launch_profiles_from_column(EnvState& e);
hipError_t launch_sounding_fill(EnvState& e, const float* levels, float groundT, float groundP);
hipError_t launch_random(EnvState& e, float* a, float mn, float mx, uint32_t seed);
hipError_t launch_stream_copy(EnvState& e, const float* src, float* dst, long n);
int blocks3d(const EnvState& e);

}  // namespace skt
