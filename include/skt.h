/*
 * skt.h — C ABI of the MI355X-native skewT cloud-simulation step.
 *
 * This is the drop-in boundary for the reference's environmentGPU class
 * (Game/environment.cuh:9-183) and its dataClass companion
 * (Game/dataClass.cuh:67-89). Every entry point returns an skt_status (0 = ok)
 * instead of breaking into the debugger; skt_env_last_error() gives the text.
 * Device arrays are owned by the environment; pointers handed out are
 * borrowed and stay valid (same address) until skt_env_destroy().
 *
 * Grid convention (Game/utils.cuh:8-11): index = x + nx*(y + ny*z), y is the
 * vertical axis; ground columns are indexed x + nx*z.
 * Work is stream-ordered on one HIP stream per environment; calls on one
 * environment are not thread-safe, separate environments are independent.
 */
#ifndef SKT_H
#define SKT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct skt_env skt_env;

typedef enum skt_status {
    SKT_OK = 0,
    SKT_ERR_ARG = 1,     /* invalid argument (shape, field id, null pointer) */
    SKT_ERR_HIP = 2,     /* a HIP runtime call failed */
    SKT_ERR_STATE = 3,   /* call order (e.g. update before init) */
    SKT_ERR_NOMEM = 4,   /* device allocation failed */
    SKT_ERR_INPUT = 5    /* sounding input missing or malformed (skt_input_last_error) */
} skt_status;

/* Boundary conditions, Game/config.h:43-46. */
enum { SKT_NEUMANN = 0, SKT_DIRICHLET = 1, SKT_CUSTOM = 2 };
/* boundsEnv, Game/config.h:52-59 (the 'up' member is carried but, as in the
 * reference, the top face uses 'sides'). */
typedef struct skt_bounds { int32_t sides, up, ground; } skt_bounds;

/* Pressure-solve modes. REFERENCE reproduces calculatePressureProject
 * (Game/environment.cu:898-987) including its sigma accumulation and stop
 * test; STANDARD is textbook Jacobi-PCG with a relative tolerance. */
enum { SKT_PCG_REFERENCE = 0, SKT_PCG_STANDARD = 1 };

typedef struct skt_config {
    int32_t nx, ny, nz;        /* ny vertical; replaces GRIDSIZESKY{X,Y,Z} (Game/config.h:3-5) */
    float dx;                  /* voxel size in m; replaces VOXELSIZE (Game/config.h:9) */
    int32_t device;            /* HIP device ordinal */
    int32_t pcg_mode;          /* SKT_PCG_REFERENCE or SKT_PCG_STANDARD */
    float pcg_tol;             /* 1e-5 absolute (reference) or relative to max|r0| */
    int32_t pcg_max_iter;      /* 200 (Game/environment.cu:907) */
    int32_t diffusion_sweeps;  /* 20 (Game/environment.cu:649) */
    int32_t max_substeps;      /* 100 (Game/environment.cu:727) */
    int32_t maxabs_mode;       /* 0: getMaxDivergence as written (|.| at z=0 only); 1: max|.| */
    uint32_t noise_seed;       /* first of the five init-noise seeds (100 in Game/environment.cu:308) */
    int32_t init_noise;        /* 1: add the reference's init perturbations */
    int32_t max_blocks;        /* upper bound on the blocks of one solver launch; 0 sizes launches from the
                                  device. A small bound gives each block more z planes (or rows). */
    /* Mixed-precision form. The five hydrometeor fields (QW, QC, QR, QS, QI;
     * Game/environment.cuh:13-25 holds them as float) may be stored in 16
     * bits: value * 2^tracer_scale_log2[k], rounded to nearest even. Every
     * kernel computes in FP32 and rounds once when it stores. */
    int32_t tracer_format;         /* SKT_FMT_F32 (default), SKT_FMT_BF16 or SKT_FMT_F16 */
    int32_t tracer_scale_log2[5];  /* per field, QW..QI; 16 by default (1e-6..2e-2 kg/kg lands in
                                      [2^-4, 2^11], inside FP16's normal range; BF16 needs no scale) */
    /* FP64 residual refinement of the pressure solve (standard solver only):
     * after the first solve, each pass forms r = b - A x in FP64 from the
     * FP64 sum x of the solves so far and solves A d = r in FP32. Every solve
     * stops at max|r| <= max(pcg_refine_tol * its own max|r_start|,
     * pcg_tol * max|b|); a pass whose residual already meets the second bound
     * runs no iteration. */
    int32_t pcg_refine;            /* refinement passes after the first solve (0: none) */
    float pcg_refine_tol;          /* relative tolerance of each solve when refining (1e-3) */
    /* Preconditioner of the standard solver: SKT_PRECON_JACOBI (the
     * reference's 1/diagonal, Game/kernelSky.cu:659-673) or SKT_PRECON_MG, a
     * geometric multigrid V-cycle on the same operator (2x2x2 coarsening,
     * coarse faces the halved sums of the fine faces they cover, two damped
     * Jacobi sweeps before and after each coarse correction, 64 sweeps on the
     * coarsest level of at most 512 cells). The V-cycle is a fixed symmetric
     * positive definite linear map, so the iteration is still conjugate
     * gradients; it stops on the same test. No counterpart in the reference. */
    int32_t pcg_precon;
} skt_config;

enum { SKT_PRECON_JACOBI = 0, SKT_PRECON_MG = 1 };
/* Storage formats of skt_config.tracer_format and skt_env_get_field_format. */
enum { SKT_FMT_F32 = 0, SKT_FMT_BF16 = 1, SKT_FMT_F16 = 2 };

/* Fills a configuration with the reference's settings for an nx*ny*nz grid. */
void skt_config_default(skt_config* cfg, int32_t nx, int32_t ny, int32_t nz, float dx);

/* Field ids: the gridDataSkyGPU / gridDataGroundGPU members
 * (Game/environment.cuh:13-37), the solver work arrays and the per-level
 * profiles the step reads. */
typedef enum skt_field {
    SKT_QV = 0, SKT_QW, SKT_QC, SKT_QR, SKT_QS, SKT_QI, SKT_POTTEMP, SKT_VELX, SKT_VELY, SKT_VELZ, SKT_PRESSURE,
    SKT_DENSITY,          /* m_densityAir */
    SKT_DENSITY_OLD,      /* m_oldDensityAir */
    SKT_PRESSURE_PROJ,    /* pressure correction of the last projection (m_outputArray) */
    SKT_QRS, SKT_QGR, SKT_QGS, SKT_QGI, SKT_GROUND_P, SKT_GROUND_TIME, SKT_GROUND_T,
    SKT_CLOUD_COVER,      /* m_dummyArrayGround after groundCoverageFactor */
    SKT_ISEN_TEMP, SKT_ISEN_VAPOR, SKT_DEFAULT_VELX, SKT_DEFAULT_VELZ, SKT_DEFAULT_PRESSURE,
    SKT_FIELD_COUNT
} skt_field;

/* 'parameter' and 'direction' of getParamArray (Game/config.h:37-40,62-65). */
enum { SKT_PARAM_POTTEMP = 0, SKT_PARAM_QV, SKT_PARAM_QW, SKT_PARAM_QC, SKT_PARAM_QR, SKT_PARAM_QS, SKT_PARAM_QI,
       SKT_PARAM_WIND, SKT_PARAM_PGROUND, SKT_PARAM_PRESSURE };
enum { SKT_DIR_LEFT = 0, SKT_DIR_RIGHT, SKT_DIR_UP, SKT_DIR_DOWN, SKT_DIR_FORWARD, SKT_DIR_BACKWARD };

/* Inputs the reference reads from its editor every step
 * (Game/environment.cu:1013-1018), plus the ground-changed flag (:992). */
typedef struct skt_env_params {
    float time;            /* seconds since midnight, 43200 at init (Game/environment.cuh:115) */
    int32_t day;           /* day of year, 130 */
    float longitude;       /* used as latitude by irridianceGPU, 52.37 */
    float sun_strength;    /* 1 */
    int32_t pause_diurnal; /* 0 */
    int32_t ground_changed;
} skt_env_params;

/* Serving diagnostics of the current state (skt_env_diagnostics). */
typedef struct skt_diagnostics {
    float max_updraft;       /* largest VELY (the reference's vertical velocity) over sky cells, m/s;
                                -inf if there is no sky cell */
    int32_t cloud_top_level; /* highest level y with QW + QC > threshold at a sky cell; -1 if none */
    float cloud_top_height;  /* cloud_top_level * dx in m (the level heights of Game/readTable.cpp:390);
                                0 if none */
    int32_t pad;
    double surface_rain;     /* ground rain water QGR summed over the nx*nz columns */
} skt_diagnostics;

typedef struct skt_step_stats {
    int32_t pcg_iterations;      /* iterations of the last pressure solve */
    float pcg_residual;          /* stop-test value after the last iteration */
    float pcg_residual0;         /* stop-test value of the initial divergence */
    int32_t n_advect;            /* advectPPMWGPU calls in the last update */
    int32_t substeps[16];        /* substep count of each of those calls */
    /* with FP64 refinement (skt_config.pcg_refine > 0); otherwise 1, pcg_iterations, 0, 0 */
    int32_t pcg_passes;          /* solves that ran iterations in the last projection */
    int32_t pcg_iterations_total;/* iterations over all its solves */
    double pcg_residual64;       /* max |b - A x| in FP64 for the final x */
    double pcg_rhs64;            /* max |b| */
} skt_step_stats;

/* environmentGPU::environmentGPU (Game/environment.cu:71-185). */
int skt_env_create(const skt_config* cfg, skt_env** out);
/* HBM bytes an env of this configuration allocates (the sum of its device
 * arrays), computed without allocating; skt_env_device_bytes gives the same
 * figure for a live env. Used to admit jobs by footprint. No counterpart in
 * the reference, which sizes its arrays at compile time (Game/config.h:3-9). */
int skt_env_footprint(const skt_config* cfg, int64_t* bytes);
int skt_env_device_bytes(skt_env* env, int64_t* bytes);
/* environmentGPU::~environmentGPU (Game/environment.cu:187-247). */
int skt_env_destroy(skt_env* env);
const char* skt_env_last_error(const skt_env* env);
/* Runs all later work on this HIP stream (hipStream_t); NULL = the env's own. */
int skt_env_set_stream(skt_env* env, void* hip_stream);
int skt_env_synchronize(skt_env* env);

/* Ground heights per column (host int32[nx*nz]); flat 0 by default. Replaces
 * the PNoise2D terrain of init (Game/environment.cu:321-340). */
int skt_env_set_ground_height(skt_env* env, const int32_t* ground_height);
/* The ground heights in use (the reference's GHeight, Game/kernelSky.cu; read
 * by isGround, Game/environment.cu:1071-1074), gh[nx*nz]. */
int skt_env_get_ground_height(skt_env* env, int32_t* gh);

/* environmentGPU::init (Game/environment.cu:249-389). All pointers are host
 * arrays: potTemps[N], velField[3N] (x,y,z per cell, MAC faces), Qv[N],
 * groundTemp[G] (written back, as the reference does), groundPres[G],
 * pressures[N] (hPa), smallPressure[ny]. */
int skt_env_init(skt_env* env, const float* potTemps, const float* velField, const float* Qv, float* groundTemp,
                 const float* groundPres, const float* pressures, const float* smallPressure);

/* environmentGPU::updateGPU (Game/environment.cu:391-555): one step. */
int skt_env_update(skt_env* env, float dt, float speed);

/* Per-phase public methods of environmentGPU (Game/environment.cuh:48-93). */
int skt_env_micro_physics_ground(skt_env* env, float dt, float speed, float irradiance);
int skt_env_micro_physics_sky(skt_env* env, float dt, float speed);
/* type: 0 pot. temp, 1 vapour, 2 vel-x, 3 vel-y, 4 vel-z, 5 other mixing ratios */
int skt_env_diffuse(skt_env* env, int32_t field, int32_t type, float dt);
int skt_env_advect_ground_water(skt_env* env, float dt, float speed);
int skt_env_set_temps_at_ground(skt_env* env, float dt, float speed);
/* profile: a per-level field id (SKT_ISEN_TEMP ...) or -1 for the zero profile */
int skt_env_advect_ppm(skt_env* env, int32_t field, int32_t profile, skt_bounds bounds, float dt);
/* fall_type: 0 rain, 1 snow, 2 ice */
int skt_env_advect_precip(skt_env* env, int32_t field, int32_t fall_type, float dt);
int skt_env_pressure_project(skt_env* env, float dt);
/* Solve only; the correction lands in SKT_PRESSURE_PROJ. */
int skt_env_calculate_pressure_project(skt_env* env, float dt);
int skt_env_editor_data(skt_env* env);
int skt_env_irradiance(skt_env* env, float* out);
int skt_env_ground_coverage_factor(skt_env* env);
int skt_env_update_ground_temps(skt_env* env, float dt, float speed, float irradiance);
int skt_env_calculate_buoyancy(skt_env* env, float dt);
/* initDensity + old-density copy (Game/environment.cu:490-492) */
int skt_env_init_density(skt_env* env);
/* calculateNewPressure (Game/kernelSky.cu:1454-1474) */
int skt_env_calculate_new_pressure(skt_env* env);

/* getParamArray (Game/environment.cu:1076-1125): borrowed device pointer. */
int skt_env_get_param_array(skt_env* env, int32_t parameter, int32_t direction, float** dev_ptr);
/* Any field: borrowed device pointer and its element count. */
int skt_env_get_field_ptr(skt_env* env, int32_t field, void** dev_ptr, int64_t* count);
int skt_env_copy_field_to_host(skt_env* env, int32_t field, float* host);
int skt_env_copy_field_from_host(skt_env* env, int32_t field, const float* host);

int skt_env_set_params(skt_env* env, const skt_env_params* params);
int skt_env_get_params(skt_env* env, skt_env_params* params);
int skt_env_get_stats(skt_env* env, skt_step_stats* stats);
/* Storage of a field: SKT_FMT_F32 for all fields but the hydrometeors of the
 * mixed form, which hold value * scale in 16 bits (skt_env_get_field_ptr then
 * points at uint16 words). The host copies convert. No counterpart in the
 * reference, whose fields are all float. */
int skt_env_get_field_format(skt_env* env, int32_t field, int32_t* format, float* scale);

/* Boundary masks in the layout of computeNeighbourGPU's Neigh struct
 * (Game/kernelSky.cu:1284-1316): 12 bytes per cell. host_out[12*N]. */
int skt_env_compute_masks(skt_env* env, uint8_t* host_out);
/* Microphysics phase mask of the last sky-microphysics call, one byte per
 * cell: bit0 sky, bit1 T<0, bit2 PTerm1, bit3 PTerm2, bit4 PTerm3, bit5 WTerm
 * (Game/microPhysics.cu:1128-1130,1177-1178). host_out[N]. */
int skt_env_phase_mask(skt_env* env, uint8_t* host_out);

/* Editor edits (environmentGPU::prepareBrushGPU / prepareSelectionGPU,
 * Game/environment.cu:1127-1260; kernels Game/kernelSky.cu:1476-1663).
 * param is a getParamArray parameter (SKT_PARAM_*): WIND edits the three
 * velocity components along dir, PGROUND raises (ground_erase 0) or erases
 * the ground at the touched cells; other parameters skip ground cells.
 * The brush adds (1 - d^2/size^2)^smoothness * dt * intensity * value over
 * the sphere of radius brush_size around pos (clipped as the reference clips
 * it); the selection sets value over the box [min_pos, max_pos]. A ground
 * change resets the cells it uncovers to the base profiles, becomes the new
 * ground height map and runs the ground-change path of editorDataGPU.
 * Positions beyond the domain are skipped (the reference lets z = nz through
 * its test, one plane past its arrays). Synchronises the env's stream. */
int skt_env_apply_brush(skt_env* env, int32_t param, float brush_size, const int32_t pos[3], float smoothness,
                        float dt, float intensity, float value, const float dir[3], int32_t ground_erase);
int skt_env_apply_selection(skt_env* env, int32_t param, const int32_t min_pos[3], const int32_t max_pos[3],
                            float value, const float dir[3], int32_t ground_erase);

/* environmentGPU::resetParameterGPU (Game/environment.cu:1274-1295), called
 * by the editor for each parameter (Game/editor.cpp:590-598). POTTEMP and QV
 * go back to their isentropic profiles, WIND to the default u and w with
 * v = 0, PRESSURE to the default pressure; the profile resets follow
 * setToDefault (Game/kernelSky.cu:1714-1728), which walks each (x, y) column
 * along z from plane 0 and stops at its first ground cell. QW .. QI are set
 * to 0 in every cell (setToValue). PGROUND is accepted and changes nothing,
 * as in the reference; other ids give SKT_ERR_ARG. Stream-ordered. */
int skt_env_reset_parameter(skt_env* env, int32_t param);

/* Terrain of environmentGPU::init (Game/environment.cu:321-340): PNoise2D
 * (Engine/source/math/geometry.cpp:196-278; seed 100, 12 octaves there) over
 * an nx*nz ground grid, and the ground heights of initGroundHeightGPU
 * (Game/kernelGround.cu:13-19), round(noise * max_height) (0.2 there, which
 * makes every height 0). Computed on the device; noise_out[nx*nz] and/or
 * heights_out[nx*nz] receive the results (either may be NULL). Heights are
 * for skt_env_set_ground_height, which accepts [0, ny). */
int skt_terrain_generate(int32_t device, int32_t seed, int32_t nx, int32_t nz, int32_t octaves, float max_height,
                         float* noise_out, int32_t* heights_out);

/* skt_terrain_generate without device allocation. No counterpart in the
 * reference: it is for callers that cannot call hipMalloc (code being captured
 * into a HIP graph, tight serving loops, harnesses that forbid allocation).
 * The noise and heights are computed into a fixed device scratch array of the
 * library (8 bytes per column, room for SKT_TERRAIN_NOALLOC_MAX_COLUMNS
 * columns, held by each GPU the library runs on until the process ends) and
 * copied to the host outputs. nx * nz above the maximum gives SKT_ERR_ARG.
 * Every call on one GPU shares that GPU's scratch, so calls that overlap on
 * one GPU (from several host threads) race; results are otherwise identical
 * to skt_terrain_generate. */
#define SKT_TERRAIN_NOALLOC_MAX_COLUMNS (1L << 20)
int skt_terrain_generate_noalloc(int32_t device, int32_t seed, int32_t nx, int32_t nz, int32_t octaves,
                                 float max_height, float* noise_out, int32_t* heights_out);

/* ---- Ensembles: several environments advanced by one launch per phase. ----
 * No counterpart in the reference, which holds one environment per process
 * (its grid and boundary data are __constant__ symbols, Game/kernelSky.cu:15-21).
 * A batch borrows initialised environments (members, any grid shapes) and
 * advances all of them by one updateGPU each per skt_batch_step: every phase
 * is one launch whose blocks walk a work list over all members' tiles, and
 * the step is replayed as a HIP graph. Each member's results are those of
 * skt_env_update on it alone, bit for bit. Members must share the solver
 * mode, the PCG iteration limit and the diffusion sweep count; while in a
 * batch a member's work runs on the batch's stream. Not thread-safe. */
typedef struct skt_batch skt_batch;
typedef struct skt_batch_options {
    int32_t use_graph;             /* 1: replay each step as a HIP graph (default); 0: issue the launches */
    int32_t advection_substep_cap; /* advection substep passes per advection in a step (default 4): a member
                                      needing more (Courant number above the cap) is reported by
                                      skt_batch_synchronize */
    int32_t launch_mode;           /* launch width of bodies not tuned: 0 one block per work item (default),
                                      1 or 2 a persistent grid of 1 or 2 rounds of resident blocks */
} skt_batch_options;
typedef struct skt_batch_tuning {
    double step_ms[3];       /* one ensemble step with every body at launch width 0, 1, 2 */
    double tuned_step_ms;    /* the sum of each body's fastest width */
    int32_t bodies;          /* distinct kernel bodies of the step */
    int32_t persistent_bodies; /* launches of the step whose fastest width is persistent */
} skt_batch_tuning;
void skt_batch_options_default(skt_batch_options* o);
int skt_batch_create(int32_t device, skt_batch** out);
int skt_batch_destroy(skt_batch* b);
const char* skt_batch_last_error(const skt_batch* b);
int skt_batch_set_options(skt_batch* b, const skt_batch_options* o);
/* Adds an initialised environment; member_id (out) names it in skt_batch_remove. */
int skt_batch_add(skt_batch* b, skt_env* env, int32_t* member_id);
int skt_batch_remove(skt_batch* b, int32_t member_id);
int skt_batch_size(const skt_batch* b, int32_t* n);
/* One updateGPU(dt, speed) of every member (Game/environment.cu:391-555). */
int skt_batch_step(skt_batch* b, float dt, float speed);
/* Waits for the batch's work; SKT_ERR_STATE if a member's advection needed
 * more substeps than the cap since the last call. */
int skt_batch_synchronize(skt_batch* b);
/* Runs three ensemble steps (ordinary steps of every member), one per launch
 * width, timing every launch, and keeps each launch's fastest width for this
 * set of member shapes (used by every later step of such an ensemble). */
int skt_batch_autotune(skt_batch* b, float dt, float speed, skt_batch_tuning* out);
/* Measurement: one ordinary ensemble step issued launch by launch with events
 * around each, summed per kernel body into 'out' as text lines
 * "body<TAB>milliseconds<TAB>launches" (at most cap - 1 characters). */
int skt_batch_profile(skt_batch* b, float dt, float speed, char* out, int64_t cap);


/* Serving diagnostics reduced on the device (maximum updraft, cloud-top level
 * and height, total ground rain) with cloud meaning QW + QC > cloud_threshold
 * (kg/kg; 1e-5 is the usual choice). No reference counterpart: it replaces the
 * editor's copy of the whole state to the host after every step
 * (editor::GPUSetEnv, Game/editor.cpp:226-269) with two small kernels and a
 * 256-byte readback. Synchronises the env's stream. */
int skt_env_diagnostics(skt_env* env, float cloud_threshold, skt_diagnostics* out);

/* dataClass microphysics region diagnostic (Game/dataClass.cu:26-81):
 * sums of the 33 limited process rates over a box of cells. */
int skt_mphys_region_set(skt_env* env, const int32_t min_pos[3], const int32_t max_pos[3]);
int skt_mphys_region_clear(skt_env* env);
int skt_mphys_region_get(skt_env* env, float out[33]);

/* Measurement (no counterpart in the reference, whose cudaEvent timing is
 * commented out at Game/environment.cu:698-702,816-823). Every launch of a
 * kernel class selected in class_mask (bit k = class k; 0 = off, 0xFF = all)
 * is bracketed by HIP events on the env's stream and accumulated per class:
 * 0 diffusion sweep, 1 PPM pass, 2 PCG search-direction + SpMV, 3 PCG update
 * + stop test, 4 microphysics, 5 max-velocity reduction, 6 precipitation
 * fall, 7 everything else, 8 the multigrid preconditioner (0x1FF = all). */
#define SKT_TIMING_CLASSES 9
typedef struct skt_timing {
    int64_t launches[SKT_TIMING_CLASSES];
    double total_ms[SKT_TIMING_CLASSES];
} skt_timing;
int skt_env_set_timing(skt_env* env, int32_t class_mask); /* also clears the accumulators */
/* Time only every stride-th launch of each enabled class (1, the default:
 * every launch); launches and total_ms then cover the timed launches. Each
 * timed launch costs about 3 us of device idle time, so a sample keeps the
 * timing's cost on a step small. */
int skt_env_set_timing_stride(skt_env* env, int32_t stride);
int skt_env_get_timing(skt_env* env, skt_timing* out);
/* Times a plain device copy of one sky field (nx*ny*nz floats) 'repeats'
 * times: the measured copy ceiling for roofline reporting. */
int skt_env_copy_probe(skt_env* env, int32_t repeats, float* ms_per_copy, double* bytes_per_copy);
int skt_device_synchronize(int32_t device);

/* Synthetic sounding (replaces readTable::initEnvironment,
 * Game/readTable.cpp:364-474, whose input files are not available):
 * conditionally unstable profile, hydrostatic base state integrated in FP64,
 * warm bubbles. Outputs are the init() arguments. */
typedef struct skt_sounding_spec {
    float surface_theta;      /* K (surface theta = surface temperature) */
    float surface_pressure;   /* hPa */
    float lapse_rate;         /* K/m, positive (0.0065) */
    float tropopause_height;  /* m (11000) */
    float surface_qv;         /* kg/kg; used when surface_rh <= 0 */
    float surface_rh;         /* 0..1; > 0 sets surface qv = rh * ws(T_s, p_s) */
    float qv_scale_height;    /* m (2500) */
    float shear_u;            /* u at and above shear_height, m/s (linear from 0) */
    float shear_height;       /* m (6000) */
    int32_t n_bubbles;        /* 0..8 */
    float bubble_amp[8];      /* K */
    float bubble_radius[8];   /* m */
    float bubble_x[8], bubble_y[8], bubble_z[8]; /* centre in m (y = height) */
} skt_sounding_spec;

/* Fills spec for the given BASE seed draw of the randomised generator
 * (surface theta 296-304 K, RH 60-95 %, shear 0-25 m/s, n bubbles of
 * 1.5-3 K); n_bubbles <= 0 gives the fixed single-bubble case
 * (theta 300 K, 14 g/kg, no wind, one 2 K bubble of 500 m at 1 km). */
int skt_sounding_spec_random(skt_sounding_spec* spec, uint64_t seed, int32_t nx, int32_t ny, int32_t nz, float dx,
                             int32_t n_bubbles);
int skt_sounding_generate(const skt_sounding_spec* spec, int32_t nx, int32_t ny, int32_t nz, float dx,
                          float* potTemps, float* velField, float* Qv, float* groundTemp, float* groundPres,
                          float* pressures, float* smallPressure);

/* ---- Sounding ingestion: readTable (Game/readTable.h:8-47, Game/readTable.cpp). ----
 * A radiosonde profile, one row per sample from the ground up: the columns of
 * radioSondeData (Game/readTable.h:8-15). Readers fill caller arrays of
 * 'capacity' rows; when the profile has more rows they return SKT_ERR_ARG with
 * 'rows' set to the count needed. Errors of these functions (no env) are
 * described by skt_input_last_error() on the calling thread. */
typedef struct skt_sounding_table {
    int32_t rows;        /* rows filled (out) or given (in) */
    int32_t capacity;    /* length of each column array */
    float* altitude;     /* m */
    float* pressure;     /* hPa */
    float* dew_point;    /* deg C */
    float* temperature;  /* deg C */
    float* wind_speed;   /* m/s */
    float* wind_dir;     /* deg, direction the wind blows from */
} skt_sounding_table;

/* Text of the last error of a reader or of skt_sounding_init_fields. */
const char* skt_input_last_error(void);

/* readTable::readKNMIFile (Game/readTable.cpp:28-129): a KNMI .mwx archive,
 * member SynchronizedSoundingData.xml. */
int skt_sounding_read_knmi(const char* mwx_path, skt_sounding_table* table);
/* The KNMI row parser on the text of SynchronizedSoundingData.xml
 * (Game/readTable.cpp:43-129). */
int skt_sounding_parse_knmi(const char* xml, int64_t len, skt_sounding_table* table);

/* readTable::readDWDFile (Game/readTable.cpp:134-354): a DWD
 * sekundenwerte_aero zip. The reference always reads launch index 161 of
 * the metadata's list (Game/readTable.cpp:206); date_index selects it here
 * (SKT_DWD_REFERENCE_DATE_INDEX for the reference's choice). */
enum { SKT_DWD_REFERENCE_DATE_INDEX = 161 };
int skt_sounding_read_dwd(const char* zip_path, int32_t date_index, skt_sounding_table* table);
/* The DWD metadata and product parsers on the member texts
 * (Game/readTable.cpp:160-354). */
int skt_sounding_parse_dwd(const char* metadata, int64_t metadata_len, const char* product, int64_t product_len,
                           int32_t date_index, skt_sounding_table* table);

/* readTable::initEnvironment (Game/readTable.cpp:364-474): the init()
 * arrays of an nx*ny*nz grid from a sounding, on the host (same outputs
 * as skt_sounding_generate). The sounding must reach (ny-1)*dx above its
 * first row. */
int skt_sounding_init_fields(const skt_sounding_table* table, int32_t nx, int32_t ny, int32_t nz, float dx,
                             float* potTemps, float* velField, float* Qv, float* groundTemp, float* groundPres,
                             float* pressures, float* smallPressure);
/* readTable::initEnvironment followed by environmentGPU::init
 * (Game/readTable.cpp:364-474, Game/environment.cu:249-389): the per-level
 * values are computed on the host and broadcast over the grid on the device,
 * with no host copy of the 3-D fields. Same state as skt_env_init on the
 * skt_sounding_init_fields arrays. */
int skt_env_init_sounding(skt_env* env, const skt_sounding_table* table);

/* ---- Serving: jobs of 'simulate sounding S for T seconds' on one GPU. ----
 * No counterpart in the reference (an interactive single-domain program).
 * A server owns an ensemble (skt_batch). skt_server_submit queues a job;
 * each skt_server_step admits queued jobs in arrival order while their HBM
 * footprint (skt_env_footprint) fits the budget and a slot is free, uploads
 * each admitted job's initial state (skt_sounding_generate on the host,
 * started by a worker thread for the first four queued jobs from their
 * submission on; then init), advances every active job by one ensemble step,
 * reduces the diagnostics of the jobs whose simulated time reached a report boundary
 * (skt_env_diagnostics; one device wait for all of them), and retires jobs
 * that have run their T seconds, freeing their memory for the queue. Every
 * job runs the step of its own configuration (the server's member_config
 * with the job's grid): its reports are those of the same job run alone,
 * bit for bit, whatever else shares the ensemble. Not thread-safe. */
typedef struct skt_server skt_server;
typedef struct skt_server_options {
    int32_t device;
    float dt;                     /* model step of every job, s (0.5) */
    int64_t hbm_budget_bytes;     /* bytes all active jobs may hold; 0: 90 % of the free device memory */
    int32_t max_active;           /* slots: jobs in flight at most (64) */
    float cloud_threshold;        /* QW + QC above which a cell is cloud in the reports (1e-5) */
    skt_config member_config;     /* solver and precision settings of every job (the grid is the job's) */
    skt_batch_options batch;      /* options of the ensemble */
} skt_server_options;
typedef struct skt_job_request {
    int32_t nx, ny, nz;           /* grid of the job (ny vertical) */
    float dx;                     /* voxel size, m */
    skt_sounding_spec sounding;   /* initial state: skt_sounding_generate of this spec */
    float sim_seconds;            /* T: simulated time to run */
    float report_every;           /* simulated seconds between reports (60); the last report is at T */
} skt_job_request;
enum { SKT_EVENT_ADMITTED = 0, SKT_EVENT_REPORT = 1, SKT_EVENT_DONE = 2 };
typedef struct skt_job_event {
    int64_t job;
    int32_t kind;                 /* SKT_EVENT_* */
    int32_t steps;                /* steps the job has run */
    float sim_time;               /* simulated seconds the job has run */
    int32_t pad;
    double wall_s;                /* server clock (s since skt_server_create) at the event */
    double submitted_s;           /* server clock when the job was submitted */
    skt_diagnostics diag;         /* for SKT_EVENT_REPORT and SKT_EVENT_DONE */
} skt_job_event;
typedef struct skt_server_status {
    int32_t active, queued;       /* jobs in flight and waiting */
    int64_t bytes_in_use;         /* footprints of the active jobs */
    int64_t budget_bytes;
    int64_t steps;                /* ensemble steps run */
    int64_t completed;            /* jobs retired */
    double cell_steps;            /* cells x steps advanced over all jobs */
    double busy_s;                /* wall seconds spent inside skt_server_step */
    double admit_s;               /*   of which building admitted jobs' initial states */
    double replan_s;              /*   of which issuing the steps that followed a change of membership (the
                                       ensemble re-records its plan and re-captures its graph) */
} skt_server_status;
void skt_server_options_default(skt_server_options* o);
int skt_server_create(const skt_server_options* o, skt_server** out);
int skt_server_destroy(skt_server* s);
const char* skt_server_last_error(const skt_server* s);
/* Queues a job; job_id (out) names it in its events. SKT_ERR_ARG when the
 * job can never run (its grid is invalid or its footprint exceeds the budget). */
int skt_server_submit(skt_server* s, const skt_job_request* req, int64_t* job_id);
/* Admission, one ensemble step of the active jobs, reports and retirement
 * (see above). With no active job it only admits. */
int skt_server_step(skt_server* s);
/* Moves up to 'capacity' pending events, oldest first, into out; count (out). */
int skt_server_poll(skt_server* s, skt_job_event* out, int32_t capacity, int32_t* count);
int skt_server_status_get(skt_server* s, skt_server_status* out);

#ifdef __cplusplus
}
#endif

#endif /* SKT_H */
