/*
 * skt_environment.hpp — header-only C++ facade over skt.h with the class and
 * method names of the reference's environmentGPU (Game/environment.cuh:9-183)
 * and the microphysics-region part of dataClass (Game/dataClass.cuh:67-89).
 *
 * A caller written against the reference swaps `#include "environment.cuh"`
 * for this header and links libskt.so. Differences a caller sees:
 *   - the grid shape, voxel size and solver settings are constructor
 *     arguments (the reference fixes them at compile time, Game/config.h:3-9);
 *   - device arrays handed to the per-phase methods must be the
 *     environment's own (obtained from getParamArray / fieldPtr), exactly the
 *     arrays the reference passes internally (m_envGrid.*, m_isentropicTemp,
 *     m_defaultVelX/Z, m_dummyArray);
 *   - errors print to std::cerr and throw skt::error instead of breaking into
 *     the debugger (the reference's checkCudaErrors + __debugbreak);
 *   - the editor inputs (time, day, latitude, sun strength, pause, ground
 *     changed) are set with setParams() instead of being read from the editor.
 */
#ifndef SKT_ENVIRONMENT_HPP
#define SKT_ENVIRONMENT_HPP

#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "skt.h"

namespace skt {

struct error : std::runtime_error {
    int status;
    error(int st, const std::string& what) : std::runtime_error(what), status(st) {}
};

// boundsEnv / boundCon (Game/config.h:43-59).
enum boundCon { NEUMANN = SKT_NEUMANN, DIRICHLET = SKT_DIRICHLET, CUSTOM = SKT_CUSTOM };
struct boundsEnv {
    boundCon sides, up, ground;
};
// parameter / direction of getParamArray (Game/config.h:37-40,62-65).
enum parameter { POTTEMP = 0, QV, QW, QC, QR, QS, QI, WIND, PGROUND, PRESSURE };
enum direction { LEFT = 0, RIGHT, UP, DOWN, FORWARD, BACKWARD };

class environmentGPU {
public:
    // environmentGPU::environmentGPU (Game/environment.cu:71-185) for an
    // nx * ny * nz grid (ny vertical) of voxel size dx [m].
    environmentGPU(int nx, int ny, int nz, float dx, int device = 0) {
        skt_config cfg;
        skt_config_default(&cfg, nx, ny, nz, dx);
        cfg.device = device;
        create(cfg);
    }
    explicit environmentGPU(const skt_config& cfg) { create(cfg); }
    ~environmentGPU() {
        if (m_env) skt_env_destroy(m_env);
    }
    environmentGPU(const environmentGPU&) = delete;
    environmentGPU& operator=(const environmentGPU&) = delete;

    // init (Game/environment.cu:249-389); V3 is any 3-float vector type such
    // as glm::vec3 (velField holds x, y, z of each cell, MAC faces).
    template <class V3>
    void init(float* potTemps, V3* velField, float* Qv, float* groundTemp, float* groundPres, float* pressures,
              float* smallPressure) {
        static_assert(sizeof(V3) == 3 * sizeof(float), "velField must be 3 floats per cell");
        check(skt_env_init(m_env, potTemps, reinterpret_cast<const float*>(velField), Qv, groundTemp, groundPres,
                           pressures, smallPressure),
              "init");
    }

    void updateGPU(float dt, const float speed) { check(skt_env_update(m_env, dt, speed), "updateGPU"); }

    void microPhysicsGroundGPU(const float dt, const float speed, const float irradiance) {
        check(skt_env_micro_physics_ground(m_env, dt, speed, irradiance), "microPhysicsGroundGPU");
    }
    void microPhysicsSkyGPU(const float dt, const float speed) {
        check(skt_env_micro_physics_sky(m_env, dt, speed), "microPhysicsSkyGPU");
    }
    // type: temp = 0, vapor = 1, velX = 2, velY = 3, velZ = 4, other = 5
    void diffuseGPU(float* diffuseArray, int type, const float dt) {
        check(skt_env_diffuse(m_env, fieldOf(diffuseArray), type, dt), "diffuseGPU");
    }
    void advectGroundWater(const float dt, const float speed) {
        check(skt_env_advect_ground_water(m_env, dt, speed), "advectGroundWater");
    }
    void setTempsAtGround(const float dt, const float speed) {
        check(skt_env_set_temps_at_ground(m_env, dt, speed), "setTempsAtGround");
    }
    // defaultVal: one of the environment's per-level profiles, or nullptr for
    // the zero profile (the reference's m_dummyArray).
    void advectPPMWGPU(float* array, const float* defaultVal, boundsEnv b, const float dt) {
        const int prof = defaultVal ? fieldOf(defaultVal) : -1;
        const skt_bounds sb{b.sides, b.up, b.ground};
        check(skt_env_advect_ppm(m_env, fieldOf(array), prof, sb, dt), "advectPPMWGPU");
    }
    // fallVelType: rain = 0, snow = 1, hail = 2
    void advectPrecip(float* array, const int fallVelType, const float dt) {
        check(skt_env_advect_precip(m_env, fieldOf(array), fallVelType, dt), "advectPrecip");
    }
    void pressureProject(const float dt) { check(skt_env_pressure_project(m_env, dt), "pressureProject"); }
    // The solve only; the correction is left in the environment's
    // pressure-correction array, which outputPressure must be.
    void calculatePressureProject(float* outputPressure, const float dt) {
        if (fieldOf(outputPressure) != SKT_PRESSURE_PROJ)
            fail(SKT_ERR_ARG, "calculatePressureProject: output must be the pressure-correction array");
        check(skt_env_calculate_pressure_project(m_env, dt), "calculatePressureProject");
    }

    void editorDataGPU() { check(skt_env_editor_data(m_env), "editorDataGPU"); }
    // prepareBrushGPU / prepareSelectionGPU (Game/environment.cu:1127-1260);
    // I3 / F3 are any 3-int / 3-float vector types (int3, float3, glm types).
    template <class I3, class F3>
    void prepareBrushGPU(parameter paramType, const float brushSize, const I3 mousePos, const float brushSmoothnes,
                         const float dt, const float brushIntensity, const float applyValue, const F3 valueDir,
                         const bool groundErase) {
        static_assert(sizeof(I3) == 3 * sizeof(int32_t) && sizeof(F3) == 3 * sizeof(float), "3-vectors");
        check(skt_env_apply_brush(m_env, paramType, brushSize, reinterpret_cast<const int32_t*>(&mousePos),
                                  brushSmoothnes, dt, brushIntensity, applyValue,
                                  reinterpret_cast<const float*>(&valueDir), groundErase),
              "prepareBrushGPU");
        refreshGround();
    }
    template <class I3, class F3>
    void prepareSelectionGPU(parameter paramType, const I3 minPos, const I3 maxPos, const float applyValue,
                             const F3 valueDir, const bool groundErase) {
        static_assert(sizeof(I3) == 3 * sizeof(int32_t) && sizeof(F3) == 3 * sizeof(float), "3-vectors");
        check(skt_env_apply_selection(m_env, paramType, reinterpret_cast<const int32_t*>(&minPos),
                                      reinterpret_cast<const int32_t*>(&maxPos), applyValue,
                                      reinterpret_cast<const float*>(&valueDir), groundErase),
              "prepareSelectionGPU");
        refreshGround();
    }
    // resetParameterGPU (Game/environment.cu:1274-1295)
    void resetParameterGPU(parameter paramType) {
        check(skt_env_reset_parameter(m_env, paramType), "resetParameterGPU");
    }
    float irridianceGPU() {
        float v = 0.0f;
        check(skt_env_irradiance(m_env, &v), "irridianceGPU");
        return v;
    }
    void groundCoverageFactor() { check(skt_env_ground_coverage_factor(m_env), "groundCoverageFactor"); }
    void updateGroundTemps(const float dt, const float speed, const float irridiance) {
        check(skt_env_update_ground_temps(m_env, dt, speed, irridiance), "updateGroundTemps");
    }
    void calculateBuoyancy(const float dt) { check(skt_env_calculate_buoyancy(m_env, dt), "calculateBuoyancy"); }
    // isGround (Game/environment.cu:1071-1074) reads the ground height of
    // column x (z = 0 plane in the reference's 2D signature) on the host.
    bool isGround(int x, int y) {
        if (m_gh.empty()) return y <= 0;
        return y <= m_gh[(size_t)x];
    }
    // Unsupported parameters print an error and give nullptr, as the
    // reference does (Game/environment.cu:1117-1124).
    float* getParamArray(parameter type, direction windDir = RIGHT) {
        float* p = nullptr;
        if (skt_env_get_param_array(m_env, type, windDir, &p) != SKT_OK) {
            std::cerr << "skt: getParamArray: " << skt_env_last_error(m_env) << std::endl;
            return nullptr;
        }
        return p;
    }

    // Additions of the drop-in (no reference counterpart).
    void setGroundHeight(const std::vector<int32_t>& gh) {
        check(skt_env_set_ground_height(m_env, gh.data()), "setGroundHeight");
        m_gh = gh;
    }
    void setParams(const skt_env_params& p) { check(skt_env_set_params(m_env, &p), "setParams"); }
    skt_env_params params() {
        skt_env_params p;
        check(skt_env_get_params(m_env, &p), "params");
        return p;
    }
    // Serving diagnostics (skt_env_diagnostics; no reference counterpart).
    skt_diagnostics diagnostics(float cloudThreshold = 1e-5f) {
        skt_diagnostics d;
        check(skt_env_diagnostics(m_env, cloudThreshold, &d), "diagnostics");
        return d;
    }
    skt_step_stats stats() {
        skt_step_stats s;
        check(skt_env_get_stats(m_env, &s), "stats");
        return s;
    }
    float* fieldPtr(skt_field f) {
        void* p = nullptr;
        int64_t n = 0;
        check(skt_env_get_field_ptr(m_env, f, &p, &n), "fieldPtr");
        return static_cast<float*>(p);
    }
    std::vector<float> copyToHost(skt_field f) {
        void* p = nullptr;
        int64_t n = 0;
        check(skt_env_get_field_ptr(m_env, f, &p, &n), "copyToHost");
        std::vector<float> h((size_t)n);
        check(skt_env_copy_field_to_host(m_env, f, h.data()), "copyToHost");
        return h;
    }
    void synchronize() { check(skt_env_synchronize(m_env), "synchronize"); }
    skt_env* handle() { return m_env; }

    // dataClass microphysics region (Game/dataClass.cu:26-81).
    void confirmMicroPhysCheckRegion(const int32_t minPos[3], const int32_t maxPos[3]) {
        check(skt_mphys_region_set(m_env, minPos, maxPos), "confirmMicroPhysCheckRegion");
    }
    void cancelMicroPhysCheckRegion() { check(skt_mphys_region_clear(m_env), "cancelMicroPhysCheckRegion"); }
    void microPhysicsData(float out[33]) { check(skt_mphys_region_get(m_env, out), "microPhysicsData"); }

private:
    skt_env* m_env = nullptr;
    std::vector<int32_t> m_gh;

    size_t m_columns = 0;  // nx * nz

    // the host copy isGround reads, after an edit may have changed the ground
    void refreshGround() {
        m_gh.resize(m_columns);
        check(skt_env_get_ground_height(m_env, m_gh.data()), "groundHeight");
    }

    void create(const skt_config& cfg) {
        m_columns = (size_t)cfg.nx * (size_t)cfg.nz;
        const int rc = skt_env_create(&cfg, &m_env);
        if (rc != SKT_OK) fail(rc, "environmentGPU: skt_env_create failed");
    }
    [[noreturn]] void fail(int st, const std::string& msg) {
        std::cerr << "skt: " << msg << std::endl;
        throw error(st, msg);
    }
    void check(int st, const char* what) {
        if (st != SKT_OK) fail(st, std::string(what) + ": " + skt_env_last_error(m_env));
    }
    // The field id of one of the environment's own device arrays.
    int fieldOf(const float* dev) {
        for (int f = 0; f < SKT_FIELD_COUNT; f++) {
            void* p = nullptr;
            int64_t n = 0;
            if (skt_env_get_field_ptr(m_env, f, &p, &n) == SKT_OK && p == dev) return f;
        }
        fail(SKT_ERR_ARG, "device array is not one of this environment's fields");
    }
};

// radioSondeData (Game/readTable.h:8-47) with owned columns.
struct radioSondeData {
    std::vector<float> altitude, pressure, dewPoint, temperature, windSpeed, windDir;
    size_t dataSize = 0;
    skt_sounding_table table() {
        return skt_sounding_table{(int32_t)dataSize, (int32_t)dataSize, altitude.data(), pressure.data(),
                                  dewPoint.data(), temperature.data(), windSpeed.data(), windDir.data()};
    }
};
struct skewTInfo {
    radioSondeData data;
};

// readTable (Game/readTable.h:49-88): radiosonde readers and the initial
// state built from the profile. Errors throw skt::error (the reference
// prints and returns without data).
class readTable {
public:
    void readKNMIFile(const char* file) {
        read([&](skt_sounding_table* t) { return skt_sounding_read_knmi(file, t); });
    }
    // The reference reads launch index 161 of the archive's metadata.
    void readDWDFile(const char* file, int dateIndex = SKT_DWD_REFERENCE_DATE_INDEX) {
        read([&](skt_sounding_table* t) { return skt_sounding_read_dwd(file, dateIndex, t); });
    }
    // initEnvironment (Game/readTable.cpp:364-474) into env.
    void initEnvironment(environmentGPU& env) {
        skt_sounding_table t = skewTData.data.table();
        const int st = skt_env_init_sounding(env.handle(), &t);
        if (st != SKT_OK) throw error(st, std::string("initEnvironment: ") + skt_env_last_error(env.handle()));
    }
    skewTInfo skewTData;

private:
    template <class F>
    void read(F call) {
        radioSondeData& d = skewTData.data;
        size_t cap = 8192;
        for (int attempt = 0; attempt < 2; attempt++) {
            for (auto* c : {&d.altitude, &d.pressure, &d.dewPoint, &d.temperature, &d.windSpeed, &d.windDir})
                c->assign(cap, 0.0f);
            skt_sounding_table t{0, (int32_t)cap, d.altitude.data(), d.pressure.data(), d.dewPoint.data(),
                                 d.temperature.data(), d.windSpeed.data(), d.windDir.data()};
            const int st = call(&t);
            if (st == SKT_ERR_ARG && (size_t)t.rows > cap) {
                cap = (size_t)t.rows;
                continue;
            }
            if (st != SKT_OK) {
                d = radioSondeData{};
                throw error(st, std::string("readTable: ") + skt_input_last_error());
            }
            for (auto* c : {&d.altitude, &d.pressure, &d.dewPoint, &d.temperature, &d.windSpeed, &d.windDir})
                c->resize((size_t)t.rows);
            d.dataSize = (size_t)t.rows;
            return;
        }
        throw error(SKT_ERR_INPUT, "readTable: profile did not fit");
    }
};

}  // namespace skt

#endif  // SKT_ENVIRONMENT_HPP
