// Drives the C++ facade (include/skt_environment.hpp) the way the reference's
// game loop drives environmentGPU (Game/game.cpp:37-56): construct, init from
// a sounding, updateGPU. A second environment is stepped through the plain C
// ABI; both must agree bit for bit. Prints "facade ok" on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "skt_environment.hpp"

struct Vec3 {
    float x, y, z;
};

// readTable through the facade: the DWD fixture read, initEnvironment on the
// device, against skt_env_init on the skt_sounding_init_fields arrays.
static int soundingCheck(const char* dwdPath) {
    const int nx = 16, ny = 12, nz = 8;
    const float dx = 100.0f;
    const size_t N = (size_t)nx * ny * nz, G = (size_t)nx * nz;
    skt::readTable reader;
    reader.readDWDFile(dwdPath);
    if (reader.skewTData.data.dataSize == 0) return 1;
    skt::environmentGPU env(nx, ny, nz, dx);
    reader.initEnvironment(env);
    std::vector<float> th(N), vel(3 * N), qv(N), gT(G), gP(G), p(N), ps(ny);
    skt_sounding_table t = reader.skewTData.data.table();
    if (skt_sounding_init_fields(&t, nx, ny, nz, dx, th.data(), vel.data(), qv.data(), gT.data(), gP.data(), p.data(),
                                 ps.data()) != SKT_OK)
        return 1;
    skt::environmentGPU ref(nx, ny, nz, dx);
    ref.init(th.data(), reinterpret_cast<Vec3*>(vel.data()), qv.data(), gT.data(), gP.data(), p.data(), ps.data());
    int bad = 0;
    for (int f = 0; f < SKT_FIELD_COUNT; f++) {
        std::vector<float> a = env.copyToHost((skt_field)f), b = ref.copyToHost((skt_field)f);
        if (a.size() != b.size() || std::memcmp(a.data(), b.data(), a.size() * 4) != 0) bad++;
    }
    bool threw = false;
    try {
        reader.readKNMIFile("/nonexistent/file.mwx");
    } catch (const skt::error& e) {
        threw = e.status == SKT_ERR_INPUT;
    }
    if (!threw) bad++;
    return bad;
}

int main(int argc, char** argv) {
    const int nx = 24, ny = 20, nz = 16;
    const float dx = 100.0f;
    const size_t N = (size_t)nx * ny * nz, G = (size_t)nx * nz;
    skt_sounding_spec spec;
    if (skt_sounding_spec_random(&spec, 20261203ull, nx, ny, nz, dx, 2) != SKT_OK) return 2;
    std::vector<float> th(N), vel(3 * N), qv(N), gT(G), gP(G), p(N), ps(ny);
    if (skt_sounding_generate(&spec, nx, ny, nz, dx, th.data(), vel.data(), qv.data(), gT.data(), gP.data(), p.data(),
                              ps.data()) != SKT_OK)
        return 2;
    std::vector<float> gT2 = gT;

    skt::environmentGPU env(nx, ny, nz, dx);
    env.init(th.data(), reinterpret_cast<Vec3*>(vel.data()), qv.data(), gT.data(), gP.data(), p.data(), ps.data());
    for (int s = 0; s < 3; s++) env.updateGPU(0.5f, 1.0f);

    skt_config cfg;
    skt_config_default(&cfg, nx, ny, nz, dx);
    skt_env* raw = nullptr;
    if (skt_env_create(&cfg, &raw) != SKT_OK) return 3;
    if (skt_env_init(raw, th.data(), vel.data(), qv.data(), gT2.data(), gP.data(), p.data(), ps.data()) != SKT_OK)
        return 3;
    for (int s = 0; s < 3; s++)
        if (skt_env_update(raw, 0.5f, 1.0f) != SKT_OK) return 3;

    int bad = 0;
    for (int f : {SKT_POTTEMP, SKT_VELX, SKT_VELY, SKT_VELZ, SKT_QV, SKT_QW, SKT_PRESSURE}) {
        std::vector<float> a = env.copyToHost((skt_field)f), b(N);
        if (skt_env_copy_field_to_host(raw, f, b.data()) != SKT_OK) return 3;
        for (size_t i = 0; i < N; i++) {
            const size_t y = (i / nx) % ny;
            if (y == 0) continue;  // ground layer
            if (!std::isfinite(a[i]) || std::memcmp(&a[i], &b[i], 4) != 0) bad++;
        }
    }
    // the facade's per-phase methods take the environment's own arrays
    float* theta = env.getParamArray(skt::POTTEMP);
    env.diffuseGPU(theta, 0, 0.5f);
    env.advectPPMWGPU(env.getParamArray(skt::QW), nullptr, skt::boundsEnv{skt::DIRICHLET, skt::DIRICHLET, skt::DIRICHLET},
                      0.5f);
    if (env.getParamArray(skt::PGROUND) != nullptr) bad++;  // unsupported, as in the reference
    bool threw = false;
    try {
        float dummy = 0.0f;
        env.diffuseGPU(&dummy, 0, 0.5f);
    } catch (const skt::error&) {
        threw = true;
    }
    if (!threw) bad++;
    env.synchronize();
    skt_env_destroy(raw);
    if (argc > 1) bad += soundingCheck(argv[1]);
    if (bad) {
        std::printf("facade mismatch: %d\n", bad);
        return 1;
    }
    std::printf("facade ok\n");
    return 0;
}
