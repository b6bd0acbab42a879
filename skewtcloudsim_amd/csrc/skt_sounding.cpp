// Synthetic soundings for skt_env_init, replacing readTable::initEnvironment
// (Game/readTable.cpp:364-474), whose DWD/KNMI input files are not available.
// It produces the same init() contract: potential temperature relative to the
// surface pressure (Game/readTable.cpp:402, theta = T (p_s/p)^(Rd/cp)),
// horizontally uniform wind on MAC faces (:442-460), vapour mixing ratio,
// pressure per cell and per level, ground temperature and pressure.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include "../../include/skt.h"

namespace {

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uni(double a, double b) { return a + (b - a) * (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

const double kRd = 287.0528, kCpd = 1003.5, kG = 9.8076;

// Saturation mixing ratio with the reference's Magnus form (Game/meteoformulas.cuh:16-45), FP64.
double wsat(double Tc, double P) {
    const double Es = 0.61078 * std::exp((17.27 * Tc) / (Tc + 237.3)) * 10.0;
    return (287.0528 / 461.5) * Es / (P - Es);
}

struct Profile {
    double T0, gamma, ztrop, qv0, H;
    double T(double z) const { return z < ztrop ? T0 - gamma * z : T0 - gamma * ztrop; }
    double qv(double z) const { return qv0 * std::exp(-z / H); }
    double Tv(double z) const { return T(z) * (1.0 + 0.608 * qv(z)); }
};

// Hydrostatic pressure dp/dz = -g p / (Rd Tv(z)) integrated in FP64 with RK4.
void hydrostatic(const Profile& pr, double ps, double dz, int levels, std::vector<double>& p) {
    p.assign((size_t)levels, ps);
    double P = ps, z = 0.0;
    const int sub = 64;
    const double h = dz / sub;
    auto f = [&](double zz, double pp) { return -kG * pp / (kRd * pr.Tv(zz)); };
    for (int k = 1; k < levels; k++) {
        for (int s = 0; s < sub; s++) {
            const double k1 = f(z, P), k2 = f(z + 0.5 * h, P + 0.5 * h * k1), k3 = f(z + 0.5 * h, P + 0.5 * h * k2),
                         k4 = f(z + h, P + h * k3);
            P += h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4);
            z += h;
        }
        p[k] = P;
    }
}

}  // namespace

extern "C" {

int skt_sounding_spec_random(skt_sounding_spec* s, uint64_t seed, int32_t nx, int32_t ny, int32_t nz, float dx,
                             int32_t n_bubbles) {
    if (!s || nx < 1 || ny < 2 || nz < 1 || !(dx > 0)) return SKT_ERR_ARG;
    std::memset(s, 0, sizeof(*s));
    s->surface_pressure = 1000.0f;
    s->lapse_rate = 0.0065f;
    s->tropopause_height = 11000.0f;
    s->qv_scale_height = 2500.0f;
    s->shear_height = 6000.0f;
    const double Lx = nx * (double)dx, Lz = nz * (double)dx, Ly = ny * (double)dx;
    if (n_bubbles <= 0) {
        // fixed case: theta 300 K, 14 g/kg, calm, one 2 K bubble of 500 m at 1 km
        s->surface_theta = 300.0f;
        s->surface_qv = 0.014f;
        s->surface_rh = 0.0f;
        s->shear_u = 0.0f;
        s->n_bubbles = 1;
        s->bubble_amp[0] = 2.0f;
        s->bubble_radius[0] = 500.0f;
        s->bubble_x[0] = (float)(0.5 * Lx);
        s->bubble_y[0] = (float)std::fmin(1000.0, 0.5 * Ly);
        s->bubble_z[0] = (float)(0.5 * Lz);
        return SKT_OK;
    }
    SplitMix rng{seed};
    s->surface_theta = (float)rng.uni(296.0, 304.0);
    s->surface_rh = (float)rng.uni(0.60, 0.95);
    s->surface_qv = 0.0f;
    s->shear_u = (float)rng.uni(0.0, 25.0);
    s->n_bubbles = n_bubbles > 8 ? 8 : n_bubbles;
    for (int k = 0; k < s->n_bubbles; k++) {
        s->bubble_amp[k] = (float)rng.uni(1.5, 3.0);
        s->bubble_radius[k] = 500.0f;
        s->bubble_x[k] = (float)rng.uni(0.2 * Lx, 0.8 * Lx);
        s->bubble_z[k] = (float)rng.uni(0.2 * Lz, 0.8 * Lz);
        s->bubble_y[k] = (float)std::fmin(rng.uni(500.0, 2000.0), 0.6 * Ly);
    }
    return SKT_OK;
}

int skt_sounding_generate(const skt_sounding_spec* s, int32_t nx, int32_t ny, int32_t nz, float dx, float* th,
                          float* vel, float* qv, float* gT, float* gP, float* p, float* pS) {
    if (!s || !th || !vel || !qv || !gT || !gP || !p || !pS || nx < 1 || ny < 2 || nz < 1 || !(dx > 0))
        return SKT_ERR_ARG;
    const double ps = s->surface_pressure;
    double qv0 = s->surface_qv;
    if (s->surface_rh > 0.0f) qv0 = s->surface_rh * wsat(s->surface_theta - 273.15, ps);
    Profile pr{s->surface_theta, s->lapse_rate, s->tropopause_height, qv0, s->qv_scale_height};
    std::vector<double> plev;
    hydrostatic(pr, ps, dx, ny, plev);
    const double kap = kRd / kCpd;
    std::vector<double> thl((size_t)ny), ql((size_t)ny), ul((size_t)ny);
    for (int y = 0; y < ny; y++) {
        const double z = y * (double)dx;
        thl[y] = pr.T(z) * std::pow(ps / plev[y], kap);
        ql[y] = pr.qv(z);
        const double zs = s->shear_height > 0 ? std::fmin(z, (double)s->shear_height) / s->shear_height : 1.0;
        ul[y] = s->shear_u * zs;
        pS[y] = (float)plev[y];
    }
    // every cell is a function of its own position: planes are filled by up
    // to 16 threads, each a contiguous range of z
    auto planes = [&](int za, int zb) {
        for (int z = za; z < zb; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    const long i = x + (long)nx * (y + (long)ny * z);
                    double t = thl[y];
                    for (int b = 0; b < s->n_bubbles && b < 8; b++) {
                        const double ddx = x * (double)dx - s->bubble_x[b], ddy = y * (double)dx - s->bubble_y[b],
                                     ddz = z * (double)dx - s->bubble_z[b];
                        const double r = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
                        if (r < s->bubble_radius[b]) {
                            const double c = std::cos(0.5 * M_PI * r / s->bubble_radius[b]);
                            t += s->bubble_amp[b] * c * c;
                        }
                    }
                    th[i] = (float)t;
                    qv[i] = (float)ql[y];
                    p[i] = (float)plev[y];
                    vel[3 * i + 0] = (float)ul[y];
                    vel[3 * i + 1] = 0.0f;
                    vel[3 * i + 2] = 0.0f;
                }
    };
    const long cells = (long)nx * ny * nz;
    const int nt = (int)std::min<long>({16L, (long)std::max(1u, std::thread::hardware_concurrency()), (long)nz,
                                        std::max(1L, cells / (1L << 18))});
    if (nt <= 1) {
        planes(0, nz);
    } else {
        std::vector<std::thread> pool;
        pool.reserve((size_t)nt);
        for (int k = 0; k < nt; k++) pool.emplace_back(planes, (int)((long)nz * k / nt), (int)((long)nz * (k + 1) / nt));
        for (std::thread& t : pool) t.join();
    }
    for (long g = 0; g < (long)nx * nz; g++) {
        gT[g] = (float)pr.T(0.0);
        gP[g] = (float)ps;
    }
    return SKT_OK;
}

}  // extern "C"
