// HIP kernels of the column-local and pointwise phases of the skewT cloud
// step for gfx950: ground phase, precipitation fall, buoyancy, equation of
// state, boundary velocity reset, bulk microphysics, masks and init helpers.
// Same layout and launch conventions as skt_kernels.hip.
#include <hip/hip_runtime.h>
#include <climits>
#include <algorithm>
#include <mutex>
#include "skt_internal.h"
#include "skt_launch.h"
#include "skt_micro.h"
#include "skt_reduce.h"

namespace skt {

// ---------------------------------------------------------- ground phase ---
// advectGroundWaterGPU, Game/kernelSky.cu:120-253, reading a pre-step snapshot
// of both stores (the reference's in-place tile update leaves its interior
// halo unfilled; every column reading its neighbours' old values is intended).
struct KGroundWater {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, const float* __restrict__ rs0, const float* __restrict__ gr0, float* __restrict__ rs,
                           float* __restrict__ gr, float dt, float speed, float dx) {
        const long g = vb.id * 256L + threadIdx.x;
        if (g >= (long)G.X * G.Z) return;
        const int x = (int)(g % G.X), zc = (int)(g / G.X);
        const long iL = x > 0 ? g - 1 : g, iR = x < G.X - 1 ? g + 1 : g;
        const long iB = zc > 0 ? g - G.X : g, iF = zc < G.Z - 1 ? g + G.X : g;
        const float n = 1 / 0.03f;
        const float kk = 5e-5f;
        const float diffR = 1 / (dx * dx) * 1.75e-3f;
        const float diffW = 1 / (dx * dx) * 1.8e-5f;
        const float grC = gr0[g], rsC = rs0[g];
        float fGr = grC, fRs = rsC;
        const float hC = float(G.gh[g]);
        const long nb[4] = {iL, iR, iB, iF};
    #pragma unroll
        for (int k = 0; k < 4; k++) {
            const long j = nb[k];
            const float hFrom = float(G.gh[j]), grFrom = gr0[j], rsFrom = rs0[j];
            const float slope = hFrom - hC / 1.0f;
            const bool fromPrev = slope > 0.0f;
            const float sRain = fromPrev ? grFrom : grC;
            const float sWater = fromPrev ? rsFrom : rsC;
            float changeRain = 0.0f;
            if (sRain != 0.0f) {
                const float sp = n * powPos(sRain, 2.0f / 3.0f);
                const float vel = sp * sqrtf(slope);
                changeRain = vel * sRain;
            }
            const float changeWater = kc::g * kk * sWater * slope;
            const float lapR = (grFrom - grC) * diffR;
            const float lapSR = (rsFrom - rsC) * diffW;
            const bool fpR = lapR > 0.0f, fpSR = lapSR > 0.0f;
            const float fgr = dt * fminf(fabsf(lapR) * speed, fpR ? grFrom : grC) * (lapR > 0.0f ? 1.0f : -1.0f);
            const float fgw = dt * fminf(fabsf(lapSR) * speed, fpSR ? rsFrom : rsC) * (lapSR > 0.0f ? 1.0f : -1.0f);
            const float sfr = dt * fminf(fabsf(changeRain) * speed, fromPrev ? grFrom : grC) * (changeRain > 0.0f ? 1.0f : -1.0f);
            const float sfw = dt * fminf(fabsf(changeWater) * speed, fromPrev ? rsFrom : rsC) * (changeWater > 0.0f ? 1.0f : -1.0f);
            fGr += fgr + sfr;
            fRs += fgw + sfw;
        }
        gr[g] = fGr;
        rs[g] = fRs;
    }
};
// Copy of two arrays of n floats (the pre-step snapshot of the ground stores).
struct KCopy2 {
    static __device__ __forceinline__ void run(VB vb, const float* __restrict__ a, const float* __restrict__ b,
                                               float* __restrict__ da, float* __restrict__ db, long n) {
        for (long i = vb.id * 256L + threadIdx.x; i < n; i += (long)vb.n * 256) {
            da[i] = a[i];
            db[i] = b[i];
        }
    }
};
hipError_t launch_ground_water_snapshot(EnvState& e, float dt, float speed) {
    float* rs0 = e.work[0];
    float* gr0 = e.work[0] + e.G;
    hipError_t err = klaunch<KCopy2>(e, grid1(e.G), dim3(256), 0, e.f[SKT_QRS], e.f[SKT_QGR], rs0, gr0, e.G);
    if (err != hipSuccess) return err;
    return klaunch<KGroundWater>(e, grid1(e.G), dim3(256), 0, geom(e), rs0, gr0, e.f[SKT_QRS], e.f[SKT_QGR], dt,
                                 speed, e.dx);
}

struct GroundArgs {
    float *Qv, *Qr, *Qs, *Qi, *Qc, *Qw, *th, *u, *p, *rho;
    float *Qrs, *Qgr, *Qgs, *Qgi, *gP, *gt, *gT, *cover;
    const float* irrDev;  // irradiance on the device (null: the irr argument)
    Store sQr, sQs, sQi, sQc, sQw;  // storage of the hydrometeor fields
};

// The column-local part of the ground phase of updateGPU
// (Game/environment.cu:427-455,469) in the reference's order. A block holds
// GC_COLS ground columns. The cloud cover's column sum is the one step that
// walks a whole column: the block's 256 threads load the levels of its
// columns side by side (level y of column c by the thread of chunk
// y mod GC_CHUNKS) and put each level's term (Qc + Qw) dx into LDS, GC_LV
// levels at a time; the chunk-0 thread of each column adds the terms in
// level order, as the reference's loop does. Everything after it is one
// thread per column (the chunk-0 threads).
constexpr int GC_COLS = 32, GC_CHUNKS = 8, GC_LV = 128;
static_assert(GC_COLS * GC_CHUNKS == 256, "ground-column blocks hold 256 threads");
template <bool MIX>
struct KGroundColumns {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, const GroundArgs& a, int stages, float dt,
                                               float speed, float irrArg, float dx, const Gammas& gm) {
        const float irr = a.irrDev ? *a.irrDev : irrArg;
        // thread (tx, ty) of the 64 x 4 block: column tx mod 32, chunk
        // tx / 32 + 2 ty (the chunk-0 threads are lanes 0-31 of the first wave)
        const int cl = (int)threadIdx.x & (GC_COLS - 1);
        const int ch = ((int)threadIdx.x >> 5) + 2 * (int)threadIdx.y;
        const long g = vb.id * (long)GC_COLS + cl;
        const bool colOk = g < (long)G.X * G.Z;
        const long gs = colOk ? g : 0;
        const int x = (int)(gs % G.X), zc = (int)(gs / G.X);
        const float kap = kc::Rsd / kc::Cpd;
        const int gh = G.gh[gs];
        if (stages & 1) {  // calculateCloudCoverGPU, Game/kernelSky.cu:1034-1044
            __shared__ float terms[GC_LV][GC_COLS];
            float tot = 0.0f;
            for (int base = 0; base < G.Y; base += GC_LV) {  // uniform
                const int top = min(G.Y, base + GC_LV);
                if (colOk) {
                    // every level of the column is in memory: the loads are
                    // unconditional, so a thread's are issued together, and
                    // the sum below reads only the sky levels' terms
#pragma unroll
                    for (int k = 0; k < GC_LV / GC_CHUNKS; k++) {
                        const int y = base + ch + k * GC_CHUNKS;
                        if (y < top) {
                            const long i = gidx(G, x, y, zc);
                            terms[y - base][cl] = (ldS<MIX>(a.Qc, i, a.sQc) + ldS<MIX>(a.Qw, i, a.sQw)) * dx;
                        }
                    }
                }
                __syncthreads();
                if (colOk && ch == 0)
                    for (int y = max(gh + 1, base); y < top; y++) tot += terms[y - base][cl];
                __syncthreads();
            }
            if (colOk && ch == 0) a.cover[g] = fminf(tot / 1.2f, 1.0f);
        }
        if (!colOk || ch != 0) return;
        if (stages & 2) {  // calculateGroundTempGPU, Game/kernelSky.cu:1046-1059 (T^4 widened to double)
            const float gtv = a.gT[g];
            const double T4 = gtv * gtv * gtv * gtv;
            a.gT[g] += (dt * speed) * ((1 - a.cover[g]) * (((1 - 0.25f) * irr - kc::ge * kc::oo * T4) / (1.0f * 1500.0f * kc::Cpds)));
        }
        if ((stages & (4 | 8 | 16)) && gh + 1 < G.Y) {
            const long i = gidx(G, x, gh + 1, zc);
            if (stages & 4) {  // calculatePrecipHittingGroundMicroPhysicsGPU, Game/microPhysics.cu:1544-1636
                const float Qgr = a.Qgr[g] > 1e-18f ? a.Qgr[g] : 0.0f;
                const float Qgs = a.Qgs[g] > 1e-18f ? a.Qgs[g] : 0.0f;
                const float Qgi = a.Qgi[g] > 1e-18f ? a.Qgi[g] : 0.0f;
                const float Qv = a.Qv[i] > 1e-18f ? a.Qv[i] : 0.0f;
                const float qr0 = ldS<MIX>(a.Qr, i, a.sQr), qs0 = ldS<MIX>(a.Qs, i, a.sQs);
                const float qi0 = ldS<MIX>(a.Qi, i, a.sQi);
                const float Qr = qr0 > 1e-18f ? qr0 : 0.0f;
                const float Qs = qs0 > 1e-18f ? qs0 : 0.0f;
                const float Qi = qi0 > 1e-18f ? qi0 : 0.0f;
                float freeze = 0.0f;
                const float tGC = a.gT[g] - 273.15f;
                const float ps = a.p[i];
                const float wind = a.u[i];
                float3 fv = make_float3(0.0f, 0.0f, 0.0f);
                const float Dair = a.rho[i];
                const float tAK = a.th[i] * powNear1(ps / a.gP[g], kap);
                if (Qgr > 0.0f || Qgs > 0.0f || Qgi > 0.0f) fv = d_fallVel(Qgr, Qgs, Qgi, Dair, 3, gm.R48, gm.S425, gm.S425);
                const float GR = dt * fminf(Qr * speed, Qr);
                const float GS = dt * fminf(Qs * speed, Qs);
                const float GI = dt * fminf(Qi * speed, Qi);
                float PGRFR = mp::PGRFR(tAK, tGC, ps, Qr, Qv, Qi, Qgr, Qgs, Dair, wind, fv);
                if (tAK < 273.15f) {
                    PGRFR = dt * fminf(Qr, speed * PGRFR);
                    freeze += PGRFR;
                }
                stS<MIX>(a.Qr, i, qr0 + -GR, a.sQr);
                stS<MIX>(a.Qs, i, qs0 + -GS, a.sQs);
                stS<MIX>(a.Qi, i, qi0 + -GI, a.sQi);
                a.Qgr[g] += GR - PGRFR;
                a.Qgs[g] += GS;
                a.Qgi[g] += GI + PGRFR;
                const float Mair = 0.02896f, Mwater = 0.01802f;
                const float XV = (Qv / Mwater) / ((Qv / Mwater) + (1 - Qv) / Mair);
                const float Mth = XV * Mwater + (1 - XV) * Mair;
                const float YV = XV * (Mwater / Mth);
                const float yth = YV * 1.33f + (1 - YV) * 1.4f;
                const float cpth = kc::R / (Mth * (yth - 1));
                float sum = 0.0f;
                sum += kc::Lf / cpth * freeze;
                a.gT[g] += sum;
            }
            if (stages & 8) {  // calculateGroundMicroPhysicsGPU, Game/microPhysics.cu:1419-1541
                const float Qrs = a.Qrs[g] > 1e-18f ? a.Qrs[g] : 0.0f;
                const float Qgr = a.Qgr[g] > 1e-18f ? a.Qgr[g] : 0.0f;
                const float Qgs = a.Qgs[g] > 1e-18f ? a.Qgs[g] : 0.0f;
                const float Qgi = a.Qgi[g] > 1e-18f ? a.Qgi[g] : 0.0f;
                const float Qv = a.Qv[i] > 1e-18f ? a.Qv[i] : 0.0f;
                float condens = 0.0f, freeze = 0.0f, depos = 0.0f;
                const float tGC = a.gT[g] - 273.15f;
                const float ps = a.p[i];
                const float wind = a.u[i];
                const float cc = a.cover[g];
                const float Dair = a.rho[i];
                float tme = a.gt[g];
                const float tAK = a.th[i] * powNear1(ps / a.gP[g], kap);
                const float sI = d_slope(Dair, Qgi, 2);
                float PGREVP = mp::PGREVP(tAK, tGC, ps, Qv, Qgr, irr, Dair, wind);
                float PGSMLT = mp::PGSMLT(tAK, tGC, ps, Qgs, Qv, Qgr, irr, cc, Dair, wind, dt);
                float PGGMLT = mp::PGGMLT(tGC, Qgi, Dair, sI);
                float PGGFR = mp::PGGFR(tAK, tGC, ps, Qgr, Qv, Qgi, Dair, irr, cc, wind, dt);
                float PGSFR = mp::PGSFR(tAK, tGC, ps, Qv, Qgr, Qgs, Dair, wind, dt);
                float PGFLW = mp::PGFLW(Qgr);
                float PGDEVP = mp::PGDEVP(&tme, Qgr, Qgs, Qgi, Qrs);
                a.gt[g] = tme;
                PGREVP = dt * fminf(Qgr, speed * PGREVP); condens -= PGREVP;
                PGSMLT = dt * fminf(Qgs, speed * PGSMLT); freeze -= PGSMLT;
                PGGMLT = dt * fminf(Qgi, speed * PGGMLT); freeze -= PGGMLT;
                PGGFR = dt * fminf(Qgr, speed * PGGFR); freeze += PGGFR;
                PGSFR = dt * fminf(Qgr, speed * PGSFR); freeze += PGSFR;
                PGFLW = dt * fminf(Qgr, speed * PGFLW);
                PGDEVP = dt * fminf(Qrs, speed * PGDEVP); condens -= PGDEVP;
                a.Qrs[g] += PGFLW - PGDEVP;
                a.Qv[i] += PGREVP + PGDEVP;
                a.Qgr[g] += PGSMLT + PGGMLT - PGREVP - PGGFR - PGSFR - PGFLW;
                a.Qgs[g] += PGSFR - PGSMLT;
                a.Qgi[g] += PGGFR - PGGMLT;
                const float cps = 1500.0f, rhos = 1500.0f, dl = 0.1f;
                const float ms = rhos * dl;
                float sum = 0.0f;
                sum += d_Lwater(tGC) / (ms * cps) * condens;
                sum += d_Lice(tGC) / (ms * cps) * depos;
                sum += kc::Lf / (ms * cps) * freeze;
                a.gT[g] += sum;
            }
            if ((stages & 16) && gh + 1 < G.Y - 1) {  // setTempsAtGroundGPU, Game/kernelSky.cu:255-268
                const float T = d_potTemp(a.gT[g] - 273.15f, a.p[i], a.gP[g]) + 273.15f;
                const float d2 = a.th[i] - T;
                a.th[i] -= d2 * fminf(1.0f, dt * speed);
            }
        }
    }
};
hipError_t launch_ground_columns(EnvState& e, int stages, float dt, float speed, float irr) {
    GroundArgs a{e.f[SKT_QV], e.f[SKT_QR], e.f[SKT_QS], e.f[SKT_QI], e.f[SKT_QC], e.f[SKT_QW], e.f[SKT_POTTEMP],
                 e.f[SKT_VELX], e.f[SKT_PRESSURE], e.f[SKT_DENSITY], e.f[SKT_QRS], e.f[SKT_QGR], e.f[SKT_QGS],
                 e.f[SKT_QGI], e.f[SKT_GROUND_P], e.f[SKT_GROUND_TIME], e.f[SKT_GROUND_T], e.f[SKT_CLOUD_COVER],
                 nullptr, e.fst[SKT_QR], e.fst[SKT_QS], e.fst[SKT_QI], e.fst[SKT_QC], e.fst[SKT_QW]};
    // a recorded ensemble step reads the member's irradiance of each step
    // from its device slot
    a.irrDev = e.rec ? e.rec->irr : nullptr;
    const dim3 grid((unsigned)((e.G + GC_COLS - 1) / GC_COLS)), block(64, 4);
    if (e.mix)
        return klaunch<KGroundColumns<true>>(e, grid, block, 0, geom(e), a, stages, dt, speed, irr, e.dx, e.gm);
    return klaunch<KGroundColumns<false>>(e, grid, block, 0, geom(e), a, stages, dt, speed, irr, e.dx, e.gm);
}

// advectPrecipGPU, Game/kernelSky.cu:576-657. Each level reads its own and
// the next level's pre-step value (the column snapshot of the reference), so
// the update can be in place. A block holds 64 columns; its four waves walk
// y upwards through the four chunks of Y / 4 levels of those columns. Every
// thread reads the pre-step value of the level above its chunk (which the
// chunk above overwrites) before the block's barrier, and writes only after
// it, so the chunks run side by side with the results of one walk per column.
// One walk carries every field of the batch (rain, snow, ice): a level's
// temperature theta (p / p_ground)^kappa, virtual temperature and density
// are computed once and serve each field's fall speed, and the level's
// theta, p and vapour are read once.
constexpr int PRECIP_CHUNKS = 4;
static_assert(64 * PRECIP_CHUNKS <= 256, "k_precip blocks (64 x PRECIP_CHUNKS) must fit the 256-thread launch bound");
template <bool MIX>
struct KPrecip {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        const PrecipBatch& pbt = nthArg<1>(a...);
        for (int f = 0; f < pbt.n; f++)
            if (pbt.dts[f] != 0.0f) return false;
        return true;
    }
    static __device__ __forceinline__ void run(VB vb, const Geom& G, const PrecipBatch& pbt, const float* __restrict__ th, const float* __restrict__ p,
                     const float* __restrict__ qv, const float* __restrict__ gP, float dx, const Gammas& gm) {
        // tiles: column blocks of 64
        const long g = vb.id * 64L + threadIdx.x;
        const bool col = g < (long)G.X * G.Z;
        const long gc = col ? g : 0;
        bool on[3];
        float dts[3];
#pragma unroll
        for (int f = 0; f < 3; f++) {
            on[f] = f < pbt.n && pbt.dts[f] != 0.0f;
            dts[f] = on[f] ? pbt.dts[f] : 0.0f;
        }
        const int x = (int)(gc % G.X), zc = (int)(gc / G.X);
        const float kap = kc::Rsd / kc::Cpd;
        const float gp = gP[gc];
        int y0 = G.gh[gc] + 1;
        if (y0 < 0) y0 = 0;
        const int yc = (G.Y + PRECIP_CHUNKS - 1) / PRECIP_CHUNKS;
        const int ylo = (int)threadIdx.y * yc, yhi = min(G.Y, ylo + yc);
        const int ya = max(ylo, y0);
        const bool work = col && ya < yhi;
        float qTop[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int f = 0; f < 3; f++)
            if (on[f] && work && yhi < G.Y) qTop[f] = ldS<MIX>(pbt.q[f], gidx(G, x, yhi, zc), pbt.st[f]);
        __syncthreads();
        if (!work) return;
        // per level: each field's Q, theta, p, qv; levels y+1 and y+2 are
        // loaded ahead
        struct Lv {
            float q[3], t, pr, v;
        };
        auto densOf = [&](const Lv& L) {
            const float T = L.t * powNear1(L.pr / gp, kap);
            const float Tv = T * (0.608f * L.v + 1);
            return L.pr * 100 / (kc::Rsd * Tv);
        };
        auto fvOf = [&](float q, float den, int f) {
            const int type = pbt.type[f];
            const float3 v3 = d_fallVel(q, q, q, den, type, gm.R48, gm.S425, gm.S425);
            return type == 0 ? v3.x : (type == 1 ? v3.y : v3.z);
        };
        const long i0 = gidx(G, x, ya, zc);
        auto load = [&](int y, Lv& L) {
            if (y <= yhi && y < G.Y) {
                const long j = i0 + (long)(y - ya) * G.X;
#pragma unroll
                for (int f = 0; f < 3; f++)
                    if (on[f]) L.q[f] = y == yhi ? qTop[f] : ldS<MIX>(pbt.q[f], j, pbt.st[f]);
                L.t = th[j];
                L.pr = p[j];
                L.v = qv[j];
            }
        };
        Lv l0{{0.0f, 0.0f, 0.0f}, 0.0f, 0.0f, 0.0f}, l1 = l0, l2 = l0;
        load(ya, l0);
        load(ya + 1, l1);
        float fv[3] = {0.0f, 0.0f, 0.0f};
        {
            const float den = densOf(l0);
#pragma unroll
            for (int f = 0; f < 3; f++)
                if (on[f]) fv[f] = fvOf(l0.q[f], den, f);
        }
        long i = i0;
        for (int y = ya; y < yhi; y++, i += G.X) {
            load(y + 2, l2);
            const bool up = y < G.Y - 1;
            const float denU = up ? densOf(l1) : 0.0f;
#pragma unroll
            for (int f = 0; f < 3; f++) {
                if (!on[f]) continue;
                const float Qc = l0.q[f], Qu = up ? l1.q[f] : 0.0f;
                const float fvu = up ? fvOf(l1.q[f], denU, f) : 0.0f;
                float nq = Qc;
                nq -= dts[f] * fminf((fv[f] / dx) * Qc, Qc);
                if (up) nq += dts[f] * fminf((fvu / dx) * Qu, Qu);
                stS<MIX>(pbt.q[f], i, nq, pbt.st[f]);
                fv[f] = fvu;
            }
            l0 = l1;
            l1 = l2;
        }
    }
};
hipError_t launch_precip_batch(EnvState& e, const PrecipBatch& pbt) {
    if (pbt.n < 1 || pbt.n > 3) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((e.G + 63) / 64)), block(64, PRECIP_CHUNKS);
    if (e.mix)
        return klaunch<KPrecip<true>>(e, grid, block, 0, geom(e), pbt, e.f[SKT_POTTEMP], e.f[SKT_PRESSURE],
                                      e.f[SKT_QV], e.f[SKT_GROUND_P], e.dx, e.gm);
    return klaunch<KPrecip<false>>(e, grid, block, 0, geom(e), pbt, e.f[SKT_POTTEMP], e.f[SKT_PRESSURE],
                                   e.f[SKT_QV], e.f[SKT_GROUND_P], e.dx, e.gm);
}

// -------------------------------------------------------------- buoyancy ---
// buoyancyGPU, Game/kernelSky.cu:1061-1247 (bounds {D,D,D}); the ice test
// 'T < 0' is on Kelvin and never holds (quirk A17).
// A march up the columns: each thread walks a chunk of YC_BUOY levels of its
// (x, z) column. A level's temperature theta (p / p_ground)^kappa and its
// saturation mixing ratio are computed once and carried: the cell below uses
// them as its upper neighbour's, the cell itself as its own, the cell above
// as its lower neighbour's (the same expressions on the same values as the
// per-cell form, so the same bits). With the bounds {D,D,D} a neighbour reads
// its raw value when it is a sky cell and 0 otherwise; the level above a sky
// cell inside the domain is sky. Blocks are 64 (x) x 4 (z) columns of one
// chunk; a chunk that starts above the ground reads the level below it.
constexpr int YC_BUOY = 8;
template <bool MIX>
struct KBuoyancy {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int yc, unsigned nyc, float* __restrict__ v, const float* __restrict__ th,
                       const float* __restrict__ qv, const float* __restrict__ Qr, const float* __restrict__ Qs,
                       const float* __restrict__ Qi, const float* __restrict__ p, const float* __restrict__ gP, float dt,
                                               const Store3& sq) {
        const unsigned lin = vb.id;
        const int chunk = (int)(lin % nyc);
        const unsigned tile = lin / nyc;
        const int gx = (G.X + 63) >> 6;
        const int x = (int)(tile % (unsigned)gx) * 64 + (int)threadIdx.x;
        const int z = (int)(tile / (unsigned)gx) * 4 + (int)threadIdx.y;
        if (x >= G.X || z >= G.Z) return;
        const int X = G.X, Y = G.Y;
        const long col = (long)X * Y * z + x;  // cell (x, 0, z)
        const long gc = x + (long)X * z;
        const int gh = G.gh[gc];
        const float gps = gP[gc];
        const float kap = kc::Rsd / kc::Cpd;
        // ground heights of the four horizontal neighbour columns (INT_MAX
        // outside the domain: no level there is sky)
        const int ghL = x > 0 ? G.gh[gc - 1] : INT_MAX, ghR = x < X - 1 ? G.gh[gc + 1] : INT_MAX;
        const int ghB = z > 0 ? G.gh[gc - X] : INT_MAX, ghF = z < G.Z - 1 ? G.gh[gc + X] : INT_MAX;
        const long oR = x < X - 1 ? 1 : 0, oL = x > 0 ? -1 : 0;
        const long oF = z < G.Z - 1 ? (long)X * Y : 0, oB = z > 0 ? -(long)X * Y : 0;
        const int y0 = max(chunk * yc, gh + 1), y1 = min(Y, chunk * yc + yc);
        if (y0 >= y1) return;
        struct Lev {
            float th, p, q, T, ws;
        };
        auto level = [&](int l, Lev& L) __attribute__((always_inline)) {
            const long j = col + (long)X * l;
            L.th = th[j];
            L.p = p[j];
            L.q = qv[j];
            L.T = L.th * powNear1(L.p / gps, kap);
            L.ws = d_ws((L.T - 273.15f), L.p);
        };
        Lev lm{0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, lc = lm, lp = lm;  // levels y - 1, y, y + 1
        if (y0 - 1 >= 0 && y0 - 1 > gh) level(y0 - 1, lm);
        level(y0, lc);
        for (int y = y0; y < y1; y++) {
            const bool up = y < Y - 1;
            if (up) level(y + 1, lp);
            const bool down = y > 0 && y - 1 > gh;
            const long i = col + (long)X * y;
            const bool l = y > ghL, r = y > ghR, bk = y > ghB, fw = y > ghF;
            const int validEnv = int(l) + int(r) + int(bk) + int(fw);
            // the cell's loads come first and unconditionally (the four
            // horizontal neighbours' vapour and theta at clamped in-domain
            // positions, taken below where the neighbour is sky); a cell with
            // no sky neighbour (validEnv 0) computes a value it does not store
            const float cur = v[i];
            const float qR = qv[i + oR], qL = qv[i + oL], qF = qv[i + oF], qB = qv[i + oB];
            const float tR = th[i + oR], tL = th[i + oL], tF = th[i + oF], tB = th[i + oB];
            const float totalQ = ldS<MIX>(Qr, i, sq.s[0]) + ldS<MIX>(Qs, i, sq.s[1]) + ldS<MIX>(Qi, i, sq.s[2]);
            {
                const float psC = lc.p;
                const float psU = up ? lp.p : psC;
                const float psD = down ? lm.p : psC;
                const float QP = lc.q;
                // the mean over the sky neighbours: dividing by 4, 2 or 1 is the
                // exact scaling (the same rounding as the division)
                auto envMean = [&](float sum) {
                    if (validEnv == 3) return sum / 3.0f;
                    return validEnv == 4 ? sum * 0.25f : (validEnv == 2 ? sum * 0.5f : sum);
                };
                const float Qenv = envMean((r ? qR : 0.0f) + (l ? qL : 0.0f) + (fw ? qF : 0.0f) + (bk ? qB : 0.0f));
                const float QenvUp = up ? lp.q : 0.0f;
                const float QenvDown = down ? lm.q : 0.0f;
                const float thD = down ? lm.th : 0.0f, thU = up ? lp.th : 0.0f;
                const float T = lc.T;
                float Tad = lc.th, TadU = lc.th, TadD = lc.th;
                const float Tenv = envMean((r ? tR : 0.0f) + (l ? tL : 0.0f) + (fw ? tF : 0.0f) + (bk ? tB : 0.0f));
                const float TenvUp = thU, TenvDown = thD;
                // lower and upper neighbours: TDown = thD (psD / gps)^kappa and TUp
                // are their carried temperatures; both adiabats take the moist
                // lapse rate of this cell, computed once
                const bool adD = down && ((T < 0.0f && QP >= d_wi((lm.T - 273.15f), psD)) || QP >= lm.ws);
                const bool adU = up && ((T < 0.0f && QP >= d_wi((lp.T - 273.15f), psU)) || QP >= lp.ws);
                if (adD || adU) {
                    const float mlr = d_MLR(T - 273.15f, psC);
                    if (adD) {
                        TadD = T - mlr * (psC - psD);
                        TadD = TadD * powNear1(gps / psD, kap);
                    }
                    if (adU) {
                        TadU = mlr * (psU - psC) + T;
                        TadU = TadU * powNear1(gps / psU, kap);
                    }
                }
                float B = 0.0f, BU = 0.0f, BD = 0.0f;
                if (down) {
                    const float VTP = TadD * (1.0f + 0.608f * QP), VTE = TenvDown * (1.0f + 0.608f * QenvDown);
                    BD = kc::g * ((VTP - VTE) / VTE) - totalQ * kc::g;
                }
                {
                    const float VTP = Tad * (1.0f + 0.608f * QP), VTE = Tenv * (1.0f + 0.608f * Qenv);
                    B = kc::g * ((VTP - VTE) / VTE) - totalQ * kc::g;
                }
                if (up) {
                    const float VTP = TadU * (1.0f + 0.608f * QP), VTE = TenvUp * (1.0f + 0.608f * QenvUp);
                    BU = kc::g * ((VTP - VTE) / VTE) - totalQ * kc::g;
                }
                float bf = B;
                if (cur > 0.0f) bf = BU;
                else if (cur < 0.0f) bf = BD;
                float change = bf * dt;
                if ((up && cur > 0.0f && -change > cur) || (down && cur < 0.0f && -change < cur)) change = -cur;
                if (validEnv > 0) v[i] = cur + change;
            }
            lm = lc;
            lc = lp;
        }
    }
};
hipError_t launch_buoyancy(EnvState& e, float dt) {
    const unsigned nyc = (unsigned)((e.Y + YC_BUOY - 1) / YC_BUOY);
    const unsigned nb = (unsigned)((e.X + 63) / 64) * (unsigned)((e.Z + 3) / 4) * nyc;
    const Store3 sq{{e.fst[SKT_QR], e.fst[SKT_QS], e.fst[SKT_QI]}};
    if (e.mix)
        return klaunch<KBuoyancy<true>>(e, dim3(nb), B3, 0, geom(e), YC_BUOY, nyc, e.f[SKT_VELY], e.f[SKT_POTTEMP],
                                        e.f[SKT_QV], e.f[SKT_QR], e.f[SKT_QS], e.f[SKT_QI], e.f[SKT_PRESSURE],
                                        e.f[SKT_GROUND_P], dt, sq);
    return klaunch<KBuoyancy<false>>(e, dim3(nb), B3, 0, geom(e), YC_BUOY, nyc, e.f[SKT_VELY], e.f[SKT_POTTEMP],
                                     e.f[SKT_QV], e.f[SKT_QR], e.f[SKT_QS], e.f[SKT_QI], e.f[SKT_PRESSURE],
                                     e.f[SKT_GROUND_P], dt, sq);
}

// ------------------------------------------------------------------- EOS ---
// initDensity, Game/kernelSky.cu:1432-1452 (+ the old-density copy). Each
// thread takes ZPT cells along z; their loads are unconditional (a ground
// cell's values are read and dropped), so they issue together.
struct KInitDensity {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, float* __restrict__ rho, float* __restrict__ rho_old, const float* __restrict__ th,
                           const float* __restrict__ p, const float* __restrict__ qv, const float* __restrict__ gP) {
        constexpr int zc = ZPT;
        ZM_SETUP_V;
        (void)bid;
        if (!inb || z0 >= z1) return;
        float t[ZPT], pr[ZPT], q[ZPT], g[ZPT];
        int gh[ZPT];
#pragma unroll
        for (int k = 0; k < ZPT; k++) {
            const int z = min(z0 + k, z1 - 1);
            const long i = gidx(G, x, y, z);
            t[k] = th[i];
            pr[k] = p[i];
            q[k] = qv[i];
            g[k] = gP[x + (long)G.X * z];
            gh[k] = G.gh[x + G.X * z];
        }
#pragma unroll
        for (int k = 0; k < ZPT; k++) {
            const int z = z0 + k;
            if (z >= z1 || y <= gh[k]) continue;
            const long i = gidx(G, x, y, z);
            const float T = t[k] * powNear1(pr[k] / g[k], kc::Rsd / kc::Cpd);
            const float Tv = T * (0.608f * q[k] + 1);
            const float d = pr[k] * 100 / (kc::Rsd * Tv);
            rho[i] = d;
            if (rho_old) rho_old[i] = d;
        }
    }
};
hipError_t launch_init_density(EnvState& e, int copy_old) {
    return klaunch<KInitDensity>(e, dim3(gridZpt(e)), B3, 0, geom(e), e.f[SKT_DENSITY],
                                 copy_old ? e.f[SKT_DENSITY_OLD] : nullptr, e.f[SKT_POTTEMP], e.f[SKT_PRESSURE],
                                 e.f[SKT_QV], e.f[SKT_GROUND_P]);
}
// calculateNewPressure, Game/kernelSky.cu:1454-1474
struct KNewPressure {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, float* __restrict__ p, const float* __restrict__ rho, const float* __restrict__ th,
                           const float* __restrict__ qv, const float* __restrict__ gP) {
        CELL_SETUP_V;
        (void)bid;
        if (!inb || isGround(G, x, y, z)) return;
        const long i = gidx(G, x, y, z);
        const float T = th[i] * powNear1(p[i] / gP[x + (long)G.X * z], kc::Rsd / kc::Cpd);
        const float Tv = T * (0.608f * qv[i] + 1);
        p[i] = rho[i] * kc::Rsd * Tv / 100.0f;
    }
};
hipError_t launch_new_pressure(EnvState& e) {
    return klaunch<KNewPressure>(e, grid3(e), B3, 0, geom(e), e.f[SKT_PRESSURE], e.f[SKT_DENSITY], e.f[SKT_POTTEMP],
                                 e.f[SKT_QV], e.f[SKT_GROUND_P]);
}

// resetVelPressProj, Game/kernelSky.cu:1648-1682. The reference walks each
// (x, y) column along z and stops at its first ground cell (quirk A16), so a
// cell is reset iff no cell of its column at or before its plane is ground:
// y > ghz, the largest ground height over planes 0..z of its (x, z) column.
// One thread per cell; the resets of different cells are independent.
// (each thread takes ZPT cells along z)
struct KResetVel {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, const int* __restrict__ ghz, float* __restrict__ u, float* __restrict__ v,
                        float* __restrict__ w, const float* __restrict__ defU, const float* __restrict__ defW) {
        constexpr int zc = ZPT;
        ZM_SETUP_V;
        (void)bid;
        if (!inb) return;
        for (int z = z0; z < z1; z++) {
            if (y <= ghz[x + (long)G.X * z]) continue;
            const long i = gidx(G, x, y, z);
            if (y == G.Y - 1 || isGround(G, x, y + 1, z)) v[i] = 0.0f;
            if (x == G.X - 1) u[i] = defU[y];
            if (x < G.X - 1 && isGround(G, x + 1, y, z)) u[i] = 0.0f;
            if (z == G.Z - 1) w[i] = defW[y];
            if (z < G.Z - 1 && isGround(G, x, y, z + 1)) w[i] = 0.0f;
        }
    }
};
hipError_t launch_reset_vel(EnvState& e) {
    return klaunch<KResetVel>(e, dim3(gridZpt(e)), B3, 0, geom(e), e.ghz, e.f[SKT_VELX], e.f[SKT_VELY], e.f[SKT_VELZ],
                              e.f[SKT_DEFAULT_VELX], e.f[SKT_DEFAULT_VELZ]);
}

// setToDefault (Game/kernelSky.cu:1714-1728) with a per-level profile: the
// reference walks each (x, y) column along z and returns at its first ground
// cell, so a cell takes prof[y] iff y > ghz (the walk reaches it, as in
// k_reset_vel). Without a profile, setToValue(0) (Game/kernelMaths.cu:12-23)
// writes every cell, ground included.
struct KResetParam {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, const int* __restrict__ ghz, float* __restrict__ a,
                                               const float* __restrict__ prof, Store st) {
        CELL_SETUP_V;
        (void)bid;
        if (!inb) return;
        const long i = gidx(G, x, y, z);
        if (!prof) stS<true>(a, i, 0.0f, st);
        else if (y > ghz[x + (long)G.X * z]) stS<true>(a, i, prof[y], st);
    }
};
hipError_t launch_reset_param(EnvState& e, int field, const float* prof) {
    return klaunch<KResetParam>(e, grid3(e), B3, 0, geom(e), e.ghz, e.f[field], prof, e.fst[field]);
}

// ---------------------------------------------------------- microphysics ---
struct MicroArgs {
    float *Qv, *Qw, *Qc, *Qr, *Qs, *Qi, *th;
    const float *rho, *p, *gP;
    uint8_t* phase;
    double* region;  // [33 * nblocks] partials or null
    int rmin[3], rmax[3];
    Store sQw, sQc, sQr, sQs, sQi;  // storage of the hydrometeor fields
};

// calculateEnvMicroPhysicsGPU (Game/microPhysics.cu:1035-1417) fused with
// addHeatGPU (Game/kernelSky.cu:1249-1282): the condens/depos/freeze arrays
// the reference passes between them are registers here.
template <bool MIX>
struct KMicrophysics {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, const MicroArgs& a, float dt, float speed, const Gammas& gm) {
        CELL_SETUP_V;
        __shared__ double shr[4];
        float rv[33];
        bool inRegion = false;
        if (inb && !isGround(G, x, y, z)) {
            const long i = gidx(G, x, y, z);
            const float qv0 = a.Qv[i], qw0 = ldS<MIX>(a.Qw, i, a.sQw), qc0 = ldS<MIX>(a.Qc, i, a.sQc);
            const float qr0 = ldS<MIX>(a.Qr, i, a.sQr), qs0 = ldS<MIX>(a.Qs, i, a.sQs);
            const float qi0 = ldS<MIX>(a.Qi, i, a.sQi);
            const float Qv = qv0 > 1e-18f ? qv0 : 0.0f;
            const float Qw = qw0 > 1e-18f ? qw0 : 0.0f;
            const float Qc = qc0 > 1e-18f ? qc0 : 0.0f;
            const float Qr = qr0 > 1e-18f ? qr0 : 0.0f;
            const float Qs = qs0 > 1e-18f ? qs0 : 0.0f;
            const float Qi = qi0 > 1e-18f ? qi0 : 0.0f;
            const float QwMin = 0.001f, QcMin = 0.001f, QiMin = 0.0006f;
            const float ps = a.p[i];
            float3 fv = make_float3(0.0f, 0.0f, 0.0f);
            const float Dair = a.rho[i];
            const float tK = a.th[i] * powNear1(ps / a.gP[x + (long)G.X * z], kc::Rsd / kc::Cpd);
            const float T = tK - 273.15f;
            if (Qr > 0.0f || Qs > 0.0f || Qi > 0.0f) fv = d_fallVel(Qr, Qs, Qi, Dair, 3, gm.R48, gm.S425, gm.S425);
            const float QWS = d_ws(T, ps);
            const float QWI = d_wi(T, ps);
            const float PT1 = T < 0.0f && Qw + Qc > 0 + 1e-6f ? 1.0f : 0.0f;
            const float PT2 = T < 0.0f && Qr < 1e-4f && Qs < 1e-4f ? 1.0f : 0.0f;
            const float PT3 = T < 0.0f && Qr < 1e-4f ? 1.0f : 0.0f;
            // Marshall-Palmer slopes. Every rate that takes the rain or ice
            // slope is zero unless that species is present (Qr > 0 / Qi > 0),
            // and the snow slope is taken without snow only by the deposition
            // of the cold branch (PSDEP), so a slope is evaluated only where a
            // rate can use it (the rates' values are unchanged)
            float sR = 0.0f, sS = 0.0f, sI = 0.0f;
            if (Qr > 0.0f) sR = d_slope(Dair, Qr, 0);
            if (T < 0.0f || Qs > 0.0f) sS = d_slope(Dair, Qs, 1);
            if (Qi > 0.0f) sI = d_slope(Dair, Qi, 2);
            const float dts = dt * speed;
            float PVCON = mp::PVCON(T, ps, Qv, QWS, dt, speed);
            float PVDEP = mp::PVDEP(T, ps, Qv, QWI, dt, speed);
            PVDEP = fminf(PVDEP, 0.0f);
            float PIMLT = mp::PIMLT(T, Qc);
            float PIDW = mp::PIDW(dt, T, Qc, Qw, Dair, ps);
            float PIHOM = mp::PIHOM(T, Qw);
            float PIACR = mp::PIACR(gm, Qr, Qc, Dair, sR);
            float PRACI = mp::PRACI(gm, Qr, Qc, Dair, sR);
            float PRAUT = mp::PRAUT(Qw, QwMin, T);
            float PRACW = mp::PRACW(gm, Qr, Qw, Dair, sR);
            float PREVP = mp::PREVP(gm, T, Qr, Qv, QWS, Dair, ps, sR, PT1);
            float PRACS = mp::PRACS(Qr, Qs, PT2, fv, Dair, sS, sR);
            float PSACW = mp::PSACW(gm, Qs, Qw, Dair, sS);
            float PSACR = mp::PSACR(Qr, Qs, fv, Dair, sR, sS);
            float PSACI = mp::PSACI(gm, T, Qs, Qc, Dair, sS);
            float PSAUT = mp::PSAUT(T, Qc, QcMin);
            float PSFW = mp::PSFW(dt, T, ps, Qs, Qw, Qc, Qv, QWI, Dair);
            float PSFI = mp::PSFI(T, ps, Qs, Qc, Qv, QWI);
            float PSDEP = mp::PSDEP(gm, T, ps, Qv, QWI, Dair, sS);
            float PSSUB = fmaxf(0.0f, -1 * PSDEP);
            PSDEP = fmaxf(0.0f, PSDEP);
            float PSMLT = mp::PSMLT(gm, T, ps, Qs, Qv, Dair, sS, PSACW, PSACR);
            float PGAUT = mp::PGAUT(T, Qs, QiMin);
            float PGFR = mp::PGFR(T, Qr, Dair, sR);
            float PGACW = mp::PGACW(gm, Qi, Qw, sI, Dair);
            float PGACI = mp::PGACI(gm, Qi, Qc, sI, Dair);
            float PGACR = mp::PGACR(Qi, Qr, Dair, fv, sR, sI);
            float PGACS = mp::PGACS(T, Qs, Qi, Dair, fv, sS, sI, false);
            float PGSUB = mp::PGSUB(gm, T, ps, Qv, QWI, Qi, Dair, sI);
            float PGMLT = mp::PGMLT(gm, T, ps, Qw, Qr, Qi, Qv, Dair, sI, PGACW, PGACR, dts);
            float PGDRY = mp::PGDRY(Qw, Qc, Qr, Qs, PGACW, PGACI, PGACR, PGACS, dts);
            float PGWET = mp::PGWET(gm, T, ps, Qc, Qs, Qi, Qv, fv, Dair, sS, sI, PGACI, PGACS, dts);
            float PGACR1 = mp::PGACR1(T, Qw, Qc, Qs, Qi, Dair, fv, sS, sI, PGWET, PGACW, PGACI, PGACS, dts);
            float W = 0.0f;
            if (PGDRY > PGWET) W = 1.0f;
            float cond = 0.0f, dep = 0.0f, frz = 0.0f;
            if (PVCON >= 0) { PVCON = dt * fminf(Qv, speed * PVCON); cond += PVCON; }
            else if (PVCON < 0) { PVCON = dt * fmaxf(-Qw, speed * PVCON); cond += PVCON; }
            if (PVDEP >= 0) { PVDEP = dt * fminf(Qv, speed * PVDEP); dep += PVDEP; }
            else if (PVDEP < 0) { PVDEP = dt * fmaxf(-Qc, speed * PVDEP); dep += PVDEP; }
            float nQv = qv0, nQw = qw0, nQc = qc0, nQr = qr0, nQs = qs0, nQi = qi0;
            if (T < 0) {
                PIDW = dt * fminf(Qw, speed * PIDW); dep += PIDW;
                PIHOM = dt * fminf(Qw, speed * PIHOM); frz += PIHOM;
                PIACR = dt * fminf(Qr, speed * PIACR); frz += PIACR;
                PRACI = dt * fminf(Qc, speed * PRACI);
                PRAUT = dt * fminf(Qw, speed * PRAUT);
                PRACW = dt * fminf(Qw, speed * PRACW);
                PREVP = dt * fminf(Qr, speed * PREVP * (1 - PT1)); cond -= PREVP;
                PRACS = dt * fminf(Qs, speed * PRACS * (1 - PT2));
                PSACR = dt * fminf(Qr, speed * PSACR); frz += PSACR;
                PSACW = dt * fminf(Qw, speed * PSACW); frz += PSACW;
                PSACI = dt * fminf(Qc, speed * PSACI);
                PSAUT = dt * fminf(Qc, speed * PSAUT);
                PSFW = dt * fminf(Qw, speed * PSFW); frz += PSFW;
                PSFI = dt * fminf(Qc, speed * PSFI);
                PSDEP = dt * fminf(Qv, speed * PSDEP * (PT1)); dep += PSDEP;
                PSSUB = dt * fminf(Qs, speed * PSSUB * (1 - PT1)); dep -= PSSUB;
                PGAUT = dt * fminf(Qs, speed * PGAUT);
                PGFR = dt * fminf(Qr, speed * PGFR); frz += PGFR;
                PGACI = dt * fminf(Qc, speed * PGACI);
                PGACR = dt * fminf(Qr, speed * PGACR * (1 - W)); frz += PGACR;
                PGACS = dt * fminf(Qs, speed * PGACS);
                PGSUB = dt * fminf(Qi, speed * PGSUB * (1 - PT1)); dep -= PGSUB;
                if (W == 1.0f) {
                    if (PGACR1 >= 0) { PGACR1 = dt * fminf(Qr, speed * PGACR1); frz += PGACR1; }
                    else if (PGACR1 < 0) { PGACR1 = dt * fmaxf(-Qi, speed * PGACR1); frz += PGACR1; }
                }
                nQv += PSSUB * (1 - PT1) + PGSUB * (1 - PT1) + PREVP * (1 - PT1) - PSDEP * (PT1)-PVCON - PVDEP;
                nQw += PVCON - PSACW - PSFW - PRAUT - PRACW - PIDW - PIHOM;
                nQc += PVDEP + PIDW + PIHOM - PSAUT - PSACI - PRACI - PSFI - PGACI;
                nQr += PRAUT + PRACW - PIACR - PSACR - PGACR * (1 - W) - PGACR1 * (W)-PGFR - PREVP * (1 - PT1);
                nQs += PSAUT + PSACI + PSACW + PSFW + PSFI + PRACI * (PT3) + PIACR * (PT3) - PGACS - PGAUT -
                       PRACS * (1 - PT2) + PSACR * (PT2) - PSSUB * (1 - PT1) + PSDEP * (PT1);
                nQi += PGAUT + PGFR + PGDRY * (1 - W) + PGWET * (W) + PSACR * (1 - PT2) + PRACS * (1 - PT2) +
                       PRACI * (1 - PT3) + PIACR * (1 - PT3) - PGSUB * (1 - PT1);
            } else {
                PIMLT = dt * fminf(Qc, speed * PIMLT); frz -= PIMLT;
                PIACR = 0.0f;
                PRAUT = dt * fminf(Qw, speed * PRAUT);
                PRACW = dt * fminf(Qw, speed * PRACW);
                PREVP = dt * fminf(Qr, speed * PREVP * (1 - PT1)); cond -= PREVP;
                PSACW = dt * fminf(Qw, speed * PSACW);
                PSMLT = dt * fminf(Qs, speed * PSMLT); frz -= PSMLT;
                PGACW = dt * fminf(Qw, speed * PGACW); frz += PGACW;
                PGACS = dt * fminf(Qs, speed * PGACS);
                PGMLT = dt * fminf(Qi, speed * PGMLT); frz -= PGMLT;
                nQv += PREVP * (1 - PT1) - PVCON - PVDEP;
                nQw += PVCON + PIMLT - PRAUT - PRACW - PGACW - PSACW;
                nQc += PVDEP - PIMLT;
                nQr += PRAUT + PRACW + PSACW + PGACW + PGMLT + PSMLT - PREVP * (1 - PT1);
                nQs += -PSMLT - PGACS;
                nQi += -PGMLT + PGACS;
            }
            a.Qv[i] = nQv;
            stS<MIX>(a.Qw, i, nQw, a.sQw);
            stS<MIX>(a.Qc, i, nQc, a.sQc);
            stS<MIX>(a.Qr, i, nQr, a.sQr);
            stS<MIX>(a.Qs, i, nQs, a.sQs);
            stS<MIX>(a.Qi, i, nQi, a.sQi);
            // addHeatGPU with the updated vapour
            const float Mair = 0.02896f, Mwater = 0.01802f;
            const float XV = (nQv / Mwater) / ((nQv / Mwater) + (1 - nQv) / Mair);
            const float Mth = XV * Mwater + (1 - XV) * Mair;
            const float YV = XV * (Mwater / Mth);
            const float yth = YV * 1.33f + (1 - YV) * 1.4f;
            const float cpth = yth * kc::R / (Mth * (yth - 1));
            const float th = a.th[i];
            float sum = 0.0f;
            sum += d_Lwater(th - 273.15f) / cpth * cond;
            sum += d_Lice(th - 273.15f) / cpth * dep;
            sum += kc::Lf / cpth * frz;
            a.th[i] = th + sum;
            if (a.phase)
                a.phase[i] = (uint8_t)(1 | (T < 0.0f ? 2 : 0) | (PT1 != 0.0f ? 4 : 0) | (PT2 != 0.0f ? 8 : 0) |
                                       (PT3 != 0.0f ? 16 : 0) | (W != 0.0f ? 32 : 0));
            if (a.region && x >= a.rmin[0] && x <= a.rmax[0] && y >= a.rmin[1] && y <= a.rmax[1] && z >= a.rmin[2] &&
                z <= a.rmax[2]) {
                inRegion = true;
                float PVVAP = 0.0f, PVSUB = 0.0f, pc = PVCON, pd = PVDEP;
                if (pc < 0.0f) { PVVAP = -pc; pc = 0.0f; }
                if (pd < 0.0f) { PVSUB = -pd; pd = 0.0f; }
                const float vals[33] = {pc, pd, PIMLT, PIDW, PIHOM, PIACR, PRACI, PRAUT, PRACW, PREVP, PRACS,
                                        PSACW, PSACR, PSACI, PSAUT, PSFW, PSFI, PSDEP, PSSUB, PSMLT, PGAUT, PGFR,
                                        PGACW, PGACI, PGACR, PGDRY * (1 - W), PGACS, PGSUB, PGMLT, PGWET * W, PGACR1, PVVAP, PVSUB};
                for (int k = 0; k < 33; k++) rv[k] = vals[k];
            }
        } else if (inb) {
            if (a.phase) a.phase[gidx(G, x, y, z)] = 0;
        }
        if (a.region) {
            const unsigned nb = vb.n;
            for (int k = 0; k < 33; k++) {
                const double s = blockSum(inRegion ? (double)rv[k] : 0.0, shr);
                if (tid256() == 0) a.region[(long)k * nb + bid] = s;
            }
        }
    }
};
hipError_t launch_microphysics(EnvState& e, float dt, float speed) {
    MicroArgs a{e.f[SKT_QV], e.f[SKT_QW], e.f[SKT_QC], e.f[SKT_QR], e.f[SKT_QS], e.f[SKT_QI], e.f[SKT_POTTEMP],
                e.f[SKT_DENSITY], e.f[SKT_PRESSURE], e.f[SKT_GROUND_P], e.phase,
                e.region_active ? e.region : nullptr, {e.rmin[0], e.rmin[1], e.rmin[2]},
                {e.rmax[0], e.rmax[1], e.rmax[2]}, e.fst[SKT_QW], e.fst[SKT_QC], e.fst[SKT_QR], e.fst[SKT_QS],
                e.fst[SKT_QI]};
    if (e.mix) return klaunch<KMicrophysics<true>>(e, grid3(e), B3, 0, geom(e), a, dt, speed, e.gm);
    return klaunch<KMicrophysics<false>>(e, grid3(e), B3, 0, geom(e), a, dt, speed, e.gm);
}
// Fixed-order sum of the region partials into 33 floats at the tail of the
// region buffer.
struct KRegionFinalize {
    static __device__ __forceinline__ void run(VB vb, double* region, unsigned nb, float* out) {
        __shared__ double sh[4];
        for (int k = 0; k < 33; k++) {
            double v = 0.0;
            for (unsigned j = tid256(); j < nb; j += 256) v += region[(long)k * nb + j];
            v = blockSum(v, sh);
            if (tid256() == 0) out[k] = (float)v;
        }
    }
};
hipError_t launch_region_finalize(EnvState& e) {
    const unsigned nb = (unsigned)blocks3d(e);
    float* out = (float*)(e.region + 33L * nb);
    return klaunch<KRegionFinalize>(e, dim3(1), dim3(64, 4, 1), 0, e.region, nb, out);
}

// ------------------------------------------------------ masks / editor ---
// computeNeighbourGPU, Game/kernelSky.cu:1284-1316 (12 bytes per cell).
KL void k_masks(Geom G, uint8_t* out) {
    CELL_SETUP;
    (void)bid;
    if (!inb) return;
    uint8_t* m = out + 12 * gidx(G, x, y, z);
    const bool cg = isGround(G, x, y, z);
    bool o;
    o = x == 0; m[0] = o; m[1] = (o && cg) || (!o && isGround(G, x - 1, y, z));
    o = x == G.X - 1; m[2] = o; m[3] = (o && cg) || (!o && isGround(G, x + 1, y, z));
    o = y == G.Y - 1; m[4] = o; m[5] = !o && isGround(G, x, y + 1, z);
    o = y == 0; m[6] = o; m[7] = (o && cg) || (!o && isGround(G, x, y - 1, z));
    o = z == G.Z - 1; m[8] = o; m[9] = (o && cg) || (!o && isGround(G, x, y, z + 1));
    o = z == 0; m[10] = o; m[11] = (o && cg) || (!o && isGround(G, x, y, z - 1));
}
hipError_t launch_masks(EnvState& e, uint8_t* dev_out) {
    hipLaunchKernelGGL(k_masks, grid3(e), B3, 0, e.stream, geom(e), dev_out);
    return hipGetLastError();
}
struct Ten {
    float* a[10];
    Store st[10];
};
// resetValueInGround over n fields (Game/kernelGround.cu:21-36)
struct KResetGround {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, const Ten& t, int n) {
        CELL_SETUP_V;
        (void)bid;
        if (!inb || !isGround(G, x, y, z)) return;
        const long i = gidx(G, x, y, z);
        for (int k = 0; k < n; k++) stS<true>(t.a[k], i, 0.0f, t.st[k]);
    }
};
hipError_t launch_editor_reset(EnvState& e, int nfields) {
    static const int ids[10] = {SKT_POTTEMP, SKT_QV,   SKT_QW,   SKT_QC,   SKT_QR,
                                SKT_QS,      SKT_QI,   SKT_VELX, SKT_VELY, SKT_VELZ};
    Ten t;
    for (int k = 0; k < 10; k++) {
        t.a[k] = e.f[ids[k]];
        t.st[k] = e.fst[ids[k]];
    }
    return klaunch<KResetGround>(e, grid3(e), B3, 0, geom(e), t, nfields);
}
// computeIsenTempGroundGPU, Game/kernelGround.cu:38-47
struct KIsenTempGround {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, float* gT, const float* isen, const float* gP, const float* p) {
        const long g = vb.id * 256L + threadIdx.x;
        if (g >= (long)G.X * G.Z) return;
        const int x = (int)(g % G.X), zc = (int)(g / G.X);
        const int GH = G.gh[g] + 1 >= G.Y ? G.Y - 1 : G.gh[g] + 1;
        const float T = d_potTemp(isen[GH - 1] - 273.15f, gP[g], p[gidx(G, x, GH, zc)]);
        gT[g] = T + 273.15f;
    }
};
hipError_t launch_isen_temp_ground(EnvState& e) {
    return klaunch<KIsenTempGround>(e, grid1(e.G), dim3(256), 0, geom(e), e.f[SKT_GROUND_T], e.f[SKT_ISEN_TEMP],
                                    e.f[SKT_GROUND_P], e.f[SKT_PRESSURE]);
}
// Default profiles = the (x=0, z=0) column before the init noise
// (cudaMemcpy2D of Game/environment.cu:299-303).
KL void k_profiles(Geom G, const float* th, const float* qv, const float* u, const float* w, float* iT, float* iQ,
                   float* dU, float* dW) {
    const int y = blockIdx.x * 256 + threadIdx.x;
    if (y >= G.Y) return;
    const long i = (long)G.X * y;
    iT[y] = th[i];
    iQ[y] = qv[i];
    dU[y] = u[i];
    dW[y] = w[i];
}
hipError_t launch_profiles_from_column(EnvState& e) {
    hipLaunchKernelGGL(k_profiles, grid1(e.Y), dim3(256), 0, e.stream, geom(e), e.f[SKT_POTTEMP], e.f[SKT_QV],
                       e.f[SKT_VELX], e.f[SKT_VELZ], e.f[SKT_ISEN_TEMP], e.f[SKT_ISEN_VAPOR], e.f[SKT_DEFAULT_VELX],
                       e.f[SKT_DEFAULT_VELZ]);
    return hipGetLastError();
}
// randomArray (Game/kernelMaths.cu:93-116) with the seed every thread intends
// to read (quirk A10); the hash is the published PCG-RXS-M-XS function.
KL void k_random(float* a, long n, float mn, float mx, uint32_t seed) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const uint32_t sd = seed + (uint32_t)i;
        const uint32_t s = sd * 747796405u + 2891336453u;
        const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
        const uint32_t h = (w >> 22u) ^ w;
        const float rnd = float(h % 1000u) / float(1000);
        a[i] += rnd * (mx - mn) + mn;
    }
}
hipError_t launch_random(EnvState& e, float* a, float mn, float mx, uint32_t seed) {
    hipLaunchKernelGGL(k_random, dim3((unsigned)std::min<long>((e.N + 255) / 256, 65536)), dim3(256), 0, e.stream, a,
                       e.N, mn, mx, seed);
    return hipGetLastError();
}

// Broadcast of a sounding's per-level initial state over every cell
// (Game/readTable.cpp:404-441,467-471). lv holds ny values each of theta,
// face u, face w, qv and pressure; the vertical face wind is 0.
KL void k_sounding_fill(long N, int X, int Y, long G, const float* __restrict__ lv, float* th, float* u, float* v,
                        float* w, float* qv, float* p, float* gT, float* gP, float groundT, float groundP) {
    const long stride = (long)gridDim.x * 256;
    for (long i = blockIdx.x * 256L + threadIdx.x; i < N; i += stride) {
        const int y = (int)((i / X) % Y);
        th[i] = lv[y];
        u[i] = lv[Y + y];
        v[i] = 0.0f;
        w[i] = lv[2 * Y + y];
        qv[i] = lv[3 * Y + y];
        p[i] = lv[4 * Y + y];
    }
    for (long i = blockIdx.x * 256L + threadIdx.x; i < G; i += stride) {
        gT[i] = groundT;
        gP[i] = groundP;
    }
}
hipError_t launch_sounding_fill(EnvState& e, const float* lv, float groundT, float groundP) {
    hipLaunchKernelGGL(k_sounding_fill, dim3((unsigned)std::min<long>((e.N + 255) / 256, 16384)), dim3(256), 0,
                       e.stream, e.N, e.X, e.Y, e.G, lv, e.f[SKT_POTTEMP], e.f[SKT_VELX], e.f[SKT_VELY],
                       e.f[SKT_VELZ], e.f[SKT_QV], e.f[SKT_PRESSURE], e.f[SKT_GROUND_T], e.f[SKT_GROUND_P], groundT,
                       groundP);
    return hipGetLastError();
}

// ------------------------------------------------------ serving diagnostics ---
// No reference counterpart: these replace the editor's full-state readback
// after every step (Game/editor.cpp:226-269) with device reductions. Per
// block: the largest VELY (the reference's vertical velocity) over its sky
// cells, the highest level whose QW + QC exceeds thr at a sky cell, and, in
// the blocks whose tile holds level 0, the ground rain water QGR of its
// columns (FP64). Each ground column lies in exactly one such block.
template <bool MIX>
struct KDiag {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int zc, const float* __restrict__ v, const float* __restrict__ qw,
                   const float* __restrict__ qc, Store sw, Store sc, const float* __restrict__ qgr, float thr, double* pw, double* pl,
                   double* pr) {
        ZM_SETUP_V;
        __shared__ double sh[4];
        float w = -__builtin_inff();
        int lev = -1;
        double rain = 0.0;
        if (inb) {
            for (int z = z0; z < z1; z++) {
                const long i = gidx(G, x, y, z);
                if (!isGround(G, x, y, z)) {
                    w = fmaxf(w, v[i]);
                    if (ldS<MIX>(qw, i, sw) + ldS<MIX>(qc, i, sc) > thr) lev = y;  // y is fixed per lane
                }
                if (y == 0) rain += (double)qgr[x + (long)G.X * z];
            }
        }
        const double m = blockMax((double)w, sh);
        const double l = blockMax((double)lev, sh);
        const double r = blockSum(rain, sh);
        if (tid256() == 0) {
            pw[bid] = m;
            pl[bid] = l;
            pr[bid] = r;
        }
    }
};
// One block: the final reductions in a fixed order.
struct KDiagFinalize {
    static __device__ __forceinline__ void run(VB vb, const double* __restrict__ pw, const double* __restrict__ pl, const double* __restrict__ pr,
                            unsigned n, Scalars* sc) {
        __shared__ double sh[4];
        double m = -__builtin_inf(), l = -1.0, r = 0.0;
        for (unsigned j = tid256(); j < n; j += 256) {
            m = fmax(m, pw[j]);
            l = fmax(l, pl[j]);
            r += pr[j];
        }
        m = blockMax(m, sh);
        l = blockMax(l, sh);
        r = blockSum(r, sh);
        if (tid256() == 0) {
            sc->dg_w = (float)m;
            sc->dg_level = (int)l;
            sc->dg_rain = r;
        }
    }
};
hipError_t launch_diagnostics(EnvState& e, float thr) {
    const ZGrid zg = zgrid(e);
    const size_t M = e.maxBlocks;
    double *pw = e.partials + PART_DG_W * M, *pl = e.partials + PART_DG_L * M, *pr = e.partials + PART_DG_R * M;
    hipError_t err =
        e.mix ? klaunch<KDiag<true>>(e, zg.grid, B3, 0, geom(e), zg.zc, (const float*)e.f[SKT_VELY],
                                     (const float*)e.f[SKT_QW], (const float*)e.f[SKT_QC], e.fst[SKT_QW],
                                     e.fst[SKT_QC], (const float*)e.f[SKT_QGR], thr, pw, pl, pr)
              : klaunch<KDiag<false>>(e, zg.grid, B3, 0, geom(e), zg.zc, (const float*)e.f[SKT_VELY],
                                      (const float*)e.f[SKT_QW], (const float*)e.f[SKT_QC], e.fst[SKT_QW],
                                      e.fst[SKT_QC], (const float*)e.f[SKT_QGR], thr, pw, pl, pr);
    if (err != hipSuccess) return err;
    return klaunch<KDiagFinalize>(e, dim3(1), B3, 0, pw, pl, pr, zg.nblocks, e.sc);
}

// ---------------------------------------------------------------- terrain ---
// PNoise2D (Engine/source/math/geometry.cpp:196-278) and initGroundHeightGPU
// (Game/kernelGround.cu:13-19): one thread per ground column. The gradient
// angle of a lattice point comes from the C runtime's rand() after
// srand(seed + ix + iy * width); the reference is a Windows (MSBuild) program,
// so that is the MSVC runtime's generator: state = state * 214013 + 2531011,
// value = (state >> 16) & 0x7FFF. Every float expression keeps the reference's
// operand order; glm::normalize is v * (1 / sqrt(dot(v, v))).
__device__ __forceinline__ float2 noiseGradient(int seed, int ix, int iy, int width) {
    const unsigned st = (unsigned)(seed + ix + iy * width) * 214013u + 2531011u;
    const int r = (int)((st >> 16) & 0x7FFFu);
    const float a = float(r % 1000) * 0.001f * 3.14159265f * 2;
    const float vx = sinf(a), vy = cosf(a);
    const float inv = 1.0f / sqrtf(vx * vx + vy * vy);
    return make_float2(vx * inv, vy * inv);
}
__device__ __forceinline__ float noiseDot(int seed, int ix, int iy, float x, float y, int width) {
    const float2 g = noiseGradient(seed, ix, iy, width);
    const float dx = fabsf(x - float(ix)), dy = fabsf(y - float(iy));
    return fabsf(dx * g.x + dy * g.y);
}
__device__ __forceinline__ float noiseInterp(float a0, float a1, float w) {
    return (a1 - a0) * (3.0f - w * 2.0f) * w * w + a0;
}
__device__ float perlin2(int seed, float x, float y, int width) {
    const int x0 = int(x), y0 = int(y), x1 = x0 + 1, y1 = y0 + 1;
    const float sx = x - float(x0), sy = y - float(y0);
    float n0 = noiseDot(seed, x0, y0, x, y, width);
    float n1 = noiseDot(seed, x1, y0, x, y, width);
    const float ix0 = noiseInterp(n0, n1, sx);
    n0 = noiseDot(seed, x0, y1, x, y, width);
    n1 = noiseDot(seed, x1, y1, x, y, width);
    const float ix1 = noiseInterp(n0, n1, sx);
    return noiseInterp(ix0, ix1, sy);
}
KL void k_terrain(int seed, int width, int depth, int octaves, float maxHeight, float* __restrict__ noise,
                  int* __restrict__ heights) {
    const long g = blockIdx.x * 256L + tid256();
    if (g >= (long)width * depth) return;
    const int x = (int)(g % width), z = (int)(g / width);
    const int FREQ0 = width;
    float val = 0.0f, freq = 1.0f, amp = 1.0f;
    for (int i = 0; i < octaves; i++) {
        val += perlin2(seed, x * freq / FREQ0, z * freq / FREQ0, width) * amp;
        freq *= 2.0f;
        amp /= 2.0f;
    }
    noise[g] = val;
    heights[g] = static_cast<int>(roundf(val * maxHeight));
}
hipError_t launch_terrain(int seed, int width, int depth, int octaves, float maxHeight, float* noise, int* heights,
                          hipStream_t stream) {
    const long n = (long)width * depth;
    hipLaunchKernelGGL(k_terrain, dim3((unsigned)((n + 255) / 256)), B3, 0, stream, seed, width, depth, octaves,
                       maxHeight, noise, heights);
    return hipGetLastError();
}

// Device scratch of the allocation-free terrain entry point
// (skt_terrain_generate_noalloc): the noise of up to kTerrainScratchColumns
// columns, then their heights. Each GPU the library runs on holds its own
// copy until the process ends. Host code never passes the array's name to a
// kernel or copy (on the host it is a stand-in, not a device address): it
// uses the address hipGetSymbolAddress gives for the device, looked up once
// per device on its first call with that device current.
constexpr long kTerrainScratchColumns = SKT_TERRAIN_NOALLOC_MAX_COLUMNS;
static __device__ __align__(256) unsigned char scratch[8 * kTerrainScratchColumns];
namespace {
constexpr int kMaxDevices = 64;
std::mutex scratchMu;
void* scratchPtr[kMaxDevices] = {};
}  // namespace
hipError_t terrain_scratch(int device, float** noise, int** heights) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratchMu);
    if (!scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t r = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (r != hipSuccess) return r;
        if (!p) return hipErrorInvalidSymbol;
        scratchPtr[device] = p;
    }
    *noise = static_cast<float*>(scratchPtr[device]);
    *heights = reinterpret_cast<int*>(static_cast<unsigned char*>(scratchPtr[device]) + 4 * kTerrainScratchColumns);
    return hipSuccess;
}

// ------------------------------------------------------------------ edits ---
// setGround (Game/kernelSky.cu:1620-1633): against the current heights
// (G.gh), erasing at a ground cell lowers the edited column to y, raising at a
// sky cell lifts it to y; either reports a change.
__device__ __forceinline__ void setGround(const Geom& G, int* ghEdit, int* changed, int x, int y, int z, int erase) {
    const int g = x + G.X * z;
    if (erase && y <= G.gh[g]) {
        atomicMin(&ghEdit[g], y);
        *changed = 1;
    } else if (!erase && y > G.gh[g]) {
        atomicMax(&ghEdit[g], y);
        *changed = 1;
    }
}
// Adds (brush) or sets (selection) the parameter value at one cell.
__device__ __forceinline__ void applyEditCell(const Geom& G, const EditArgs& a, int* ghEdit, int* changed, int mX,
                                              int mY, int mZ, float v1, bool add) {
    float v2 = 0.0f, v3 = 0.0f;
    if (a.param == 7) {
        v3 = v1 * a.dir[2];
        v2 = v1 * a.dir[1];
        v1 *= a.dir[0];
    }
    const long i = gidx(G, mX, mY, mZ);
    if (a.param == 8) {
        setGround(G, ghEdit, changed, mX, mY, mZ, a.erase);
    } else if (a.param == 7) {
        if (add) {
            a.a[0][i] += v1;
            a.a[1][i] += v2;
            a.a[2][i] += v3;
        } else {
            a.a[0][i] = v1;
            a.a[1][i] = v2;
            a.a[2][i] = v3;
        }
    } else {
        // a hydrometeor may be stored in 16 bits
        stS<true>(a.a[0], i, add ? ldS<true>(a.a[0], i, a.st[0]) + v1 : v1, a.st[0]);
    }
}
// applyBrushGPU (Game/kernelSky.cu:1476-1556): thread (x, y) of the
// n x n square, n = ceil(size) - floor(-size), and every z of the domain; the
// cell at (x, y, z) - ceil(size) from the brush position gets
// (1 - d^2/size^2)^smooth * dt * intensity * value inside the sphere. A
// position beyond the domain is skipped (the reference's z test lets z = nz
// through, one plane past its arrays).
KL void k_apply_brush(Geom G, EditArgs a, int n, int* __restrict__ ghEdit, int* changed) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = blockIdx.z;
    if (x >= n || y >= n) return;
    const int off = int(ceilf(a.size));
    const int lX = x - off, lY = y - off, lZ = z - off;
    const int mX = x + a.pos[0] - off, mY = y + a.pos[1] - off, mZ = z + a.pos[2] - off;
    if (mX >= G.X || mY >= G.Y || mZ >= G.Z || mX < 0 || mY < 0 || mZ < 0) return;
    if (a.param != 8 && mY <= G.gh[mX + G.X * mZ]) return;  // can't brush the ground
    const float eX = lX, eY = lY, eZ = lZ;
    const float distance = eX * eX + eY * eY + eZ * eZ;
    const float radius = a.size * a.size;
    if (distance <= radius) {
        float v1 = -(distance / radius - 1);
        v1 = powf(v1, a.smooth);
        v1 *= a.dt * a.intensity * a.value;
        applyEditCell(G, a, ghEdit, changed, mX, mY, mZ, v1, true);
    }
}
// applySelectionGPU (Game/kernelSky.cu:1558-1618): the box [mn, mx] (an
// empty x or y range still runs one thread at mn) gets the value.
KL void k_apply_selection(Geom G, EditArgs a, int nx, int ny, int* __restrict__ ghEdit, int* changed) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, z = a.mn[2] + (int)blockIdx.z;
    if (x >= nx || y >= ny) return;
    const int mX = x + a.mn[0], mY = y + a.mn[1], mZ = z;
    if (mX >= G.X || mY >= G.Y || mZ >= G.Z || mX < 0 || mY < 0 || mZ < 0) return;
    if (a.param != 8 && mY <= G.gh[mX + G.X * mZ]) return;
    applyEditCell(G, a, ghEdit, changed, mX, mY, mZ, a.value, false);
}
// compareAndResetValuesOutGround (Game/kernelSky.cu:1635-1663): cells that an
// erase turned from ground into sky (new height < y <= old height) get the
// base profiles: vapour and potential temperature of the isentropic state,
// the default pressure, no condensate and no wind.
KL void k_reset_uncovered(Geom G, const int* __restrict__ ghNew, const float* __restrict__ isT,
                          const float* __restrict__ isV, const float* __restrict__ defP, float* Qv, float* Qw,
                          float* Qc, float* Qr, float* Qs, float* Qi, float* th, float* u, float* v, float* w,
                          float* p, Store sw, Store sc, Store sr, Store ss, Store si) {
    CELL_SETUP;
    (void)bid;
    if (!inb) return;
    const int g = x + G.X * z;
    if (y > G.gh[g] || y <= ghNew[g]) return;
    const long i = gidx(G, x, y, z);
    Qv[i] = isV[y];
    stS<true>(Qw, i, 0.0f, sw);
    stS<true>(Qc, i, 0.0f, sc);
    stS<true>(Qr, i, 0.0f, sr);
    stS<true>(Qs, i, 0.0f, ss);
    stS<true>(Qi, i, 0.0f, si);
    th[i] = isT[y];
    u[i] = 0.0f;
    v[i] = 0.0f;
    w[i] = 0.0f;
    p[i] = defP[y];
}
hipError_t launch_apply_brush(EnvState& e, const EditArgs& a) {
    const int n = int(std::ceil(a.size)) - int(std::floor(-a.size));
    if (n <= 0) return hipSuccess;
    const dim3 grid((n + 63) / 64, (n + 3) / 4, e.Z);
    hipLaunchKernelGGL(k_apply_brush, grid, B3, 0, e.stream, geom(e), a, n, e.ghEdit, &e.sc->edit_ground);
    return hipGetLastError();
}
hipError_t launch_apply_selection(EnvState& e, const EditArgs& a) {
    const int nx = std::max(a.mx[0] - a.mn[0] + 1, 1), ny = std::max(a.mx[1] - a.mn[1] + 1, 1);
    const int nz = a.mx[2] - a.mn[2] + 1;
    if (nz <= 0) return hipSuccess;
    const dim3 grid((nx + 63) / 64, (ny + 3) / 4, nz);
    hipLaunchKernelGGL(k_apply_selection, grid, B3, 0, e.stream, geom(e), a, nx, ny, e.ghEdit, &e.sc->edit_ground);
    return hipGetLastError();
}
hipError_t launch_reset_uncovered(EnvState& e) {
    hipLaunchKernelGGL(k_reset_uncovered, grid3(e), B3, 0, e.stream, geom(e), e.ghEdit, e.f[SKT_ISEN_TEMP],
                       e.f[SKT_ISEN_VAPOR], e.f[SKT_DEFAULT_PRESSURE], e.f[SKT_QV], e.f[SKT_QW], e.f[SKT_QC],
                       e.f[SKT_QR], e.f[SKT_QS], e.f[SKT_QI], e.f[SKT_POTTEMP], e.f[SKT_VELX], e.f[SKT_VELY],
                       e.f[SKT_VELZ], e.f[SKT_PRESSURE], e.fst[SKT_QW], e.fst[SKT_QC], e.fst[SKT_QR],
                       e.fst[SKT_QS], e.fst[SKT_QI]);
    return hipGetLastError();
}

}  // namespace skt
