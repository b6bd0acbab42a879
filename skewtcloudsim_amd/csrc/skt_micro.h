// Bulk microphysics process rates for gfx950, following the reference's GPU
// forms in Game/microPhysics.cu:67-1032 (operand order kept term by term).
#pragma once
#include "skt_device.h"

namespace skt {
namespace mp {
using namespace kc;

// Game/microPhysics.cu:67-75
__device__ __forceinline__ float PVCON(float T, float ps, float Qv, float QWS, float dt, float speed) {
    const float ES = d_es(T) * 10.0f;
    const float der = ((E * ps) / ((ps - ES) * (ps - ES))) * (ES * (17.27f * 237.3f / ((T + 237.3f) * (T + 237.3f))));
    return 1 / (dt * speed) * (Qv - QWS) / (1 + E0v / Cpd * der);
}
// :77-86
__device__ __forceinline__ float PVDEP(float T, float ps, float Qv, float QWI, float dt, float speed) {
    const float EI = d_ei(T) * 10.0f;
    const float der = ((E * ps) / ((ps - EI) * (ps - EI))) * (EI * (21.875f * 265.5f / ((T + 265.5f) * (T + 265.5f))));
    return 1 / (dt * speed) * (Qv - QWI) / (1 + Ls / Cpd * der);
}
// :88-97
__device__ __forceinline__ float PIMLT(float T, float Qc) { return T >= 0.0f ? Qc : 0.0f; }
// :99-113
__device__ __forceinline__ float PIDW(float dt, float T, float Qc, float Qw, float Dair, float ps) {
    if (T < 0.0f && Qw > 0.0f) {
        const float a = 0.5f;
        const float quu = fmaxf(1e-12f * d_Ni(T) / (Dair * 0.001f), Qc);
        return powPos((1 - a) * d_cvd(T, ps, Dair * 0.001f) * dt + powPos(quu, 1 - a), 1 / (1 - a)) - Qc;
    }
    return 0.0f;
}
// :115-122
__device__ __forceinline__ float PIHOM(float T, float Qw) { return (T < -40 && Qw > 0.0f) ? Qw : 0.0f; }
// :124-145
__device__ __forceinline__ float PIACR(const Gammas& gm, float Qr, float Qc, float Dair, float sR) {
    if (Qr > 0 && Qc > 0.0f) {
        const float density = sqrtf(1.225f / Dair);
        return (PI * PI * 0.3f * 8e-2f * 2115.0f * Qc * 0.99f * gm.gRC) / (24 * 4.19e-10f * powPos(sR, 6 + 0.8f)) * density;
    }
    return 0.0f;
}
// :147-165
__device__ __forceinline__ float PRACI(const Gammas& gm, float Qr, float Qc, float Dair, float sR) {
    if (Qr > 0 && Qc > 0.0f) {
        const float density = sqrtf(1.225f / Dair);
        return (PI * 0.3f * 8e-2f * 2115.0f * Qc * gm.gR) / (4 * powPos(sR, 3 + 0.8f)) * density;
    }
    return 0.0f;
}
// :167-187 ('4 / 3' is integer division: quirk A5)
__device__ __forceinline__ float PRAUT(float Qw, float QwMin, float T) {
    if (Qw >= QwMin) {
        const float rc = 10e-6f;
        float Nc = (Qw / (4 / 3 * PI * d_pwater(T) * rc * rc * rc)) * 1.225f;
        Nc /= 1e6f;
        const float disE = 0.5f * 0.5f;
        const float dispersion = ((1 + 3 * disE) * (1 + 4 * disE) * (1 + 5 * disE)) / ((1 + disE) * (1 + 2 * disE));
        const float L = Qw * 1.225f * 1e-3f;
        return 1e3f * 1.1e10f * dispersion * (L * L * L) * (1.0f / Nc) *
               static_cast<float>(1 - expf(-(1.03e16f * powPos(Nc, -2.0f / 3.0f) * (L * L))));
    }
    return 0.0f;
}
// :189-208
__device__ __forceinline__ float PRACW(const Gammas& gm, float Qr, float Qw, float Dair, float sR) {
    if (Qr > 0 && Qw > 0.0f) {
        const float density = sqrtf(1.225f / Dair);
        return (PI * 1.0f * 8e-2f * 2115.0f * Qw * gm.gR) / (4 * powPos(sR, 3 + 0.8f)) * density;
    }
    return 0.0f;
}
// :210-236
__device__ __forceinline__ float PREVP(const Gammas& gm, float T, float Qr, float Qv, float QWS, float Dair, float ps,
                                       float sR, float PT1) {
    if (Qr > 0.0f && Qv / QWS < 1.0f && (1 - PT1) != 0.0f) {
        const float tK = T + 273.15f;
        const float S = Qv / QWS;
        const float b = 0.8f, a = 2115.0f;
        const float dQv = d_DQVair(T, ps);
        const float kv = d_ViscAir(T);
        float Sc = kv / dQv;
        const float density = powPos(1.225f / (Dair * 0.001f), 0.25f);
        return fmaxf(0.0f, -1 * 2 * PI * (S - 1) * 8e-2f *
                               (0.78f * powPos(sR, -2) + 0.31f * powPos(Sc, 1.0f / 3.0f) * gm.gER * sqrtf(a) *
                                                           powPos(kv, -0.5f) * density * powPos(sR, -((b + 5) / 2))) *
                               1 / (Dair * 0.001f) *
                               (1.0f / (E0v * E0v / (Ka * Rsw * tK * tK) + (1 / (Dair * 0.001f * QWS * dQv)))));
    }
    return 0.0f;
}
__device__ __forceinline__ float coll(float sA, float sB) {
    return (5 / (powPos(sA, 6) * sB) + (2 / (powPos(sA, 5) * powPos(sB, 2))) + (0.5f / (powPos(sA, 4) * powPos(sB, 3))));
}
// :238-261
__device__ __forceinline__ float PRACS(float Qr, float Qs, float PT2, float3 fv, float Dair, float sS, float sR) {
    if (Qr > 0.0f && Qs > 0.0f && PT2 == 0) {
        const float size = PI * PI * 1.0f * 8e-2f * 3e-2f;
        return size * fabsf(fv.x - fv.y) * (0.11f / (Dair * 0.001f)) * coll(sS, sR);
    }
    return 0.0f;
}
// :263-281
__device__ __forceinline__ float PSACW(const Gammas& gm, float Qs, float Qw, float Dair, float sS) {
    if (Qs > 0 && Qw > 0) {
        const float density = sqrtf(1.225f / Dair);
        return (PI * 1.0f * 3e-2f * 152.93f * Qw * gm.gS) / (4 * powPos(sS, 3 + 0.25f)) * density;
    }
    return 0.0f;
}
// :283-305
__device__ __forceinline__ float PSACR(float Qr, float Qs, float3 fv, float Dair, float sR, float sS) {
    if (Qr > 0.0f && Qs > 0.0f) {
        const float size = PI * PI * 1.0f * 8e-2f * 3e-2f;
        return size * fabsf(fv.y - fv.x) * (0.99f / (Dair * 0.001f)) * coll(sR, sS);
    }
    return 0.0f;
}
// :307-325
__device__ __forceinline__ float PSACI(const Gammas& gm, float T, float Qs, float Qc, float Dair, float sS) {
    if (Qs > 0 && Qc > 0) {
        const float ESI = expf(0.025f * (T));
        const float density = sqrtf(1.225f / Dair);
        return (PI * ESI * 3e-2f * 152.93f * Qc * gm.gS) / (4 * powPos(sS, 3 + 0.25f)) * density;
    }
    return 0.0f;
}
// :327-334
__device__ __forceinline__ float PSAUT(float T, float Qc, float QcMin) {
    if (Qc - QcMin > 0.0f && T < 0.0f) return 1e-3f * expf(0.025f * (T)) * (Qc - QcMin);
    return 0.0f;
}
// Bergeron coefficients shared by PSFW / PSFI (:351-359, :383-390)
__device__ __forceinline__ void bergeron(float T, float ps, float Qv, float QWI, float& a1g, float& a2g) {
    const float tK = T + 273.15f;
    float Si = Qv / QWI;
    const float X = d_DQVair(T, ps);
    const float A = E0v / (Ka * tK) * (Ls * Mw * 0.001f / (R * tK) - 1);
    const float B = R * tK * X * Mw * 0.001f * d_ei(T) * 1000;
    a1g = 206.2f * (Si - 1) / (A + B);
    a2g = expf(0.09f * T);
}
// :336-369
__device__ __forceinline__ float PSFW(float dt, float T, float ps, float Qs, float Qw, float Qc, float Qv, float QWI,
                                      float Dair) {
    if (Qs > 0 && Qw > 0) {
        const float RI50 = 5e-3f, mI50 = 4.8e-7f, mI40 = 2.46e-7f, UI50 = 100.0f;
        float a1g, a2g;
        bergeron(T, ps, Qv, QWI, a1g, a2g);
        const float dt1 = 1 / (a1g * (1 - a2g)) * (powPos(mI50, 1 - a2g) - powPos(mI40, 1 - a2g));
        const float Qc50 = Qc * (dt / dt1) / 100;
        const float NI50 = Qc50 / mI50;
        return NI50 * (a1g * powPos(mI50, a2g) + PI * 1.0f * (Dair * 0.001f) * Qw * RI50 * RI50 * UI50);
    }
    return 0.0f;
}
// :371-395
__device__ __forceinline__ float PSFI(float T, float ps, float Qs, float Qc, float Qv, float QWI) {
    if (Qs > 0 && Qc > 0) {
        const float mI50 = 4.8e-7f, mI40 = 2.46e-7f;
        float a1g, a2g;
        bergeron(T, ps, Qv, QWI, a1g, a2g);
        if (a1g > 0) return a2g * a1g * Qc / (sqrtf(mI50) - sqrtf(mI40));
    }
    return 0.0f;
}
// :397-422
__device__ __forceinline__ float PSDEP(const Gammas& gm, float T, float ps, float Qv, float QWI, float Dair, float sS) {
    if (T < 0.0f) {
        const float tK = T + 273.15f;
        const float c = 152.93f, d = 0.25f;
        const float dQv = d_DQVair(T, ps);
        const float kv = d_ViscAir(T);
        float Sc = kv / dQv;
        float Si = Qv / QWI;
        const float A = Ls * Ls / (Ka * Rsw * tK * tK);
        const float B = 1 / (Dair * 0.001f * d_wi(T, ps) * dQv);
        const float density = powPos(1.225f / Dair, 0.25f);
        return PI * PI * (Si - 1) / (Dair * 0.001f * (A + B)) * 3e-2f *
               (0.78f * powPos(sS, -2) + 0.31f * powPos(Sc, 1.0f / 3.0f) * gm.gSS * sqrtf(c) * density * powPos(kv, -0.5f) *
                                           powPos(sS, -(d + 5) / 2));
    }
    return 0.0f;
}
// :429-460 (the limited LPSACW/LPSACR are computed but unused there)
__device__ __forceinline__ float PSMLT(const Gammas& gm, float T, float ps, float Qs, float Qv, float Dair, float sS,
                                       float psacw, float psacr) {
    if (T >= 0.0f && Qs > 0.0f) {
        const float c = 152.93f, d = 0.25f;
        const float dQv = d_DQVair(T, ps);
        const float kv = d_ViscAir(T);
        float Sc = kv / dQv;
        const float Drs = d_ws(0.0f, ps) - Qv;
        const float density = powPos(1.225f / Dair, 0.25f);
        return fmaxf(0.0f, -1 * (-(2 * PI / (density * 0.001f * Lf)) * (Ka * T - E0v * dQv * density * 0.001f * Drs) * 3e-2f *
                                     (0.78f * powPos(sS, -2) + 0.31f * powPos(Sc, 1.0f / 3.0f) * gm.gSS * sqrtf(c) *
                                                                 density * powPos(kv, -0.5f) * powPos(sS, -(d + 5) / 2)) -
                                 (Cvl * T / Lf) * (psacw + psacr)));
    }
    return 0.0f;
}
// :462-470
__device__ __forceinline__ float PGAUT(float T, float Qs, float QiMin) {
    const float surplus = Qs - QiMin;
    if (surplus > 0.0f && T < 0.0f) return (1e-3f * expf(0.09f * T)) * (surplus);
    return 0.0f;
}
// :472-487
__device__ __forceinline__ float PGFR(float T, float Qr, float Dair, float sR) {
    if (Qr > 0.0f && T < 0.0f)
        return 20 * PI * PI * 100.0f * 8e-2f * (0.99f / (Dair * 0.001f)) * (expf(0.66f * (-T)) - 1) * powPos(sR, -7);
    return 0.0f;
}
__device__ __forceinline__ float hailTerm(float Dair) {
    return sqrtf(4 * g * 100 * 0.91f / (3 * 0.6f * Dair * 0.001f));
}
// :489-504
__device__ __forceinline__ float PGACW(const Gammas& gm, float Qi, float Qw, float sI, float Dair) {
    if (Qi > 0.0f && Qw > 0.0f) return PI * 1.0f * 4e-4f * Qw * gm.gI / (4 * powPos(sI, 3.5f)) * hailTerm(Dair);
    return 0.0f;
}
// :506-521
__device__ __forceinline__ float PGACI(const Gammas& gm, float Qi, float Qc, float sI, float Dair) {
    if (Qi > 0.0f && Qc > 0.0f) return PI * 1.0f * 4e-4f * Qc * gm.gI / (4 * powPos(sI, 3.5f)) * hailTerm(Dair);
    return 0.0f;
}
// :523-539
__device__ __forceinline__ float PGACR(float Qi, float Qr, float Dair, float3 fv, float sR, float sI) {
    if (Qi > 0.0f && Qr > 0.0f)
        return PI * PI * 0.1f * 8e-2f * 4e-4f * fabsf(fv.z - fv.x) * (0.99f / (Dair * 0.001f)) * coll(sR, sI);
    return 0.0f;
}
// :541-558
__device__ __forceinline__ float PGACS(float T, float Qs, float Qi, float Dair, float3 fv, float sS, float sI, bool EGS1) {
    if (Qs > 0.0f && Qi > 0.0f) {
        const float EGS = T >= 0.0f || EGS1 ? 1.0f : expf(0.09f * T);
        return PI * PI * EGS * 3e-2f * 4e-4f * fabsf(fv.z - fv.y) * (0.11f / (Dair * 0.001f)) * coll(sS, sI);
    }
    return 0.0f;
}
// :560-570
__device__ __forceinline__ float PGDRY(float Qw, float Qc, float Qr, float Qs, float gacw, float gaci, float gacr,
                                      float gacs, float dts) {
    const float a = fminf(Qw, gacw * dts), b = fminf(Qc, gaci * dts), c = fminf(Qr, gacr * dts), d = fminf(Qs, gacs * dts);
    return a + b + c + d;
}
__device__ __forceinline__ float brackI(const Gammas& gm, float sI, float Sc, float kv, float density) {
    return (0.78f * powPos(sI, -2) + 0.31f * powPos(Sc, 1.0f / 3.0f) * gm.gRI * density * powPos(kv, -0.5f) * powPos(sI, -2.75f));
}
// :572-600
__device__ __forceinline__ float PGSUB(const Gammas& gm, float T, float ps, float Qv, float QWI, float Qi, float Dair,
                                      float sI) {
    float Si = Qv / QWI;
    if (Qi > 0.0f && Si < 1.0f) {
        const float tK = T + 273.15f;
        const float dQv = d_DQVair(T, ps);
        const float kv = d_ViscAir(T);
        float Sc = kv / dQv;
        const float A = Ls * Ls / (Ka * Rsw * tK * tK);
        const float B = 1 / (Dair * 0.001f * d_wi(T, ps) * dQv);
        const float density = powPos(4 * g * 100 * 0.91f / (3 * 0.6f * Dair * 0.001f), 0.25f);
        return fmaxf(0.0f, -1 * PI * PI * (Si - 1) / (Dair * 0.001f * (A + B)) * 4e-4f * brackI(gm, sI, Sc, kv, density));
    }
    return 0.0f;
}
// :602-629
__device__ __forceinline__ float PGMLT(const Gammas& gm, float T, float ps, float Qw, float Qr, float Qi, float Qv,
                                      float Dair, float sI, float gacw, float gacr, float dts) {
    if (Qi > 0.0f) {
        const float LW = fminf(Qw, gacw * dts);
        const float LR = fminf(Qr, gacr * dts);
        const float dQv = d_DQVair(T, ps);
        const float kv = d_ViscAir(T);
        float Sc = kv / dQv;
        const float Drs = d_ws(0.0f, ps) - Qv;
        const float density = powf(4 * g * 100 * 0.91f / (3 * 0.6f), 0.25f);
        return fmaxf(0.0f, -1 * (-2 * PI / (Dair * 0.001f * Lf) * (Ka * T - E0v * dQv * Dair * 0.001f * Drs) * 4e-4f *
                                     brackI(gm, sI, Sc, kv, density) -
                                 Cvl * T / Lf * (LW + LR)));
    }
    return 0.0f;
}
// :632-665
__device__ __forceinline__ float PGWET(const Gammas& gm, float T, float ps, float Qc, float Qs, float Qi, float Qv,
                                      float3 fv, float Dair, float sS, float sI, float gaci, float gacs, float dts) {
    if (Qi > 0.0f) {
        const float gaci1 = gaci * 10.0f;
        const float gacs1 = T >= 0.0f ? gacs : PGACS(T, Qs, Qi, Dair, fv, sS, sI, true);
        const float LI = fminf(Qc, gaci1 * dts);
        const float LS = fminf(Qs, gacs1 * dts);
        const float dQv = d_DQVair(T, ps);
        const float kv = d_ViscAir(T);
        float Sc = kv / dQv;
        const float Drs = d_ws(0.0f, ps) - Qv;
        const float density = powf(4 * g * 100.0f * 0.91f / (3 * 0.6f), 0.25f);
        return 2 * PI * 4e-4f * (Dair * 0.001f * E0v * dQv * Drs - Ka * T) / (Dair * 0.001f * (Lf + Cvl * T)) *
                   brackI(gm, sI, Sc, kv, density) +
               (LI + LS) * (1 - Cpi * T / (Lf + Cvl * T));
    }
    return 0.0f;
}
// :667-685
__device__ __forceinline__ float PGACR1(float T, float Qw, float Qc, float Qs, float Qi, float Dair, float3 fv, float sS,
                                       float sI, float gwet, float gacw, float gaci, float gacs, float dts) {
    if (Qi > 0.0f) {
        const float gaci1 = gaci * 10.0f;
        const float gacs1 = T >= 0.0f ? gacs : PGACS(T, Qs, Qi, Dair, fv, sS, sI, true);
        const float LI = fminf(Qc, gaci1 * dts);
        const float LS = fminf(Qs, gacs1 * dts);
        const float LW = fminf(Qw, gacw * dts);
        return gwet - LW - LI - LS;
    }
    return 0.0f;
}

// ------------------------------------------------------- ground surface ---
// :687-696
__device__ __forceinline__ float PGFLW(float Qr) { return Qr > 0.0f ? 100.0f * 5e-5f * Qr : 0.0f; }
// :698-729
__device__ __forceinline__ float PGGMLT(float tGC, float Qi, float Dair, float sI) {
    if (Qi > 0.0f) {
        const float k4 = 0.955f * 1.0e-2f;
        const float currentDens = Qi * Dair * 0.001f;
        const float amountIce = currentDens / 0.91f;
        const float iceHeight = amountIce * 100.0f;
        const float af = 0.45f;
        const float QGround = af * k4 * tGC / iceHeight;
        const float meltRate = fmaxf(0.0f, QGround / Lf);
        return meltRate * 4e-4f * powPos(sI, -2) * af / (Dair * 0.001f);
    }
    return 0.0f;
}
// :731-795. Within the block, 'Qs' is the local shortwave term that shadows
// the snow argument (Game/microPhysics.cu:736).
__device__ __forceinline__ float PGSMLT(float tAK, float tGC, float ps, float snow, float Qv, float Qr, float irr,
                                       float cc, float Dair, float wind, float dt) {
    if (snow > 0.0f) {
        float Qs = 0.0f, Ql = 0.0f, Qh = 0.0f, Qe = 0.0f, Qg = 0.0f, Qp = 0.0f;
        const float Tac = tAK - 273.15f;
        const float Ts = 0.0f;
        const float Tsk = Ts + 273.15f;
        const float Si = irr;
        const float A = 0.9f;
        Qs = (1 - A) * Si;
        const float Qlout = 310.0f;
        const float eac = 9.2e-6f * tAK * tAK;
        const float ea = (1 - 0.84f * cc) * eac + 0.84f * cc;
        const float Qlin = ea * oo * Tsk * Tsk * Tsk * Tsk;
        Ql = Qlin - Qlout;
        const float k = 0.41f, z0m = 0.001f, z0h = 0.0002f;
        const float Ch = k * k / (logf(1.0f / z0m) * logf(2.0f / z0h));
        Qh = Dair * Cpd * Ch * wind * (Tac - Ts);
        const float Ce = Ch;
        const float Qvs = Qv / (1 + Qv);
        Qe = Dair * E0v * Ce * wind * (Qvs - d_qs(Ts, ps));
        const float Td = tGC - Ts;
        if (Td > 0.0f) {
            const float currentDens = Qs * Dair * 0.001f;
            const float amountSnow = currentDens / 0.11f;
            const float snowHeight = amountSnow;
            Qg = 0.3f * (tGC - Ts) / snowHeight;
        }
        Qp = Cvl * 0.001f * Qr * Dair * tAK;
        const float Q = Qs + Ql + Qh + Qe + Qg + Qp;
        const float Qss = Qs * Dair;
        const float QE = (Q * dt) / Qss;
        const float meltFrac = QE / Lf;
        return fmaxf(0.0f, meltFrac * Qs);
    }
    return 0.0f;
}
// :797-845
__device__ __forceinline__ float PGREVP(float tAK, float tGC, float ps, float Qv, float Qr, float irr, float Dair,
                                       float wind) {
    if (Qr != 0.0f) {
        const float tGK = tGC + 273.15f;
        const float VPD = d_es(tGC) * 1000.0f * (1 - Qv / d_ws(tGC, ps));
        const float k = 0.41f, z0m = 0.0001f, z0h = z0m * 0.1f;
        const float ga = wind * k * k / (logf(1.0f / z0m) * logf(2.0f / z0h));
        const float y = Cpd * ps * 100.0f / (0.622f * E0v);
        const float esw = 0.96f, a = 0.08f;
        const float aP = d_es(tAK - 273.15f) * 10.0f * (Qv / d_ws(tAK - 273.15f, ps));
        const float ea = 1.24f * powPos((aP / tAK), 1.0f / 7.0f);
        const float Rn = irr * (1 - a) + oo * esw * (ea * (tAK * tAK * tAK * tAK) - (tGK * tGK * tGK * tGK));
        const float RnG = Rn - 0.2f * Rn;
        const float m = d_es(tGC) * 1000.0f * E0v / (Rsw * tGK * tGK);
        const float Em = (m * RnG + Dair * Cpd * VPD * ga) / (E0v * (m + y));
        const float EmD = fmaxf(0.0f, Em / Dair);
        const float areaW = fminf(1.0f, 0.01f * sqrtf(Qr));
        return EmD * areaW;
    }
    return 0.0f;
}
// :847-871
__device__ __forceinline__ float PGDEVP(float* time, float Qr, float Qs, float Qi, float Qrs) {
    if (Qrs != 0.0f) {
        if (Qr == 0 && Qs == 0 && Qi == 0) {
            const float secsInDay = 60 * 60 * 24;
            const float O_ = 0.21f * secsInDay;
            return 200.0f * 1e-6f * Qrs * expf(-*time / O_);
        } else {
            *time = 0.0f;
        }
    }
    return 0.0f;
}
// :873-934
__device__ __forceinline__ float PGGFR(float tAK, float tGC, float ps, float Qr, float Qv, float Qi, float Dair, float irr,
                                      float cc, float wind, float dt) {
    const float Tac = tAK - 273.15f;
    if (Qr != 0.0f && Tac < 0.0f) {
        float Qw = 0.0f, Qc = 0.0f, Qe = 0.0f, Ql = 0.0f, Qs = 0.0f, Qg = 0.0f, Qf = 0.0f;
        Qf = Lf * Qr * Dair / dt;
        Qw = Cvl * (273.15f - tAK) * Qr * Dair / dt;
        const float k = 0.41f, z0m = 0.0001f, z0h = z0m * 0.1f;
        const float ga = wind * k * k / (logf(1.0f / z0m) * logf(1.0f / z0h));
        Qc = Dair * Cpd * ga * (tGC - Tac);
        const float Ce = k * k;
        const float Qvs = Qv / (1 + Qv);
        Qe = Dair * E0v * Ce * wind * fmaxf(0.0f, Qvs - d_qs(0, ps));
        const float Si = irr, A = 0.06f;
        Qs = -(1 - A) * Si;
        const float Qlout = 310.0f;
        const float eac = 9.2e-6f * tAK * tAK;
        const float ea = (1 - 0.84f * cc) * eac + 0.84f * cc;
        const float Qlin = ea * oo * 273.15f * 273.15f * 273.15f * 273.15f;
        Ql = Qlin - Qlout;
        if (Qi > 0.0f) {
            const float currentDens = Qi * Dair * 0.001f;
            const float amountIce = currentDens / 0.91f;
            const float iceHeight = amountIce;
            Qg = 0.3f * (tGC - 0) / iceHeight;
            Qg *= -1;
        }
        const float f = fminf(fmaxf((Qw + Qe + Qc + Qs + Ql + Qg) / Qf, 0.0f), 1.0f);
        return f * Qr;
    }
    return 0.0f;
}
// :936-980
__device__ __forceinline__ float PGSFR(float tAK, float tGC, float ps, float Qv, float Qr, float Qs, float Dair, float wind,
                                      float dt) {
    if (Qr != 0.0f && Qs != 0.0f) {
        float Qw = 0.0f, Qe = 0.0f, Qg = 0.0f, Qf = 0.0f;
        Qf = Lf * Qr * Dair / dt;
        Qw = Cvl * (273.15f - tAK) * Qr * Dair / dt;
        const float k = 0.41f;
        const float Ce = k * k;
        const float Qvs = Qv / (1 + Qv);
        Qe = Dair * E0v * Ce * wind * fmaxf(0.0f, Qvs - d_qs(0, ps));
        const float currentDens = Qs * Dair * 0.001f;
        const float amountSnow = currentDens / 0.11f;
        const float snowHeight = amountSnow;
        Qg = 0.3f * (tGC - 0) / snowHeight;
        Qg *= -1;
        const float f = fminf(fmaxf((Qw + Qe + Qg) / Qf, 0.0f), 1.0f);
        return f * Qr;
    }
    return 0.0f;
}
// :982-1032
__device__ __forceinline__ float PGRFR(float tAK, float tGC, float ps, float Qr, float Qv, float Qi, float Qgr, float Qgs,
                                      float Dair, float wind, float3 fv) {
    const float Tac = tAK - 273.15f;
    if (!(Qgr != 0.0f) && !(Qgs != 0.0f) && Qr != 0.0f && Tac < 0.0f) {
        float Qw = 0.0f, Qc = 0.0f, Qe = 0.0f, Ql = 0.0f, Qs = 0.0f, Qg = 0.0f, Qf = 0.0f;
        const float densW = 0.99f;
        const float Rr = Qr * Dair * fv.x;
        const float P = Rr * 3600.0f / (densW * 1000.0f) * 1000;
        const float FFW = sqrtf(powPos(P * densW / 3.6f, 2.0f) + powPos(0.067f * powPos(P, 0.846f) * wind, 2.0f));
        Qf = Lf * 0.001f * FFW;
        Qw = Cvl * 0.001f * (0 - Tac) * FFW;
        const float k = 0.41f, z0m = 0.1f, z0h = z0m * 0.1f;
        const float ga = wind * k * k / (logf(1.0f / z0m) * logf(1.0f / z0h));
        Qc = Dair * Cpd * ga * (tGC - Tac);
        const float Ce = k * k;
        const float Qvs = Qv / (1 + Qv);
        Qe = Dair * E0v * Ce * wind * fmaxf(0.0f, Qvs - d_qs(0, ps));
        const float Td = tGC - 0;
        if (Td > 0.0f) {
            const float currentDens = Qi * Dair * 0.001f;
            const float amountIce = currentDens / 0.91f;
            const float iceHeight = amountIce;
            Qg = 0.3f * (tGC - 0) / iceHeight;
            Qg *= -1;
        }
        const float f = fminf(fmaxf((Qw + Qc + Qe + Ql + Qs + Qg) / Qf, 0.0f), 1.0f);
        return f * Qr;
    }
    return 0.0f;
}

}  // namespace mp
}  // namespace skt
