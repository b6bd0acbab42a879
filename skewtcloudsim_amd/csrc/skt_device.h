// Device-side physics and boundary helpers for the skewT cloud step on gfx950.
//
// Formulas follow the reference's GPU path (Game/meteoformulas.cuh,
// Game/microPhysics.cu, Game/kernelSky.cu); each expression keeps the
// reference's operand order so that float rounding matches term by term.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace skt {

// Physical constants, Game/environment.cu:28-60.
namespace kc {
constexpr float g = 9.8076f, Hv = 2501000.0f, R = 8.3145f, Rsd = 287.0528f, Rsw = 461.5f;
constexpr float E = 287.0528f / 461.5f;
constexpr float Cpd = 1003.5f, Cvl = 4119.0f, Cpi = 2093.0f, Cpds = 800.0f, Mw = 18.02f;
constexpr float E0v = 2.374e+6f, Lf = 3.3355e+5f, Ls = 2.834e+6f, Ka = 2.40e-2f;
constexpr float PI = 3.14159265359f, oo = 5.67e-8f, ge = 0.95f;
}  // namespace kc

enum { BC_NEUMANN = 0, BC_DIRICHLET = 1, BC_CUSTOM = 2 };

struct Bounds {
    int sides, up, ground;
};

// Storage of a field. FMT_F32 holds floats. The 16-bit forms hold
// value * scale (scale a power of two, so the scaling itself is exact) as a
// bfloat16 (FMT_BF16) or an IEEE half (FMT_F16), rounded to nearest even;
// the mixed-precision form stores the five hydrometeor fields this way.
enum { FMT_F32 = 0, FMT_BF16 = 1, FMT_F16 = 2 };
struct Store {
    int fmt;
    float scale, inv;
};
__host__ __device__ inline Store storeF32() { return Store{FMT_F32, 1.0f, 1.0f}; }
struct Store3 {
    Store s[3];
};
__host__ __device__ inline float fromH16(unsigned short h, const Store& st) {
    const float v = st.fmt == FMT_BF16 ? __builtin_bit_cast(float, (unsigned)h << 16)
                                       : (float)__builtin_bit_cast(_Float16, h);
    return v * st.inv;
}
__host__ __device__ inline unsigned short toH16(float v, const Store& st) {
    const float s = v * st.scale;
    if (st.fmt == FMT_BF16) {
        const unsigned u = __builtin_bit_cast(unsigned, s);
        if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // quiet NaN
        return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
    return __builtin_bit_cast(unsigned short, (_Float16)s);
}
// Element i of a field stored as st says (MIX: the storage may be 16-bit;
// otherwise it is float and no format test is compiled).
template <bool MIX>
__device__ __forceinline__ float ldS(const void* p, long i, const Store& st) {
    if constexpr (MIX) {
        if (st.fmt != FMT_F32) return fromH16(static_cast<const unsigned short*>(p)[i], st);
    }
    return static_cast<const float*>(p)[i];
}
template <bool MIX>
__device__ __forceinline__ void stS(void* p, long i, float v, const Store& st) {
    if constexpr (MIX) {
        if (st.fmt != FMT_F32) {
            static_cast<unsigned short*>(p)[i] = toH16(v, st);
            return;
        }
    }
    static_cast<float*>(p)[i] = v;
}

// Grid geometry. Index = x + X*(y + Y*z); y is vertical; gh[x + X*z] is the
// highest ground level of each column (cells with y <= gh are ground).
// ghm1 / ghm2 hold the largest gh over the 3x3 / 5x5 neighbourhood of each
// column: a cell at y >= ghm + 2 (resp. + 3) has only sky within 1 (2) cells,
// which selects the boundary-free fast path of the stencil kernels.
struct Geom {
    int X, Y, Z;
    long N;
    const int* __restrict__ gh;
    const int* __restrict__ ghm1;
    const int* __restrict__ ghm2;
};

struct Gammas {
    float R48, S425;                    // fall speeds: Gamma(4+b), Gamma(4+d)
    float gR, gER, gS, gRC, gSS, gI, gRI;  // microphysics gammas
};

__device__ __forceinline__ long gidx(const Geom& G, int x, int y, int z) {
    return (long)x + (long)G.X * ((long)y + (long)G.Y * z);
}
__device__ __forceinline__ bool isGround(const Geom& G, int x, int y, int z) {
    return y <= G.gh[x + G.X * z];
}
__device__ __forceinline__ float bcv(int cond, float data, float cust) {
    return cond == BC_NEUMANN ? data : (cond == BC_DIRICHLET ? 0.0f : cust);
}
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Stencil read of field q at (px,py,pz) from a sky cell: raw for inside sky,
// the ground condition for ground cells, and the condition of the last inside
// cell for positions beyond the domain (halo of Game/kernelSky.cu:1765-2041;
// the top face takes the sides condition).
__device__ __forceinline__ float extv(const Geom& G, const float* __restrict__ q, const float* __restrict__ prof,
                                      Bounds b, int px, int py, int pz) {
    const int bx = clampi(px, 0, G.X - 1), by = clampi(py, 0, G.Y - 1), bz = clampi(pz, 0, G.Z - 1);
    const bool out = (bx != px) | (by != py) | (bz != pz);
    const long bi = gidx(G, bx, by, bz);
    const float cv = prof ? prof[by] : 0.0f;
    if (isGround(G, bx, by, bz)) return bcv(b.ground, q[bi], cv);
    if (out) return bcv(b.sides, q[bi], cv);
    return q[bi];
}

// Neighbour read through fillNeighbourData with the cell's own value as the
// boundary datum (Game/kernelSky.cu:1844-1874).
__device__ __forceinline__ float nbOwn(const Geom& G, const float* __restrict__ v, Bounds b, float cust, int x, int y,
                                       int z, int ox, int oy, int oz) {
    const int px = x + ox, py = y + oy, pz = z + oz;
    const long c = gidx(G, x, y, z);
    if (px < 0 || px >= G.X || py < 0 || py >= G.Y || pz < 0 || pz >= G.Z) return bcv(b.sides, v[c], cust);
    if (isGround(G, px, py, pz)) return bcv(b.ground, v[c], cust);
    return v[gidx(G, px, py, pz)];
}

__device__ __forceinline__ bool skyAt(const Geom& G, int x, int y, int z) {
    return x >= 0 && x < G.X && y >= 0 && y < G.Y && z >= 0 && z < G.Z && !isGround(G, x, y, z);
}

// ------------------------------------------------------------ powers ---
// x^y for the physics' powers from the hardware base-2 logarithm and
// exponential (v_log_f32, v_exp_f32), in place of the library powf, which
// spends about 160 instructions on an extended-precision logarithm and on
// the special cases of every sign and size. The per-cell phases that raise
// p / p_ground to Rd/cp or take Marshall-Palmer slopes to fractional powers
// are bound by those instructions. tests/test_pow.py measures both forms
// against the exact power over the ranges the phases use.
//
// powNear1: x within a few octaves of 1 and |y log2 x| of order 1 (the
// pressure ratio to Rd/cp, the temperature ratios of the vapour
// diffusivity and the air viscosity): log2 x is small, so its rounding and
// the product's stay well under an ulp of the result. Three instructions.
__device__ __forceinline__ float powNear1(float x, float y) {
    return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
}
// powPos: any normal x > 0 and any y. With x = m 2^e, m in [sqrt(1/2),
// sqrt(2)): y log2 x = y e + y log2 m. Both products are split exactly into
// their rounded values and rounding errors (fma); the integer n nearest y e
// is taken out before the fraction (y e - n, exact) is summed with y log2 m
// and the two errors, so the exponential's argument keeps a fraction's
// precision however large y log2 x is, and 2^n is applied exactly (ldexp).
// What remains is the logarithm's own error, |y| times an ulp of |log2 m|
// <= 1/2, and the exponential's. Other x (0, subnormal, negative, infinite,
// NaN) take the library powf.
__device__ __forceinline__ float powPos(float x, float y) {
    if (!(x >= 0x1p-126f && x <= 3.40282347e38f)) return powf(x, y);
    float m = __builtin_amdgcn_frexp_mantf(x);
    int e = __builtin_amdgcn_frexp_expf(x);
    if (m < 0.70710678f) {
        m *= 2.0f;
        e -= 1;
    }
    const float fe = (float)e;
    const float a = y * fe;
    const float ae = __builtin_fmaf(y, fe, -a);
    const float l = __builtin_amdgcn_logf(m);
    const float b = y * l;
    const float be = __builtin_fmaf(y, l, -b);
    const float n = __builtin_rintf(a);
    const float f = ((a - n) + b) + (ae + be);
    const float nf = __builtin_rintf(f);
    return __builtin_amdgcn_ldexpf(__builtin_amdgcn_exp2f(f - nf), (int)(n + nf));
}

// ------------------------------------------------------------ thermo ---
// Game/meteoformulas.cuh:16-316
__device__ __forceinline__ float d_es(float T) { return 0.61078f * expf((17.27f * T) / (T + 237.3f)); }
__device__ __forceinline__ float d_ei(float T) { return 0.61078f * expf((21.875f * T) / (T + 265.5f)); }
__device__ __forceinline__ float d_ws(float T, float P) {
    float Es = d_es(T) * 10.0f;
    return (kc::E * Es) / (P - Es);
}
__device__ __forceinline__ float d_wi(float T, float P) {
    float Es = d_ei(T) * 10.0f;
    return (kc::E * Es) / (P - Es);
}
__device__ __forceinline__ float d_qs(float T, float P) {
    float Es = d_ei(T) * 10.0f;
    return (kc::E * Es) / (P - (1 - kc::E) * Es);
}
__device__ __forceinline__ float d_Ni(float T) {
    const float EI = d_ei(T) * 1000;
    const float ES = d_es(T) * 1000;
    return 10000 * expf((12.96f * (ES - EI)) / (EI - 0.639f));
}
__device__ __forceinline__ float d_pwater(float T) {
    return 999.83311f + (0.0752f * T) - (0.0089f * T * T) + (7.36413e-5f * T * T * T) - (4.74639e-7f * T * T * T * T) +
           (1.34888e-9f * T * T * T * T * T);
}
__device__ __forceinline__ float d_Lwater(float T) {
    return (2500.8f - 2.36f * T + 0.0016f * T * T - 0.00006f * T * T * T) * 1e3f;
}
__device__ __forceinline__ float d_Lice(float T) { return (2834.1f - 0.29f * T - 0.004f * T * T) * 1e3f; }
__device__ __forceinline__ float d_DQVair(float T, float P) {
    return 2.26e-5f * powNear1((T + 273.15f) / 273.15f, 1.94f) * (1013.25f / P);
}
__device__ __forceinline__ float d_ViscAir(float T) {
    return powNear1((T + 273.15f) / 273.15f, 3.0f / 2.0f) * ((273.15f + 111.0f) / (T + 111.0f)) * 1.716e-5f;
}
__device__ __forceinline__ float d_slope(float D, float Qj, int type) {
    float num;
    if (type == 0) num = kc::PI * 0.99f * 8e-2f;
    else if (type == 1) num = kc::PI * 0.11f * 3e-2f;
    else if (type == 2) num = kc::PI * 0.91f * 4e-4f;
    else return 0.0f;
    const float den = fmaxf(D * Qj * 0.001f, 1e-14f);
    return powPos((num / den), 0.25f);
}
// calculateFallingVelocityGPU: type 3 ('all') evaluates the rain branch only.
__device__ __forceinline__ float3 d_fallVel(float Qr, float Qs, float Qi, float dens, int type, float GR, float GS,
                                            float GI) {
    const bool all = type == 3;
    float UR = 0.0f, US = 0.0f, UI = 0.0f;
    if (type == 0 || all) {
        float sR = d_slope(dens, Qr, 0);
        UR = 2115.0f * (GR / (6 * powPos(sR, 0.8f))) * sqrtf(1.225f / dens);
        UR *= 0.01f;
    } else if (type == 1 || all) {
        float sS = d_slope(dens, Qs, 1);
        US = 152.93f * (GS / (6 * powPos(sS, 0.25f))) * sqrtf(1.225f / dens);
        US *= 0.01f;
    } else if (type == 2 || all) {
        float sI = d_slope(dens, Qi, 2);
        UI = (GI / (6 * sqrtf(sI))) * sqrtf(4 * kc::g * 100 * 0.91f / (3 * 0.6f * dens * 0.001f));
        UI *= 0.01f;
    }
    return make_float3(UR, US, UI);
}
__device__ __forceinline__ float d_cvd(float T, float P, float dens) {
    const float EI = d_ei(T) * 1000;
    const float ES = d_es(T) * 1000;
    const float A = (kc::Ls / (kc::Ka * (T + 273.15f))) * (kc::Ls / (kc::Rsw * (T + 273.15f)) - 1);
    const float B = (kc::Rsw * (T + 273.15f) * P) / (2.21f * EI);
    const float Ni = 10000 * expf((12.96f * (ES - EI)) / (EI - 0.639f));
    return 65.2f * ((sqrtf(Ni) * (ES - EI)) / (sqrtf(dens) * (A + B) * EI));
}
__device__ __forceinline__ float d_potTemp(float T, float Pk, float Pt) {
    return (T + 273.15f) * powNear1((Pt / Pk), kc::Rsd / kc::Cpd) - 273.15f;
}
__device__ __forceinline__ float d_MLR(float T, float P) {
    float TK = T + 273.15f;
    float W = d_ws(T, P);
    float Tw = ((kc::Rsd * TK + kc::Hv * W) / (kc::Cpd + (kc::Hv * kc::Hv * W * kc::E / (kc::Rsd * TK * TK))));
    return Tw / P;
}

// PPM flux with van Leer slope and Walcek limiter, Game/kernelSky.cu:270-366.
__device__ __forceinline__ float ppmFlux(float vel, float L, float Cc, float Rr, float dt, float dx) {
    float Cn = vel * dt / dx;
    Cn = Cn > 1.0f ? 1.0f : (Cn < -1.0f ? -1.0f : Cn);
    const float qMin = fminf(L, fminf(Cc, Rr));
    const float qMax = fmaxf(L, fmaxf(Cc, Rr));
    float s = 0.0f;
    const float dl = Cc - L, dr = Rr - Cc;
    if (dl * dr > 0.0f) {
        const float db = Rr - L;
        const float md = fminf(fabsf(dr), fabsf(dl));
        const float sl = fminf(0.5f * fabsf(db), md);
        s = db >= 0.0f ? sl : -sl;
    }
    const float qLS = Cc - 0.5f * s, qRS = Cc + 0.5f * s;
    float B = 1.0f;
    if (qRS > qMax) {
        const float den = qRS - Cc;
        if (den > 1e-16f) B = fminf(B, (qMax - Cc) / den);
    }
    if (qLS < qMin) {
        const float den = Cc - qLS;
        if (den > 1e-16f) B = fminf(B, (Cc - qMin) / den);
    }
    float qR = Cc + B * (qRS - Cc);
    float qL = Cc + B * (qLS - Cc);
    if ((qR - Cc) * (Cc - qL) <= 0.0f) {
        qR = Cc;
        qL = Cc;
    }
    float qq;
    if (Cn > 0.0f) qq = qR + Cn * (-qL - 2 * qR + 3 * Cc) + Cn * Cn * (qL + qR - 2 * Cc);
    else qq = qR + Cn * (-2 * qL - qR + 3 * Cc) + Cn * Cn * (qL + qR - 2 * Cc);
    return vel * qq;
}

// ---------------------------------------------------------- reciprocals ---
// The correctly rounded reciprocal in three instructions: the hardware
// reciprocal estimate with one FMA refinement. For |s| in [2^-126, 2^126]
// it equals the IEEE 1.0f / s bit for bit (checked for every such float by
// tests/test_face_coefficient.py); other s take the IEEE division (eleven
// instructions).
__device__ __forceinline__ bool recipOk(float s) {
    const float m = fabsf(s);
    return m >= 0x1p-126f && m <= 0x1p126f;
}
__device__ __forceinline__ float recipRNfast(float s) {
    const float e = __builtin_amdgcn_rcpf(s);
    return __builtin_fmaf(__builtin_fmaf(-s, e, 1.0f), e, e);
}
__device__ __forceinline__ float recipRN(float s) {
    if (!recipOk(s)) return 1.0f / s;
    return recipRNfast(s);
}
// 2 / s rounded to nearest: 2 RN(1/s) (scaling by 2 is exact there).
__device__ __forceinline__ float twoOverRN(float s) {
    if (!recipOk(s)) return 2.0f / s;
    return 2.0f * recipRNfast(s);
}

// ------------------------------------------------------ pressure operator ---
// Face coefficient of initAMatrix (Game/kernelSky.cu:1372-1377),
// nb == 0 ? 0 : -max(1, 1.225 / (2 / (1/cur + 1/nb))), as a function of
// s = 1/cur + 1/nb (the sum the reference forms): -max(1, 1.225 / (2 / s))
// with both divisions rounded to
// nearest. Scaling by 2 is exact, so 2 / s rounds to 2 RN(1/s) and the second
// quotient equals RN(0.6125 / RN(1/s)) (0.6125f is 1.225f / 2 exactly). The
// short form takes t = RN(1/s) from the hardware reciprocal estimate with one
// FMA refinement, and the quotient from q = RN(0.6125 s) with one remainder
// correction, q + (0.6125 - t q) s, in which s itself stands for 1/t
// (Markstein's iteration; s and 1/t differ by one rounding): six instructions
// and the max, one of them a reciprocal. For s in [2^-126, 2^126] every float
// gives the same bits as the IEEE expression (checked exhaustively over all
// 2^32 inputs by tests/test_face_coefficient.py); other s take the IEEE path.
__device__ __forceinline__ bool faceFastOk(float s) { return s >= 0x1p-126f && s <= 0x1p126f; }
// The second half of the short form, from s and t = RN(1/s) (recipRNfast of s).
__device__ __forceinline__ float faceOfRecip(float t, float s) {
    const float q = 0.6125f * s;
    return -fmaxf(1.0f, __builtin_fmaf(__builtin_fmaf(-t, q, 0.6125f), s, q));
}
__device__ __forceinline__ float faceFast(float s) { return faceOfRecip(recipRNfast(s), s); }
__device__ __forceinline__ float faceOfSum(float s) {
    if (faceFastOk(s)) return faceFast(s);
    return -fmaxf(1.0f, 1.225f / (2.0f / s));
}
// Densities whose reciprocals, face sums, harmonic means and operator
// diagonals all stay inside the ranges where recipRNfast and faceFast equal
// the IEEE expressions: |d| in [2^-100, 2^100] gives reciprocals there, sums
// of two in [2^-100, 2^101], faces and diagonals below 2^104. A pass checks
// this once per cell and takes the fast forms for the whole cell.
__device__ __forceinline__ bool densOk(float d) {
    const float m = fabsf(d);
    return m >= 0x1p-100f && m <= 0x1p100f;
}
// Fast face from two reciprocals of densities that pass densOk (an infinite
// one, a zero neighbour density or a ground face, gives 0).
__device__ __forceinline__ float faceIfast(float ic, float in) {
    const float v = faceFast(ic + in);
    return isinf(in) ? 0.0f : v;
}

// A sky cell's 1/density for which every face sum with another such value
// lies in the short form's range (or is infinite, a zero face).
__device__ __forceinline__ bool faceValueOk(float i) { return i >= 0x1p-127f && i <= 0x1p125f; }

// The face coefficient from the reciprocal densities of the two cells. A
// zero neighbour density (infinite reciprocal) gives 0, as nb == 0 does in
// the reference.
__device__ __forceinline__ float faceI(float ic, float in) { return isinf(in) ? 0.0f : faceOfSum(ic + in); }

// One row of the symmetric pressure operator at sky cell i = (x, y, z), as
// the SpMV forms it: the six face coefficients (faceI of the two cells'
// 1/density; a neighbour beyond the domain takes the cell's own 1/density
// and counts 0, a ground neighbour gives a zero face; the back face at z = 0
// is the domain boundary) and the diagonal w = -(sum of the six) summed in
// float in the SpMV's order. s* say which neighbours are sky cells in the
// domain (the off-diagonal entries). gh* are the ground heights of the
// cell's column and of its four neighbour columns (unused outside).
struct OpRow {
    float fr, fl, fu, fd, ff, fb, w;
    bool sR, sL, sU, sD, sF, sB;
};
__device__ __forceinline__ OpRow opRow(const float* __restrict__ irho, long i, int x, int y, int z, int X, int Y,
                                       int Z, long XY, int gh, int ghL, int ghR, int ghB, int ghF) {
    const float INF = __builtin_inff();
    const bool inL = x > 0, inR = x < X - 1, inB = z > 0, inF = z < Z - 1;
    OpRow o;
    o.sR = inR && y > ghR;
    o.sL = inL && y > ghL;
    o.sU = y + 1 < Y;
    o.sD = y > 0 && y - 1 > gh;
    o.sF = inF && y > ghF;
    o.sB = inB && y > ghB;
    // the six neighbours are read at offsets clamped into the domain
    // (unconditional loads, issued together) and replaced by the rules below
    const float ic = irho[i];
    const float vR = irho[i + (inR ? 1 : 0)], vL = irho[i - (inL ? 1 : 0)];
    const float vU = irho[i + (o.sU ? X : 0)], vD = irho[i - (y > 0 ? X : 0)];
    const float vF = irho[i + (inF ? XY : 0)], vB = irho[i - (inB ? XY : 0)];
    const float iR = !inR ? ic : (o.sR ? vR : INF);
    const float iL = !inL ? ic : (o.sL ? vL : INF);
    const float iU = o.sU ? vU : ic;
    const float iD = y == 0 ? ic : (o.sD ? vD : INF);
    const float iF = !inF ? ic : (o.sF ? vF : INF);
    o.fr = faceI(ic, iR);
    o.fl = faceI(ic, iL);
    o.fu = faceI(ic, iU);
    o.fd = faceI(ic, iD);
    o.ff = faceI(ic, iF);
    o.fb = z == 0 ? faceI(ic, ic) : (o.sB ? faceI(vB, ic) : 0.0f);
    float w = 0.0f;
    w -= o.fr;
    w -= o.fl;
    w -= o.fu;
    w -= o.fd;
    w -= o.ff;
    w -= o.fb;
    o.w = w;
    return o;
}

}  // namespace skt
