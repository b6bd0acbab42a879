// Geometric multigrid preconditioner of the standard pressure solve
// (skt_config.pcg_precon = SKT_PRECON_MG). No counterpart in the reference,
// whose PCG is preconditioned by 1/diagonal (Game/kernelSky.cu:659-673).
//
// The operator on level 0 is the solver's own: per sky cell the positive
// couplings to its +x, +y and +z neighbours (minus the face coefficients of
// opRow) and the diagonal w; a coupling to a -x/-y/-z neighbour is that
// neighbour's + coupling (the operator is symmetric). Level l + 1 coarsens
// 2x2x2 (a last odd cell alone); a coarse coupling is half the sum of the
// fine couplings across the coarse face, the coarse diagonal half the sum of
// all fine couplings that leave the coarse cell (domain-boundary faces
// included). A coarse cell with no sky cell has diagonal 0 and is excluded.
// The rows are stored as 16-bit floats (8 B per cell instead of 16), level l
// divided by 2^l (its right-hand side too; power-of-two scaling is exact),
// with the couplings rounded toward zero and the diagonal away from it, so
// each stored level stays diagonally dominant.
// The V-cycle: two damped Jacobi sweeps (the first from zero), the residual
// summed over the 2x2x2 children onto the coarse level, recursion, the
// coarse correction added to every child, two more sweeps; on the coarsest
// level (at most 512 cells) 64 sweeps from zero in LDS by one block. The
// restriction is the transpose of the prolongation and the pre- and
// post-smoothers are equal, so the cycle is a symmetric positive definite
// linear map and conjugate gradients keep their stop test.
//
// The launch sequence covers kMgLevels levels whatever the grid (a level
// past the grid's depth has no tiles), so members of any size share it in
// an ensemble. Every body returns at once when the solve has finished.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "skt_device.h"
#include "skt_launch.h"
#include "skt_reduce.h"

namespace skt {
namespace {

constexpr float kOmega = 0.8f;     // Jacobi damping
constexpr int kCoarseSweeps = 64;  // sweeps on the coarsest level (even: the result lands in xb)

// omega / d with the hardware reciprocal (about 1 ulp): the smoother's
// diagonal only has to be a fixed positive diagonal for the cycle to stay
// symmetric positive definite, and one instruction replaces a division's ten.
__device__ __forceinline__ float omegaOver(float w) { return kOmega * __builtin_amdgcn_rcpf(w); }

__device__ __forceinline__ long gstride(const VB& vb) { return (long)vb.n * 256; }
__device__ __forceinline__ long gfirst(const VB& vb) { return (long)vb.id * 256 + tid256(); }

// The 16-bit operator rows: loads, and stores that round the couplings
// toward zero and the diagonal away from it, so that every stored level
// stays diagonally dominant (its smoother and cycle stay symmetric positive
// definite).
__device__ __forceinline__ float4 ldC(const MgC* __restrict__ C, long i) {
    const MgC c = C[i];
    return float4{(float)c.x, (float)c.y, (float)c.z, (float)c.w};
}
__device__ __forceinline__ float ldCx(const MgC* __restrict__ C, long i) { return (float)C[i].x; }
__device__ __forceinline__ float ldCy(const MgC* __restrict__ C, long i) { return (float)C[i].y; }
__device__ __forceinline__ float ldCz(const MgC* __restrict__ C, long i) { return (float)C[i].z; }
__device__ __forceinline__ float ldCw(const MgC* __restrict__ C, long i) { return (float)C[i].w; }
__device__ __forceinline__ _Float16 halfStep(_Float16 h, int d) {
    return __builtin_bit_cast(_Float16, (unsigned short)(__builtin_bit_cast(unsigned short, h) + d));
}
__device__ __forceinline__ _Float16 halfDown(float v) {  // v >= 0
    _Float16 h = (_Float16)v;
    return (float)h > v ? halfStep(h, -1) : h;
}
__device__ __forceinline__ _Float16 halfUp(float v) {  // v >= 0
    _Float16 h = (_Float16)v;
    return (float)h < v ? halfStep(h, 1) : h;
}
__device__ __forceinline__ void stC(MgC* __restrict__ C, long i, float x, float y, float z, float w) {
    MgC c;
    c.x = halfDown(x);
    c.y = halfDown(y);
    c.z = halfDown(z);
    c.w = halfUp(w);
    C[i] = c;
}

// A level's dims and the z-chunk of its tiles. Tiles are 64 (x) x 4 (y)
// columns of zc planes; a thread marches its column through the chunk, so a
// cell's coordinates are never divided out of its index (levels hold fewer
// than 2^31 cells, so 32-bit indices serve).
struct MgDims {
    int X, Y, Z;
    long N;
    int zc;
};
struct MgCol {
    int x, y, z0, z1;
    bool in;
};
__device__ __forceinline__ MgCol mgCol(const VB& vb, const MgDims& D) {
    const unsigned gx = (unsigned)(D.X + 63) >> 6, gy = (unsigned)(D.Y + 3) >> 2;
    const unsigned bx = vb.id % gx, by = (vb.id / gx) % gy, bz = vb.id / (gx * gy);
    MgCol c;
    c.x = (int)bx * 64 + (int)threadIdx.x;
    c.y = (int)by * 4 + (int)threadIdx.y;
    c.z0 = (int)bz * D.zc;
    c.z1 = min(D.Z, c.z0 + D.zc);
    c.in = c.x < D.X && c.y < D.Y && c.z0 < c.z1;
    return c;
}

// The LDS tile of a plane for the level kernels whose stencil reads a value
// at the four in-plane neighbours: 64 x 4 cells of the block and a one-cell
// halo. Each lane publishes its own cell's value; the halo is read by the row
// below the tile (wave 0), the row above it (wave 1) and the columns left and
// right of it (lanes 0-3 and 4-7 of wave 2). Lanes past the level work on the
// cell clamped into it (their neighbours drop them by a select), take part
// in every barrier and store nothing.
struct MgTile {
    int ib, hb;            // the lane's cell and halo cell, index within a plane (clamped)
    int hx, hy;            // the halo cell's coordinates (clamped)
    int hrow, hcol;
    bool hasH, in;
};
__device__ __forceinline__ MgTile mgTile(const MgCol& t, int X, int Y) {
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
    const int x0 = t.x - tx, y0 = t.y - ty;
    MgTile g;
    int hx = t.x, hy = y0 - 1;
    g.hrow = 0;
    g.hcol = tx + 1;
    if (ty == 1) hy = y0 + 4, g.hrow = 5;
    if (ty == 2) hx = tx < 4 ? x0 - 1 : x0 + 64, hy = y0 + (tx & 3), g.hrow = 1 + (tx & 3), g.hcol = tx < 4 ? 0 : 65;
    g.hasH = ty < 2 || (ty == 2 && tx < 8);
    g.in = t.x < X && t.y < Y;
    g.ib = min(t.x, X - 1) + X * min(t.y, Y - 1);
    g.hx = g.hasH ? min(max(hx, 0), X - 1) : min(t.x, X - 1);
    g.hy = g.hasH ? min(max(hy, 0), Y - 1) : min(t.y, Y - 1);
    g.hb = g.hx + X * g.hy;
    return g;
}

// Level-0 operator from 1/density (after the divergence pass wrote irho).
struct KMgCoef0 {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, const float* __restrict__ irho,
                                               MgC* __restrict__ C) {
        const int X = G.X, Y = G.Y, Z = G.Z;
        const long XY = (long)X * Y, N = XY * Z;
        for (long i = gfirst(vb); i < N; i += gstride(vb)) {
            const int x = (int)(i % X), y = (int)((i / X) % Y), z = (int)(i / XY);
            const long gc = x + (long)X * z;
            const int gh = G.gh[gc];
            if (y <= gh) {
                stC(C, i, 0.0f, 0.0f, 0.0f, 0.0f);
                continue;
            }
            const int ghL = x > 0 ? G.gh[gc - 1] : 0, ghR = x < X - 1 ? G.gh[gc + 1] : 0;
            const int ghB = z > 0 ? G.gh[gc - X] : 0, ghF = z < Z - 1 ? G.gh[gc + X] : 0;
            const OpRow o = opRow(irho, i, x, y, z, X, Y, Z, XY, gh, ghL, ghR, ghB, ghF);
            stC(C, i, o.sR ? -o.fr : 0.0f, o.sU ? -o.fu : 0.0f, o.sF ? -o.ff : 0.0f, o.w);
        }
    }
};

// Level l + 1's operator from level l's.
struct KMgCoarsen {
    static __device__ __forceinline__ void run(VB vb, MgDims F, const MgC* __restrict__ Cf, MgDims D,
                                               MgC* __restrict__ Cc) {
        const long FXY = (long)F.X * F.Y, XY = (long)D.X * D.Y;
        for (long k = gfirst(vb); k < D.N; k += gstride(vb)) {
            const int I = (int)(k % D.X), J = (int)((k / D.X) % D.Y), K = (int)(k / XY);
            const int x0 = 2 * I, y0 = 2 * J, z0 = 2 * K;
            const int nx = min(2, F.X - x0), ny = min(2, F.Y - y0), nz = min(2, F.Z - z0);
            float sumd = 0.0f, inner = 0.0f, bx = 0.0f, by = 0.0f, bz = 0.0f;
            for (int c = 0; c < nz; c++)
                for (int b = 0; b < ny; b++)
                    for (int a = 0; a < nx; a++) {
                        const float4 f = ldC(Cf, (x0 + a) + (long)F.X * (y0 + b) + FXY * (z0 + c));
                        sumd += f.w;
                        if (a == 0 && nx == 2) inner += f.x;
                        if (b == 0 && ny == 2) inner += f.y;
                        if (c == 0 && nz == 2) inner += f.z;
                        if (a == nx - 1) bx += f.x;
                        if (b == ny - 1) by += f.y;
                        if (c == nz - 1) bz += f.z;
                    }
            // a sibling face is in both children's diagonals; the coarse
            // operator (half sums) is stored divided by 2 once more, the
            // level's power-of-two scale
            const float d = 0.25f * (sumd - 2.0f * inner);
            if (d > 0.0f) stC(Cc, k, 0.25f * bx, 0.25f * by, 0.25f * bz, d);
            else stC(Cc, k, 0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
};

// One damped Jacobi sweep xout = xin + (r - A xin) omega / d (first: from
// xin = 0, xout = r omega / d). The column's x at z - 1, z and z + 1 and the
// coupling to z - 1 ride in registers.
struct KMgSmooth {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        return nthArg<5>(a...)->done != 0;  // st
    }
    static __device__ __forceinline__ void run(VB vb, MgDims D, const MgC* __restrict__ C,
                                               const float* __restrict__ r, const float* __restrict__ xin,
                                               float* __restrict__ xout, const PcgState* st, int first,
                                               double* dpart) {
        if (st->done) return;
        const MgCol t = mgCol(vb, D);
        const int X = D.X, XY = D.X * D.Y;
        if (first) {
            if (!t.in) return;
            int i = t.x + X * t.y + XY * t.z0;
            for (int z = t.z0; z < t.z1; z++, i += XY) {
                const float w = ldCw(C, i);
                xout[i] = w > 0.0f ? r[i] * omegaOver(w) : 0.0f;
            }
            return;
        }
        if (t.z0 >= t.z1) {  // the whole block
            if (dpart && tid256() == 0) dpart[vb.id] = 0.0;
            return;
        }
        // x of the in-plane neighbours from the plane's LDS tile
        __shared__ float sX[2][6][66];
        __shared__ double sh[4];
        const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
        const MgTile g = mgTile(t, X, D.Y);
        double dz = 0.0;  // this lane's share of z.r (level 0's last sweep)
        int i = g.ib + XY * t.z0;
        float xm = t.z0 > 0 ? xin[i - XY] : 0.0f, czm = t.z0 > 0 ? ldCz(C, i - XY) : 0.0f;
        float xc = xin[i], hv = xin[g.hb + XY * t.z0];
        // a neighbour beyond the level (or a ground cell, w = 0) is dropped
        // by a select; couplings are read at offsets clamped into the level
        const bool hxR = t.x + 1 < X, hxL = t.x > 0, hyU = t.y + 1 < D.Y, hyD = t.y > 0;
        const int oL = hxL ? -1 : 0, oD = hyD ? -X : 0;
        for (int z = t.z0; z < t.z1; z++, i += XY) {
            const int buf = (z - t.z0) & 1;
            const float4 c = ldC(C, i);
            const bool hzP = z + 1 < D.Z;
            const float xpl = xin[hzP ? i + XY : i];
            const float cl = ldCx(C, i + oL), cd = ldCy(C, i + oD), rv = r[i];
            const float hvn = xin[g.hb + XY * min(z + 1, D.Z - 1)];
            sX[buf][ty + 1][tx + 1] = xc;
            if (g.hasH) sX[buf][g.hrow][g.hcol] = hv;
            __syncthreads();
            const float xr = sX[buf][ty + 1][tx + 2], xl = sX[buf][ty + 1][tx];
            const float xu = sX[buf][ty + 2][tx + 1], xd = sX[buf][ty][tx + 1];
            const float xp = hzP ? xpl : 0.0f;
            float s = czm * xm + c.z * xp;
            s = hxR ? s + c.x * xr : s;
            s = hxL ? s + cl * xl : s;
            s = hyU ? s + c.y * xu : s;
            s = hyD ? s + cd * xd : s;
            const float v = c.w > 0.0f ? xc + (rv - (c.w * xc - s)) * omegaOver(c.w) : 0.0f;
            if (g.in) {
                xout[i] = v;
                dz += (double)(v * rv);
            }
            xm = xc;
            xc = xp;
            czm = c.z;
            hv = hvn;
        }
        if (dpart) {
            dz = blockSum(dz, sh);
            if (tid256() == 0) dpart[vb.id] = dz;
        }
    }
};

// The two pre-smoothing sweeps in one pass: the first sweep from zero gives
// a_j = omega r_j / d_j at every cell, which the second sweep's stencil
// reads at the cell and its six neighbours: xout = a + omega (r - A a) / d.
struct KMgPreSmooth {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        return nthArg<4>(a...)->done != 0;  // st
    }
    static __device__ __forceinline__ void run(VB vb, MgDims D, const MgC* __restrict__ C,
                                               const float* __restrict__ r, float* __restrict__ xout,
                                               const PcgState* st) {
        if (st->done) return;
        const MgCol t = mgCol(vb, D);
        if (t.z0 >= t.z1) return;  // the whole block
        // Each cell forms its own a (the first sweep's value) once per plane
        // and publishes it in the plane's LDS tile (two buffers, one barrier
        // per plane); its four in-plane neighbours read it there. Along z, a
        // stays in registers.
        __shared__ float sA[2][6][66];
        const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
        const int X = D.X, Y = D.Y, XY = D.X * D.Y;
        auto a0 = [&](float w, float rv) { return w > 0.0f ? rv * omegaOver(w) : 0.0f; };
        const MgTile g = mgTile(t, X, Y);
        const int hb = g.hb;
        int i = g.ib + XY * t.z0;
        float4 cm{0.0f, 0.0f, 0.0f, 0.0f}, cc = ldC(C, i);
        float am = 0.0f, ac = a0(cc.w, r[i]);
        if (t.z0 > 0) {
            cm = ldC(C, i - XY);
            am = a0(cm.w, r[i - XY]);
        }
        float hw = ldCw(C, hb + XY * t.z0), hr = r[hb + XY * t.z0];
        // a neighbour beyond the level (or a ground cell, w = 0) is dropped
        // by a select; its coupling is read at an offset clamped into the level
        const bool hxR = t.x + 1 < X, hxL = t.x > 0, hyU = t.y + 1 < Y, hyD = t.y > 0;
        const int oL = hxL ? -1 : 0, oD = hyD ? -X : 0;
        for (int z = t.z0; z < t.z1; z++, i += XY) {
            const int buf = (z - t.z0) & 1;
            const bool hzP = z + 1 < D.Z;
            const int ip = hzP ? i + XY : i;
            const float4 cpl = ldC(C, ip);
            const float rpl = r[ip], rv = r[i], cl = ldCx(C, i + oL), cd = ldCy(C, i + oD);
            const int hn = hb + XY * min(z + 1, D.Z - 1);
            const float hwn = ldCw(C, hn), hrn = r[hn];
            sA[buf][ty + 1][tx + 1] = ac;
            if (g.hasH) sA[buf][g.hrow][g.hcol] = a0(hw, hr);
            __syncthreads();
            const float aR = sA[buf][ty + 1][tx + 2], aL = sA[buf][ty + 1][tx];
            const float aU = sA[buf][ty + 2][tx + 1], aD = sA[buf][ty][tx + 1];
            const float4 cp = hzP ? cpl : float4{0.0f, 0.0f, 0.0f, 0.0f};
            const float ap = hzP ? a0(cp.w, rpl) : 0.0f;
            float s = cm.z * am + cc.z * ap;
            s = hxR ? s + cc.x * aR : s;
            s = hxL ? s + cl * aL : s;
            s = hyU ? s + cc.y * aU : s;
            s = hyD ? s + cd * aD : s;
            const float v = cc.w > 0.0f ? ac + (rv - (cc.w * ac - s)) * omegaOver(cc.w) : 0.0f;
            if (g.in) xout[i] = v;
            cm = cc;
            am = ac;
            cc = cp;
            ac = ap;
            hw = hwn;
            hr = hrn;
        }
    }
};

// The coarse correction added to x and the first post-smoothing sweep in one
// pass: the stencil reads x + (the parent's correction) at the cell and its
// six neighbours; xout = xe + omega (r - A xe) / d with xe corrected.
struct KMgProlongSmooth {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        return nthArg<7>(a...)->done != 0;  // st
    }
    static __device__ __forceinline__ void run(VB vb, MgDims F, const MgC* __restrict__ C,
                                               const float* __restrict__ r, const float* __restrict__ xin,
                                               float* __restrict__ xout, MgDims D, const float* __restrict__ xc,
                                               const PcgState* st) {
        if (st->done) return;
        const MgCol t = mgCol(vb, F);
        if (t.z0 >= t.z1) return;  // the whole block
        // the corrected x of the in-plane neighbours from the plane's LDS tile
        __shared__ float sX[2][6][66];
        const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
        const int X = F.X, XY = F.X * F.Y, DXY = D.X * D.Y;
        const MgTile g = mgTile(t, X, F.Y);
        // the parent rows of the cell and of the halo cell
        const int px = (g.ib % X) >> 1, py = (g.ib / X) >> 1, hpx = g.hx >> 1, hpy = g.hy >> 1;
        auto corr = [&](int cx, int cy, int z) { return xc[cx + D.X * cy + DXY * (z >> 1)]; };
        int i = g.ib + XY * t.z0;
        float xm = 0.0f, czm = 0.0f;
        if (t.z0 > 0) {
            xm = xin[i - XY] + corr(px, py, t.z0 - 1);
            czm = ldCz(C, i - XY);
        }
        float xe = xin[i] + corr(px, py, t.z0);
        float hv = xin[g.hb + XY * t.z0] + corr(hpx, hpy, t.z0);
        // a neighbour beyond the level (or a ground cell, w = 0) is dropped
        // by a select; couplings are read at offsets clamped into the level
        const bool hxR = t.x + 1 < X, hxL = t.x > 0, hyU = t.y + 1 < F.Y, hyD = t.y > 0;
        const int oL = hxL ? -1 : 0, oD = hyD ? -X : 0;
        for (int z = t.z0; z < t.z1; z++, i += XY) {
            const int buf = (z - t.z0) & 1;
            const float4 c = ldC(C, i);
            const bool hzP = z + 1 < F.Z;
            const float xpl = xin[hzP ? i + XY : i] + corr(px, py, hzP ? z + 1 : z);
            const float cl = ldCx(C, i + oL), cd = ldCy(C, i + oD), rv = r[i];
            const int zn = min(z + 1, F.Z - 1);
            const float hvn = xin[g.hb + XY * zn] + corr(hpx, hpy, zn);
            sX[buf][ty + 1][tx + 1] = xe;
            if (g.hasH) sX[buf][g.hrow][g.hcol] = hv;
            __syncthreads();
            const float xR = sX[buf][ty + 1][tx + 2], xL = sX[buf][ty + 1][tx];
            const float xU = sX[buf][ty + 2][tx + 1], xD = sX[buf][ty][tx + 1];
            const float xp = hzP ? xpl : 0.0f;
            float s = czm * xm + c.z * xp;
            s = hxR ? s + c.x * xR : s;
            s = hxL ? s + cl * xL : s;
            s = hyU ? s + c.y * xU : s;
            s = hyD ? s + cd * xD : s;
            const float v = c.w > 0.0f ? xe + (rv - (c.w * xe - s)) * omegaOver(c.w) : 0.0f;
            if (g.in) xout[i] = v;
            xm = xe;
            xe = xp;
            czm = c.z;
            hv = hvn;
        }
    }
};

// The residual r - A x of level l summed over each coarse cell's children
// into level l + 1's right-hand side, in one pass over the fine level: a
// block is 64 x 4 fine columns (32 x 2 coarse) marching the fine planes of
// a coarse z-chunk; each thread forms its cells' residuals (the column's x
// at z - 1, z, z + 1 in registers, the in-plane neighbours' x from the
// plane's LDS tile), sums each pair of planes, adds its x neighbour's sum by
// a lane exchange and its y neighbour's through LDS.
struct KMgRestrict {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        return nthArg<6>(a...)->done != 0;  // st
    }
    static __device__ __forceinline__ void run(VB vb, MgDims F, const MgC* __restrict__ C,
                                               const float* __restrict__ r, const float* __restrict__ x, MgDims D,
                                               float* __restrict__ rc, const PcgState* st) {
        if (st->done) return;
        __shared__ float sh[4][64];
        const unsigned gx = (unsigned)(F.X + 63) >> 6, gy = (unsigned)(F.Y + 3) >> 2;
        const unsigned bx = vb.id % gx, by = (vb.id / gx) % gy, bz = vb.id / (gx * gy);
        const int tx = threadIdx.x, ty = threadIdx.y;
        const int fx = (int)bx * 64 + tx, fy = (int)by * 4 + ty;
        const int zs = 2 * (int)bz * D.zc, ze = min(F.Z, 2 * ((int)bz + 1) * D.zc);
        if (zs >= ze) return;  // the whole block
        // x of the in-plane neighbours from the plane's LDS tile
        __shared__ float sX[2][6][66];
        const int X = F.X, XY = F.X * F.Y;
        const MgTile g = mgTile(MgCol{fx, fy, zs, ze, true}, X, F.Y);
        int i = g.ib + XY * zs;
        float xm = 0.0f, czm = 0.0f, xc = x[i], hv = x[g.hb + XY * zs];
        if (zs > 0) {
            xm = x[i - XY];
            czm = ldCz(C, i - XY);
        }
        const int cx = fx >> 1, cy = fy >> 1, DXY = D.X * D.Y;
        float acc = 0.0f;
        // a neighbour beyond the level (or a ground cell, w = 0) is dropped
        // by a select; couplings are read at offsets clamped into the level;
        // lanes past the level add nothing
        const bool hxR = fx + 1 < X, hxL = fx > 0, hyU = fy + 1 < F.Y, hyD = fy > 0;
        const int oL = hxL ? -1 : 0, oD = hyD ? -X : 0;
        for (int z = zs; z < ze; z++, i += XY) {
            const int buf = (z - zs) & 1;
            const float4 c = ldC(C, i);
            const bool hzP = z + 1 < F.Z;
            const float xpl = x[hzP ? i + XY : i];
            const float cl = ldCx(C, i + oL), cd = ldCy(C, i + oD), rv = r[i];
            const float hvn = x[g.hb + XY * min(z + 1, F.Z - 1)];
            sX[buf][ty + 1][tx + 1] = xc;
            if (g.hasH) sX[buf][g.hrow][g.hcol] = hv;
            __syncthreads();
            const float xr = sX[buf][ty + 1][tx + 2], xl = sX[buf][ty + 1][tx];
            const float xu = sX[buf][ty + 2][tx + 1], xd = sX[buf][ty][tx + 1];
            const float xp = hzP ? xpl : 0.0f;
            float s = czm * xm + c.z * xp;
            s = hxR ? s + c.x * xr : s;
            s = hxL ? s + cl * xl : s;
            s = hyU ? s + c.y * xu : s;
            s = hyD ? s + cd * xd : s;
            acc = g.in && c.w > 0.0f ? acc + (rv - (c.w * xc - s)) : acc;
            xm = xc;
            xc = xp;
            czm = c.z;
            hv = hvn;
            if ((z & 1) || z == F.Z - 1) {  // a coarse plane is complete (uniform over the block)
                const float v = acc + __shfl_xor(acc, 1);
                sh[ty][tx] = v;
                __syncthreads();
                // the coarse level's right-hand side is divided by 2 once more
                if (!(ty & 1) && !(tx & 1) && cx < D.X && cy < D.Y)
                    rc[cx + D.X * cy + DXY * (z >> 1)] = 0.5f * (v + sh[ty + 1][tx]);
                __syncthreads();
                acc = 0.0f;
            }
        }
    }
};

// The coarsest level: kCoarseSweeps damped Jacobi sweeps from zero, in LDS,
// by one block; the result is written to x.
struct KMgCoarsest {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        return nthArg<4>(a...)->done != 0;  // st
    }
    static __device__ __forceinline__ void run(VB vb, MgDims D, const MgC* __restrict__ C,
                                               const float* __restrict__ r, float* __restrict__ x,
                                               const PcgState* st) {
        (void)vb;
        if (st->done) return;
        __shared__ float4 sC[kMgCoarsest];
        __shared__ float sr[kMgCoarsest], sx[2][kMgCoarsest];
        const int t = tid256();
        const int n = (int)D.N;  // at most kMgCoarsest (checked on the host)
        for (int i = t; i < n; i += 256) {
            sC[i] = ldC(C, i);
            sr[i] = r[i];
        }
        __syncthreads();
        const int X = D.X, Y = D.Y, Z = D.Z, XY = X * Y;
        for (int i = t; i < n; i += 256) sx[0][i] = sC[i].w > 0.0f ? sr[i] * omegaOver(sC[i].w) : 0.0f;
        __syncthreads();
        for (int k = 1; k < kCoarseSweeps; k++) {
            const float* xin = sx[(k - 1) & 1];
            float* xo = sx[k & 1];
            for (int i = t; i < n; i += 256) {
                const float4 c = sC[i];
                float v = 0.0f;
                if (c.w > 0.0f) {
                    const int xx = i % X, yy = (i / X) % Y, zz = i / XY;
                    float s = 0.0f;
                    if (xx + 1 < X) s += c.x * xin[i + 1];
                    if (xx > 0) s += sC[i - 1].x * xin[i - 1];
                    if (yy + 1 < Y) s += c.y * xin[i + X];
                    if (yy > 0) s += sC[i - X].y * xin[i - X];
                    if (zz + 1 < Z) s += c.z * xin[i + XY];
                    if (zz > 0) s += sC[i - XY].z * xin[i - XY];
                    v = xin[i] + (sr[i] - (c.w * xin[i] - s)) * omegaOver(c.w);
                }
                xo[i] = v;
            }
            __syncthreads();
        }
        for (int i = t; i < n; i += 256) x[i] = sx[(kCoarseSweeps - 1) & 1][i];
    }
};

// z.r partials of the preconditioned residual for a grid of one level (no
// level-0 sweep to form them): n blocks, block b summing cells
// [b chunk, (b + 1) chunk) of float products in double.
struct KMgDot {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        return nthArg<4>(a...)->done != 0;  // st
    }
    static __device__ __forceinline__ void run(VB vb, long N, const float* __restrict__ z,
                                               const float* __restrict__ r, double* part, const PcgState* st) {
        __shared__ double sh[4];
        if (st->done) return;
        const long chunk = (N + vb.n - 1) / vb.n;
        const long a = (long)vb.id * chunk, b = min(N, a + chunk);
        double s = 0.0;
        for (long i = a + tid256(); i < b; i += 256) s += (double)(z[i] * r[i]);
        s = blockSum(s, sh);
        if (tid256() == 0) part[vb.id] = s;
    }
};

// A level's dims with its z-chunk (about 2048 tiles when the level has the
// planes for it), and its tile count.
// In an ensemble step (ens) the chunks are 16 planes: the members together
// fill the device (no level kernel's values depend on its chunking).
MgDims dims(const MgLevel& L, bool ens = false) {
    MgDims d{L.X, L.Y, L.Z, L.N, 1};
    if (L.N <= 0) return d;
    const long cols = (long)((L.X + 63) / 64) * ((L.Y + 3) / 4);
    const long nzc = ens ? (L.Z + 15) / 16 : std::max(1L, std::min((long)L.Z, (2048 + cols - 1) / cols));
    d.zc = (int)((L.Z + nzc - 1) / nzc);
    return d;
}
dim3 colTiles(const MgDims& d, bool on) {
    if (!on || d.N <= 0) return dim3(0);
    const long cols = (long)((d.X + 63) / 64) * ((d.Y + 3) / 4);
    return dim3((unsigned)(cols * ((d.Z + d.zc - 1) / d.zc)));
}
// Tiles of a restriction: the fine level's columns times the coarse z-chunks.
dim3 restrictTiles(const MgDims& f, const MgDims& c, bool on) {
    if (!on || c.N <= 0) return dim3(0);
    const long cols = (long)((f.X + 63) / 64) * ((f.Y + 3) / 4);
    return dim3((unsigned)(cols * ((c.Z + c.zc - 1) / c.zc)));
}
// Tiles of an elementwise pass over n cells (none past the grid's depth).
dim3 tiles(long n, bool on) {
    if (!on || n <= 0) return dim3(0);
    return dim3((unsigned)std::min<long>((n + 255) / 256, 8192));
}
constexpr dim3 kB3(64, 4, 1);

}  // namespace

hipError_t launch_mg_setup(EnvState& e) {
    const MgLevel* L = e.mg;
    hipError_t r = klaunch<KMgCoef0>(e, tiles(e.N, true), dim3(256), 0, geom(e), (const float*)e.irho, L[0].C);
    for (int l = 0; l + 1 < kMgLevels && r == hipSuccess; l++) {
        const bool on = l + 1 < e.mg_depth;
        r = klaunch<KMgCoarsen>(e, tiles(L[l + 1].N, on), dim3(256), 0, dims(L[l]), (const MgC*)L[l].C,
                                dims(L[l + 1]), L[l + 1].C);
    }
    return r;
}

hipError_t launch_mg_vcycle(EnvState& e, int start) {
    const MgLevel* L = e.mg;
    const bool ens = e.rec != nullptr;
    const PcgState* st = &e.sc->st[1];
    const int D = e.mg_depth;
    hipError_t r = hipSuccess;
    auto go = [&](hipError_t x) {
        if (r == hipSuccess) r = x;
    };
    // the z.r partials of the next search direction: one per tile of level
    // 0's last sweep, which forms z; a grid of one level (no level-0 sweep)
    // reduces them in a pass of its own
    const size_t M = e.maxBlocks;
    double* part = e.partials + (start ? PART_START : PART_BSUM) * M;
    const unsigned t0 = colTiles(dims(L[0], ens), D > 1).x;
    const bool fused = D > 1 && t0 <= M;
    // two Jacobi sweeps before (one fused pass into xb) and after (the
    // coarse correction and the first sweep fused into xa, the second back
    // into xb, each level's result)
    for (int l = 0; l + 1 < kMgLevels; l++) {
        const bool on = l + 1 < D;
        const MgDims f = dims(L[l], ens), c = dims(L[l + 1], ens);
        go(klaunch<KMgPreSmooth>(e, colTiles(f, on), kB3, 0, f, (const MgC*)L[l].C, (const float*)L[l].r, L[l].xb,
                                 st));
        go(klaunch<KMgRestrict>(e, restrictTiles(f, c, on), kB3, 0, f, (const MgC*)L[l].C, (const float*)L[l].r,
                                (const float*)L[l].xb, c, L[l + 1].r, st));
    }
    const MgLevel& B = L[D - 1];
    go(klaunch<KMgCoarsest>(e, dim3(1), dim3(256), 0, dims(B), (const MgC*)B.C, (const float*)B.r, B.xb, st));
    for (int l = kMgLevels - 2; l >= 0; l--) {
        const bool on = l + 1 < D;
        const MgDims f = dims(L[l], ens), c = dims(L[l + 1], ens);
        const dim3 g = colTiles(f, on);
        go(klaunch<KMgProlongSmooth>(e, g, kB3, 0, f, (const MgC*)L[l].C, (const float*)L[l].r,
                                     (const float*)L[l].xb, L[l].xa, c, (const float*)L[l + 1].xb, st));
        go(klaunch<KMgSmooth>(e, g, kB3, 0, f, (const MgC*)L[l].C, (const float*)L[l].r, (const float*)L[l].xa,
                              L[l].xb, st, 0, l == 0 && fused ? part : (double*)nullptr));
    }
    // the pass of its own has no tiles when level 0's sweep formed the
    // partials (every grid records the same launch sequence)
    e.mg_dot_n = fused ? t0 : (start ? e.gZ.nblocks : e.gB.nblocks);
    go(klaunch<KMgDot>(e, dim3(fused ? 0u : e.mg_dot_n), dim3(256), 0, e.N, (const float*)e.pcg[1],
                       (const float*)e.pcg[0], part, (const PcgState*)&e.sc->st[1]));
    return r;
}

}  // namespace skt
