// HIP kernels of the skewT cloud step for gfx950 (MI355X): diffusion, PPM
// advection, the pressure solve and the reductions that steer the step.
//
// Layout: every voxel field is one float array in HBM, index x + X*(y + Y*z)
// (x fastest, y vertical), exactly the reference's getIdx (Game/utils.cuh:8-11)
// so device pointers are interchangeable with the reference's.
//
// Launch shape: 256-thread blocks own a 64 (x) x 4 (y) tile; a wavefront
// covers 64 consecutive floats of a row (256-byte coalesced lines) and the
// block walks a run of zc planes (about 2048 blocks on the large domains,
// which fills the 256 CUs and leaves one reduction partial per block).
// The kernels are bound by memory latency unless each wave keeps many loads
// in flight, so every z-marching kernel works on groups of ZB planes: it
// first issues all loads of the group (stencil neighbours at clamped, always
// valid addresses, plus the ground heights that decide the boundary rule),
// then applies the boundary conditions with selects and computes. Column
// values along z stay in registers between planes.
//
// Reductions are deterministic: per-block partials (FP64), summed in a fixed
// order. The scalar steps of the pressure solve (alpha, beta, the stop test)
// run in the prologue of the kernel that consumes them: every block reduces
// the previous kernel's partials to the same value and block 0 stores the
// updated state in the other of two slots.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <type_traits>
#include "skt_internal.h"
#include "skt_launch.h"
#include "skt_reduce.h"

namespace skt {

int blocks3d(const EnvState& e) {
    dim3 g = grid3(e);
    return (int)(g.x * g.y * g.z);
}
Geom geom(const EnvState& e) { return Geom{e.X, e.Y, e.Z, e.N, e.gh, e.ghm1, e.ghm2}; }
static long zcols(const EnvState& e, int rows = 4) { return (long)((e.X + 63) / 64) * ((e.Y + rows - 1) / rows); }
// nzc chunks of equal length along z (the last one shorter), tiles of 64 x rows
static ZGrid zgridChunks(const EnvState& e, long nzc, int rows = 4) {
    nzc = std::max(1L, std::min((long)e.Z, nzc));
    const int zc = (int)((e.Z + nzc - 1) / nzc);
    nzc = (e.Z + zc - 1) / zc;
    const unsigned nb = (unsigned)(zcols(e, rows) * nzc);
    return ZGrid{dim3(nb), zc, nb};
}
ZGrid zgrid(const EnvState& e) { return e.gZ; }
// Chunks of the stencil passes in an ensemble step: the members together
// fill the device, so each member's tiles take 16 planes where its own
// launch splits z finer to fill the device alone (a pure stencil gives the
// same values for any chunking).
constexpr int kEnsembleZc = 16;
static long ensembleChunks(const EnvState& e) { return (e.Z + kEnsembleZc - 1) / kEnsembleZc; }

void setup_grids(EnvState& e, ResidentBlocks rb) {
    const long cols = zcols(e), cap = e.cfg.max_blocks;
    // about 2048 blocks (at least one chunk per tile column), or at most cap
    e.gZ = zgridChunks(e, cap > 0 ? cap / cols : (2048 + cols - 1) / cols);
    // the pressure operator and a diffusion sweep (per field): one round of
    // resident blocks
    auto round = [&](long r) {
        long a = r > 0 ? r : 2048;
        if (cap > 0) a = std::min(a, cap);
        return zgridChunks(e, a / cols);
    };
    // The pressure operator takes at most five blocks per CU (of the seven it
    // could hold): on MI355X its launch then measured 4 % faster at 256 x 128
    // x 256 and 320 x 128 x 256, and no slower at the other shapes timed.
    e.gA = round(rb.cus > 0 ? std::min(rb.pcgA, 5L * rb.cus) : rb.pcgA);
    // diffusion: tiles of 64 x 8; a chunk is at most 64 planes (one bit per
    // plane marks the cells next to ground)
    {
        long a = rb.diffuse > 0 ? rb.diffuse : 2048;
        if (cap > 0) a = std::min(a, cap);
        e.gD = zgridChunks(e, a / zcols(e, 8), 8);
        if (e.gD.zc > 64) e.gD = zgridChunks(e, (e.Z + 63) / 64, 8);
    }
    // the flat walks: about 2048 blocks of whole rows
    const long rows = (long)e.Y * e.Z, nb = std::max(1L, std::min(rows, cap > 0 ? std::min(2048L, cap) : 2048L));
    const long rpb = (rows + nb - 1) / nb;
    e.gB = FGrid{(unsigned)((rows + rpb - 1) / rpb), (unsigned)rpb};
}

// Boundary value of a stencil read (the extv rule): ground condition if the
// clamped position is ground, the sides condition if the position lies
// outside the domain (the top face too), else the raw value.
__device__ __forceinline__ float nbv(float raw, bool out, bool gnd, float pv, Bounds b) {
    return gnd ? bcv(b.ground, raw, pv) : (out ? bcv(b.sides, raw, pv) : raw);
}
// The same rule for the unit field (profile value 1).
__device__ __forceinline__ float unitv(bool out, bool gnd, Bounds b) {
    return gnd ? (b.ground == BC_DIRICHLET ? 0.0f : 1.0f) : (out ? (b.sides == BC_DIRICHLET ? 0.0f : 1.0f) : 1.0f);
}

// -------------------------------------------------------------- utility ---
KL void k_fill(float* a, float v, long n) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) a[i] = v;
}
hipError_t launch_fill(EnvState& e, float* a, float v, long n) {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)std::min<long>((n + 255) / 256, 65536)), dim3(256), 0, e.stream, a, v, n);
    return hipGetLastError();
}
template <bool MIX>
struct KCopySky {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, const float* __restrict__ s, float* __restrict__ d,
                                               Store ss, Store ds) {
        CELL_SETUP_V;
        (void)bid;
        if (!inb || isGround(G, x, y, z)) return;
        const long i = gidx(G, x, y, z);
        stS<MIX>(d, i, ldS<MIX>(s, i, ss), ds);
    }
};
hipError_t launch_copy_sky(EnvState& e, const float* src, float* dst, Store ss, Store ds) {
    if (ss.fmt != FMT_F32 || ds.fmt != FMT_F32)
        return klaunch<KCopySky<true>>(e, grid3(e), B3, 0, geom(e), src, dst, ss, ds);
    return klaunch<KCopySky<false>>(e, grid3(e), B3, 0, geom(e), src, dst, ss, ds);
}
// Plain streaming copy (float4 loads/stores): the measured bandwidth ceiling.
KL void k_stream_copy(const float4* __restrict__ s, float4* __restrict__ d, long n4) {
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n4; i += (long)gridDim.x * 256) d[i] = s[i];
}
hipError_t launch_stream_copy(EnvState& e, const float* src, float* dst, long n) {
    const long n4 = n / 4;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)std::min<long>((n4 + 255) / 256, 8192)), dim3(256), 0, e.stream,
                       (const float4*)src, (float4*)dst, n4);
    return hipGetLastError();
}

// Raw buffer access: the resource (base address) and a wave-uniform byte
// offset sit in scalar registers, the lane supplies a 32-bit byte offset, so
// a lane keeps one VGPR per offset instead of a 64-bit address per array, and
// moving to the next plane is one scalar add. The records limit is the
// largest (no range check), so every offset must lie inside the array.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc arrayRsrc(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, -1, 0x00020000);
}
__device__ __forceinline__ float bldf(Rsrc r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ int bldi(Rsrc r, unsigned voff, unsigned soff) {
    return (int)__builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void bstf(Rsrc r, unsigned voff, unsigned soff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ unsigned short bldh(Rsrc r, unsigned voff, unsigned soff) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void bsth(Rsrc r, unsigned voff, unsigned soff, unsigned short v) {
    __builtin_amdgcn_raw_buffer_store_b16(v, r, (int)voff, (int)soff, 0);
}

// LDS tile of one plane for the 64 x 4 block: rows y0-1 .. y0+4 and columns
// x0-1 .. x0+64 (the block's cells plus a one-cell halo). Each cell's value
// is read from HBM once per plane: by its own thread, plus the halo rows by
// waves 0 and 1 and the halo columns by lanes 0-7 of wave 2. Two buffers and
// one barrier per plane: a wave can only overwrite a buffer after every wave
// has passed the next plane's barrier, i.e. finished reading it.
constexpr int TH = 6, TW = 66;

// ------------------------------------------------------------- diffusion ---
// One Jacobi sweep = the red+black pair of diffuseRedBlack
// (Game/kernelSky.cu:43-115; both colours read 'input', quirk A3):
// out = (in + k (l + r + u + d + f + b)) / (1 + 6k) on sky cells.
// One launch sweeps up to SKT_BATCH independent fields (the grid is the
// z-march grid repeated per field). Each input value is read from HBM once:
// a plane's values go through the LDS tile (in-plane neighbours; positions
// outside the domain hold the clamped cell's value, which is what the
// boundary rules start from), along z they stay in a ring of four register
// slots (read three planes ahead, four planes per loop trip, no moves).
// Cells with y >= ghm1 + 2 (ghm1: largest ground height of the 3x3 columns
// around) have no ground neighbour, so only the domain faces need a boundary
// value; the other sky cells are marked and done after the z march, with
// every neighbour tested for ground.
// One field per block. A block owns a 64 (x) x 8 (y) tile: each thread
// sweeps two cells of its column, rows yb + ty and yb + ty + 4, so the
// per-plane scalar work (offsets, loop, barrier) and the ground-height loads
// of a column serve two cells, and the tile's halo is 144 cells for 512.
// Blocks of two or three fields (sharing the ground loads and barriers)
// needed over 100 VGPRs and ran slower at half the occupancy; direct
// neighbour loads without the tile ran slower too.
constexpr int DTH = 10;  // tile rows yb - 1 .. yb + 8
template <bool MIX>
struct KDiffuse {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int zc, unsigned nbf, const DiffBatch& db, float k, const float* __restrict__ gT,
                      const float* __restrict__ gP, const float* __restrict__ p) {
        const unsigned lin = xcdRemap(vb.id, vb.n);
        const DiffField& fd = db.f[lin / nbf];
        const unsigned bl = lin % nbf;
        const int X = G.X, Y = G.Y, Z = G.Z;
        const int gx = (X + 63) >> 6, gy = (Y + 7) >> 3;
        const int tx = threadIdx.x, ty = threadIdx.y;
        const int x = (int)(bl % (unsigned)gx) * 64 + tx;
        const int yb = (int)((bl / (unsigned)gx) % (unsigned)gy) * 8;
        const int z0 = (int)(bl / (unsigned)(gx * gy)) * zc, z1 = min(Z, z0 + zc);
        if (z0 >= z1) return;  // whole block
        __shared__ float tV[2][DTH][TW];
        const long XY = (long)X * Y;
        const int ya = yb + ty, yb4 = yb + ty + 4;  // the thread's two rows
        const bool inA = x < X && ya < Y, inB = x < X && yb4 < Y;
        const int xc = min(x, X - 1), yca = min(ya, Y - 1), ycb = min(yb4, Y - 1);
        const unsigned obA = 4u * (unsigned)(xc + X * yca), obB = 4u * (unsigned)(xc + X * ycb), gb = 4u * (unsigned)xc;
        // halo cells: the row below the tile (wave 0), the row above it (wave 1),
        // the columns left and right of it (lanes 0-7 and 8-15 of wave 2, one row
        // each); each reads its position clamped into the domain
        const int x0 = x - tx;
        int hx = x, hy = yb - 1, hrow = 0, hcol = tx + 1;
        if (ty == 1) hy = yb + 8, hrow = 9;
        if (ty == 2) hx = tx < 8 ? x0 - 1 : x0 + 64, hy = yb + (tx & 7), hrow = 1 + (tx & 7), hcol = tx < 8 ? 0 : 65;
        const bool hasH = ty < 2 || (ty == 2 && tx < 16);
        const unsigned hb = 4u * (unsigned)(min(max(hx, 0), X - 1) + X * min(max(hy, 0), Y - 1));
        // the same loads in every plane, as in the pressure operator
        const unsigned hbL = hasH ? hb : obA;
        const int zb = max(z0 - 1, 0);
        const unsigned PB = 4u * (unsigned)XY, RB = 4u * (unsigned)X;
        // 16-bit storage (the mixed form's tracers): half the byte offsets
        const bool in16 = MIX && fd.sin.fmt != FMT_F32, out16 = MIX && fd.sout.fmt != FMT_F32;
        const Rsrc rV = arrayRsrc(static_cast<const char*>(static_cast<const void*>(fd.in)) + (in16 ? 2 : 4) * XY * zb);
        const Rsrc wV = arrayRsrc(static_cast<char*>(static_cast<void*>(fd.out)) + (out16 ? 2 : 4) * XY * zb);
        auto ldV = [&](unsigned off, unsigned soff) __attribute__((always_inline)) -> float {
            if constexpr (MIX) {
                if (in16) return fromH16(bldh(rV, off >> 1, soff >> 1), fd.sin);
            }
            return bldf(rV, off, soff);
        };
        auto stV = [&](unsigned off, unsigned soff, float v) __attribute__((always_inline)) {
            if constexpr (MIX) {
                if (out16) {
                    bsth(wV, off >> 1, soff >> 1, toH16(v, fd.sout));
                    return;
                }
            }
            bstf(wV, off, soff, v);
        };
        const Bounds b = fd.b;
        auto prof = [&](int yy) { return fd.prof ? fd.prof[yy] : 0.0f; };
        const float pvA = prof(yca), pvuA = prof(min(yca + 1, Y - 1));
        const float pvB = prof(ycb), pvuB = prof(min(ycb + 1, Y - 1));
        const Rsrc rG = arrayRsrc(G.gh + (long)X * zb), rM = arrayRsrc(G.ghm1 + (long)X * zb);
        auto pOff = [&](int pl) { return PB * (unsigned)(pl - zb); };
        auto gOff = [&](int pl) { return RB * (unsigned)(pl - zb); };
        const float den = 1 + 6 * k;
        struct Own {
            float a, b;  // the two rows
            int gm, gf;  // ghm1 of the plane, ground height of the next plane
        };
        auto loadOwn = [&](int pl, Own& o) __attribute__((always_inline)) {
            o.a = ldV(obA, pOff(pl));
            o.b = ldV(obB, pOff(pl));
            o.gm = bldi(rM, gb, gOff(pl));
            o.gf = bldi(rG, gb, gOff(min(pl + 1, Z - 1)));
        };
        // plane below the chunk (clamped: plane 0 for the first chunk)
        float vbA = ldV(obA, pOff(zb)), vbB = ldV(obB, pOff(zb));
        int ghc = bldi(rG, gb, gOff(z0));
        Own o0{}, o1{}, o2{}, o3{};
        float h0 = 0.0f, h1 = 0.0f;
        // prefetches past the chunk read the next chunk's first planes: measured
        // faster than clamping them to the chunk's last plane
        const int zlast = Z - 1;
        loadOwn(z0, o0);
        loadOwn(min(z0 + 1, zlast), o1);
        loadOwn(min(z0 + 2, zlast), o2);
        h0 = ldV(hbL, pOff(z0));
        // planes (bit z - z0) whose cell of row a / row b is next to ground
        unsigned long long slowA = 0, slowB = 0;
        const bool oL = x == 0, oR = x == X - 1;
        // fast sweep of one cell: tile row r, this plane's value vc, the front
        // plane's F (this cell's value on the last plane), the back plane's vb
        auto sweep = [&](int buf, int r, int z, int y, float vc, float F, float vb, float pv, float pvu, unsigned ob)
            __attribute__((always_inline)) {
            const float L = tV[buf][r][tx], R = tV[buf][r][tx + 2];
            const float D = tV[buf][r - 1][tx + 1], U = tV[buf][r + 1][tx + 1];
            const bool oU = y == Y - 1, oF = z == Z - 1, oB = z == 0;
            const float l = oL ? bcv(b.sides, L, pv) : L, rr = oR ? bcv(b.sides, R, pv) : R;
            const float u = oU ? bcv(b.sides, U, pvu) : U;
            const float fr = oF ? bcv(b.sides, F, pv) : F, bb = oB ? bcv(b.sides, vb, pv) : vb;
            stV(ob, pOff(z), (vc + k * (l + rr + u + D + fr + bb)) / den);
        };
        auto step = [&](int z, int buf, Own& oc, Own& of, Own& o3n, float& hc, float& hn) __attribute__((always_inline)) {
            // loads complete in issue order: the halo (used next plane) goes first,
            // so waiting for it does not wait for the plane read three ahead
            hn = ldV(hbL, pOff(min(z + 1, zlast)));
            loadOwn(min(z + 3, zlast), o3n);
            tV[buf][ty + 1][tx + 1] = oc.a;
            tV[buf][ty + 5][tx + 1] = oc.b;
            if (hasH) tV[buf][hrow][hcol] = hc;
            __syncthreads();
            const bool inZ = z + 1 < Z;
            if (inA && ya > ghc) {
                if (ya >= oc.gm + 2) sweep(buf, ty + 1, z, ya, oc.a, inZ ? of.a : oc.a, vbA, pvA, pvuA, obA);
                else slowA |= 1ull << (z - z0);
            }
            if (inB && yb4 > ghc) {
                if (yb4 >= oc.gm + 2) sweep(buf, ty + 5, z, yb4, oc.b, inZ ? of.b : oc.b, vbB, pvB, pvuB, obB);
                else slowB |= 1ull << (z - z0);
            }
            vbA = oc.a;
            vbB = oc.b;
            ghc = oc.gf;
        };
        for (int zg = z0; zg < z1; zg += 4) {
            step(zg, 0, o0, o1, o3, h0, h1);
            if (zg + 1 < z1) step(zg + 1, 1, o1, o2, o0, h1, h0);
            if (zg + 2 < z1) step(zg + 2, 0, o2, o3, o1, h0, h1);
            if (zg + 3 < z1) step(zg + 3, 1, o3, o0, o2, h1, h0);
        }
        // Sky cells next to ground (marked in slowA / slowB, one bit per plane of
        // the chunk): the sweep reads only 'in', so they can follow the fast
        // cells. The block's four waves share them: wave w takes the marked planes
        // z0 + w, z0 + w + 4, ... of all eight rows of its columns, so a row of
        // marked cells (all of them in one wave) does not hold the block for the
        // whole chunk's length. Each neighbour is tested and takes the ground or sides condition
        // (nbv); for the potential temperature a ground neighbour reads the
        // potential temperature of the ground temperature at this cell's pressure
        // (:95-100). Neighbour positions are clamped into the domain, as in the tile.
        const int xl = max(xc - 1, 0), xr = min(xc + 1, X - 1);
        auto in = [&](long i) { return ldS<MIX>(fd.in, i, fd.sin); };
        auto slowCells = [&](unsigned long long slow, int y) {
            const int yd = max(y - 1, 0), yu = min(y + 1, Y - 1);
            const float pv = prof(y), pvd = prof(yd), pvu = prof(yu);
            while (slow) {
                const int z = z0 + __builtin_ctzll(slow);
                slow &= slow - 1;
                const long pz = XY * z;
                const int zf = min(z + 1, Z - 1), zbk = max(z - 1, 0);
                const int* gz = G.gh + X * z;
                const int gc = gz[xc], gf = G.gh[xc + X * zf], gbk = G.gh[xc + X * zbk];
                const bool oU = y == Y - 1, oF = z == Z - 1, oB = z == 0;
                // a position outside the domain clamps onto this sky cell, so the
                // ground test at the clamped column is the in-domain test
                const bool gL = y <= gz[xl], gR = y <= gz[xr], gD = y - 1 <= gc;
                const bool gF = !oF && y <= gf, gB = !oB && y <= gbk;
                const float vc = in(pz + xc + X * y);
                const float L = in(pz + xl + X * y), R = in(pz + xr + X * y);
                const float D = in(pz + xc + X * yd), U = in(pz + xc + X * yu);
                const float F = in(XY * zf + xc + X * y), Bk = in(XY * zbk + xc + X * y);
                float l = nbv(L, oL, gL, pv, b), r = nbv(R, oR, gR, pv, b);
                float d = nbv(D, false, gD, pvd, b), u = nbv(U, oU, false, pvu, b);
                float fr = nbv(F, oF, gF, pv, b), bb = nbv(Bk, oB, gB, pv, b);
                if (fd.type == 0 && (gL | gR | gD | gF | gB)) {
                    const float pc = p[pz + xc + X * y];
                    auto tg = [&](int col) { return d_potTemp(gT[col] - 273.15f, pc, gP[col]) + 273.15f; };
                    if (gD) d = tg(x + X * z);
                    if (gL) l = tg(xl + X * z);
                    if (gR) r = tg(xr + X * z);
                    if (gF) fr = tg(x + X * (z + 1));
                    if (gB) bb = tg(x + X * (z - 1));
                }
                stS<MIX>(fd.out, pz + xc + X * y, (vc + k * (l + r + u + d + fr + bb)) / den, fd.sout);
            }
        };
        __shared__ unsigned long long sMask[8][64];
        sMask[ty][tx] = slowA;
        sMask[ty + 4][tx] = slowB;
        __syncthreads();
        const unsigned long long mine = 0x1111111111111111ull << ty;
        for (int r = 0; r < 8; r++) {
            const unsigned long long m = sMask[r][tx] & mine;
            if (m) slowCells(m, yb + r);
        }
    }
};
// A batch with a field stored in 16 bits needs the kernels' 16-bit form; a
// batch of FP32 fields (in the mixed form: every field but the hydrometeors,
// and the hydrometeors' sweeps between their first and last) runs the FP32
// form, whose loads and stores are the same for FP32 storage without the
// format tests and conversions.
static bool has16(const DiffBatch& db) {
    for (int j = 0; j < db.n; j++)
        if (db.f[j].sin.fmt != FMT_F32 || db.f[j].sout.fmt != FMT_F32) return true;
    return false;
}
static bool has16(const PpmBatch& pb) {
    for (int j = 0; j < pb.n; j++)
        if (pb.f[j].sin.fmt != FMT_F32 || pb.f[j].sout.fmt != FMT_F32 ||
            (pb.f[j].gsrc && pb.f[j].sg.fmt != FMT_F32))
            return true;
    return false;
}

hipError_t launch_diffuse_batch(EnvState& e, const DiffBatch& db, float k) {
    const ZGrid zg = e.rec ? zgridChunks(e, ensembleChunks(e), 8) : e.gD;
    if ((double)(zg.zc + 3) * (double)e.X * (double)e.Y * 4.0 >= 4294967296.0) return hipErrorInvalidValue;
    if (e.mix && has16(db))
        return klaunch<KDiffuse<true>>(e, dim3(zg.nblocks * db.n), B3, 0, geom(e), zg.zc, zg.nblocks, db, k,
                                       e.f[SKT_GROUND_T], e.f[SKT_GROUND_P], e.f[SKT_PRESSURE]);
    return klaunch<KDiffuse<false>>(e, dim3(zg.nblocks * db.n), B3, 0, geom(e), zg.zc, zg.nblocks, db, k,
                                    e.f[SKT_GROUND_T], e.f[SKT_GROUND_P], e.f[SKT_PRESSURE]);
}

// ------------------------------------------------------------- advection ---
// getMaxDivergence over u, v, w (Game/kernelSky.cu:910-953; the three calls of
// Game/environment.cu:720-722 share one maximum): |a| at z=0 and the signed
// value elsewhere (mode 0, quirk A6) or |a| everywhere (mode 1).
constexpr int ZB_MAXV = 4;
template <int ZB>
struct KMaxvel {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int zc, const float* __restrict__ u, const float* __restrict__ v, const float* __restrict__ w,
                     int mode, double* part) {
        ZM_SETUP_V;
        __shared__ double sh[4];
        double m = 0.0;
        if (inb) {
            const long XY = (long)G.X * G.Y;
            const int rc = x + G.X * y;
            for (int zg = z0; zg < z1; zg += ZB) {
                float a0[ZB], a1[ZB], a2[ZB];
                int gh[ZB];
    #pragma unroll
                for (int j = 0; j < ZB; j++) {
                    const int z = zg + j;
                    if (z < z1) {
                        const long i = XY * z + rc;
                        a0[j] = u[i];
                        a1[j] = v[i];
                        a2[j] = w[i];
                        gh[j] = G.gh[x + G.X * z];
                    }
                }
    #pragma unroll
                for (int j = 0; j < ZB; j++) {
                    const int z = zg + j;
                    if (z < z1 && y > gh[j]) {
                        const bool ab = mode != 0 || z == 0;
                        m = fmax(m, (double)(ab ? fabsf(a0[j]) : a0[j]));
                        m = fmax(m, (double)(ab ? fabsf(a1[j]) : a1[j]));
                        m = fmax(m, (double)(ab ? fabsf(a2[j]) : a2[j]));
                    }
                }
            }
        }
        m = blockMax(m, sh);
        if (tid256() == 0) part[bid] = m;
    }
};
// Substep count and length of advectPPMWGPU (Game/environment.cu:726-730).
struct KMaxvelFinalize {
    static __device__ __forceinline__ void run(VB vb, const double* __restrict__ part, unsigned n, Scalars* sc,
                                               float dt, float dx, int maxsub, int cap) {
        __shared__ double sh[4];
        double vmax = 0.0;
        for (unsigned j = tid256(); j < n; j += 256) vmax = fmax(vmax, part[j]);
        vmax = blockMax(vmax, sh);
        if (tid256() != 0) return;
        const float maxVel = (float)vmax;
        const float C = maxVel * dt / dx;
        const int ns = int(fminf(ceilf(fabsf(C)), (float)maxsub));
        sc->maxvel = maxVel;
        sc->nsub = ns;
        sc->dtsub = ns > 0 ? dt / ns : 0.0f;
        if (cap > 0 && ns > cap) sc->sub_overflow = 1;
    }
};
hipError_t launch_maxvel(EnvState& e, float dt) {
    const ZGrid zg = zgrid(e);
    hipError_t r = klaunch<KMaxvel<ZB_MAXV>>(e, zg.grid, B3, 0, geom(e), zg.zc, e.f[SKT_VELX], e.f[SKT_VELY],
                                             e.f[SKT_VELZ], e.cfg.maxabs_mode, e.partials);
    if (r != hipSuccess) return r;
    // a recorded ensemble step runs adv_cap substep passes: flag a member
    // that needs more
    const int cap = e.rec ? e.rec->adv_cap : 0;
    return klaunch<KMaxvelFinalize>(e, dim3(1), B3, 0, e.partials, zg.nblocks, e.sc, dt, e.dx, e.cfg.max_substeps,
                                    cap);
}

// One directional PPM pass of advectPPMX/Y/Z (Game/kernelSky.cu:368-574) fused
// with the unit-density pass and divideValues (Game/kernelMaths.cu:38-49):
// out = (q - c dF(q)) / (1 - c dF(1)) on sky cells, c = dt/dx. The second
// face velocity reads through the advected field's bounds (:415,486,559).
// q[0..4] / o[0..4]: the field and the unit field at offsets -2..2 along the
// pass direction; vc, vm: the face velocities at +1/2 and -1/2.
// The upwind side is chosen by selecting the stencil operands of a single
// ppmFlux call (the same function of the same values as the reference's two
// calls), so a wave whose face velocities differ in sign evaluates one flux,
// not both.
__device__ __forceinline__ float ppmFaceR(const float* q, float vc, float dt, float dx) {
    const bool up = vc >= 0.0f;
    return ppmFlux(vc, up ? q[1] : q[4], up ? q[2] : q[3], up ? q[3] : q[2], dt, dx);
}
__device__ __forceinline__ float ppmFaceL(const float* q, float vm, float dt, float dx) {
    const bool dn = vm < 0.0f;
    return ppmFlux(vm, dn ? q[3] : q[0], dn ? q[2] : q[1], dn ? q[1] : q[2], dt, dx);
}
// The cell update from its two face fluxes of the field (Fr, Fl).
__device__ __forceinline__ float ppmFinish(const float* q, const float* o, float vc, float vm, float Fr, float Fl,
                                          float dt, float dx) {
    float Gr = vc, Gl = vm;
    // A unit stencil (1,1,1) reconstructs the face value 1 exactly, so away
    // from boundaries the unit field's flux is the face velocity itself.
    if (!(o[0] == 1.0f && o[1] == 1.0f && o[3] == 1.0f && o[4] == 1.0f)) {
        Gr = ppmFaceR(o, vc, dt, dx);
        Gl = ppmFaceL(o, vm, dt, dx);
    }
    const float c = dt / dx;
    const float num = q[2] - c * (Fr - Fl);
    const float den = 1.0f - c * (Gr - Gl);
    return num / den;
}
__device__ __forceinline__ float ppmCell(const float* q, const float* o, float vc, float vm, float dt, float dx) {
    return ppmFinish(q, o, vc, vm, ppmFaceR(q, vc, dt, dx), ppmFaceL(q, vm, dt, dx), dt, dx);
}

// The pass's time step: the given one or, in a recorded ensemble step
// (pb.dsc set), the member's advection substep length read on the device
// (half of it for the x and y passes); false for a pass past the member's
// substep count, which the block skips.
__device__ __forceinline__ bool ppmPassDt(const PpmBatch& pb, float& dt) {
    if (!pb.dsc) return true;
    if (pb.sub >= pb.dsc->nsub) return false;
    dt = pb.half ? pb.dsc->dtsub * 0.5f : pb.dsc->dtsub;
    return true;
}

// Field of this block in a batched pass (the grid repeats the z-march grid
// once per field).
#define PPM_FIELD_SETUP                                                   \
    const unsigned lin_ = xcdRemap(vb.id, vb.n);                          \
    const PpmField& fd_ = pb.f[lin_ / nbf];                               \
    ZM_SETUP_B(lin_ % nbf);                                               \
    (void)bid;                                                            \
    const float* __restrict__ vel = fd_.vel;                              \
    const float* __restrict__ prof = fd_.prof;                            \
    const Bounds b = fd_.b;                                               \
    /* value that stands at a ground position: the field's own ground */  \
    /* value or 0 (read only under a NEUMANN ground condition)        */  \
    const bool gNeu_ = b.ground == BC_NEUMANN && fd_.gsrc != nullptr;     \
    auto graw = [&](long gi) { return gNeu_ ? ldS<MIX>(fd_.gsrc, gi, fd_.sg) : 0.0f; }; \
    auto in = [&](long i) { return ldS<MIX>(fd_.in, i, fd_.sin); }

// x pass as a march along the rows: each wave owns one row (y, z) of a
// plane at a time and walks its 64-cell segments from x = 0 upward. The flux
// through a cell's left face is the right-face flux of the cell to its left
// (same face velocity, same three stencil values, bit for bit): lane l takes
// lane l - 1's, lane 0 takes lane 63's from the segment before. Only x = 0
// and a cell whose left neighbour is ground (its left face velocity takes the
// ground condition) evaluate it directly, so a row of X cells evaluates about
// X field fluxes instead of 2 X. Cells with y > ghm2 (largest ground height
// over the 5x5 columns around) see no ground in the stencil. A block is four
// rows (one per wave) over a chunk of zc planes of one field; the loads of up
// to XSEG_PPM segments are issued before any is used.
constexpr int XSEG_PPM = 2;
template <bool MIX>
struct KPpmXr {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        float dt = 0.0f;
        return !ppmPassDt(nthArg<3>(a...), dt);  // pb
    }
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int zc, unsigned nbf, const PpmBatch& pb,
                                               float dtArg, float dx, const float* __restrict__ cust) {
        float dt = dtArg;
        if (!ppmPassDt(pb, dt)) return;  // the whole block
        const unsigned lin = xcdRemap(vb.id, vb.n);
        const PpmField& fd = pb.f[lin / nbf];
        const unsigned rem = lin % nbf;
        const int gy = (G.Y + 3) >> 2;
        const int y = (int)(rem % (unsigned)gy) * 4 + (int)threadIdx.y;
        const int z0 = (int)(rem / (unsigned)gy) * zc, z1 = min(G.Z, z0 + zc);
        if (y >= G.Y || z0 >= z1) return;  // whole waves (one row per wave)
        auto in = [&](long i) { return ldS<MIX>(fd.in, i, fd.sin); };
        const float* __restrict__ vel = fd.vel;
        const float* __restrict__ prof = fd.prof;
        const Bounds b = fd.b;
        // value that stands at a ground position: the field's own ground value
        // or 0 (read only under a NEUMANN ground condition)
        const bool gNeu = b.ground == BC_NEUMANN && fd.gsrc != nullptr;
        const int X = G.X, tx = (int)threadIdx.x;
        const long XY = (long)X * G.Y;
        const unsigned rowy = X * y;
        const float pv = prof ? prof[y] : 0.0f, cv = cust[y];
        const int nseg = (X + 63) >> 6;
        for (int z = z0; z < z1; z++) {
            const long pl = XY * z + rowy;
            const float* vl = vel + XY * z + rowy;
            const int* gz = G.gh + X * z;
            const int* gmz = G.ghm2 + X * z;
            float carry = 0.0f;  // right-face flux of the last cell of the segment before
            for (int k0 = 0; k0 < nseg; k0 += XSEG_PPM) {
                float Q[XSEG_PPM][5], VC[XSEG_PPM], VM[XSEG_PPM];
                int GM[XSEG_PPM], GC[XSEG_PPM];
    #pragma unroll
                for (int j = 0; j < XSEG_PPM; j++) {
                    // lanes past the row (and segments past the last) read its last cell
                    const int xr = min((k0 + j) * 64 + tx, X - 1);
    #pragma unroll
                    for (int s = 0; s < 5; s++) Q[j][s] = in(pl + clampi(xr + s - 2, 0, X - 1));
                    VC[j] = vl[xr];
                    VM[j] = vl[max(xr - 1, 0)];
                    GM[j] = gmz[xr];
                    GC[j] = gz[xr];
                }
    #pragma unroll
                for (int j = 0; j < XSEG_PPM; j++) {
                    if (k0 + j >= nseg) break;  // uniform
                    const int x = (k0 + j) * 64 + tx;
                    const bool sky = x < X && y > GC[j];
                    float q[5], o[5], vm = 0.0f, Fr = 0.0f;
                    bool leftSky = false;
                    if (sky) {
                        if (y > GM[j]) {
    #pragma unroll
                            for (int s = 0; s < 5; s++) {
                                const bool outp = (x + s - 2 < 0) || (x + s - 2 > X - 1);
                                q[s] = outp ? bcv(b.sides, Q[j][s], pv) : Q[j][s];
                                o[s] = outp ? (b.sides == BC_DIRICHLET ? 0.0f : 1.0f) : 1.0f;
                            }
                            vm = x == 0 ? bcv(b.sides, VC[j], cv) : VM[j];
                            leftSky = x > 0;
                        } else {
                            int gh[5];
    #pragma unroll
                            for (int s = 0; s < 5; s++) gh[s] = gz[clampi(x + s - 2, 0, X - 1)];
    #pragma unroll
                            for (int s = 0; s < 5; s++) {
                                const bool outp = (x + s - 2 < 0) || (x + s - 2 > X - 1);
                                const bool gnd = y <= gh[s];
                                const float raw =
                                    gnd ? (gNeu ? ldS<MIX>(fd.gsrc, XY * z + rowy + clampi(x + s - 2, 0, X - 1), fd.sg) : 0.0f) : Q[j][s];
                                q[s] = nbv(raw, outp, gnd, pv, b);
                                o[s] = unitv(outp, gnd, b);
                            }
                            vm = x == 0 ? bcv(b.sides, VC[j], cv) : (y <= gh[1] ? bcv(b.ground, VC[j], cv) : VM[j]);
                            leftSky = x > 0 && y > gh[1];
                        }
                        Fr = ppmFaceR(q, VC[j], dt, dx);
                    }
                    // every lane of the wave takes part in both exchanges
                    const float up = __shfl_up(Fr, 1);
                    const float last = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, Fr), 63));
                    if (sky) {
                        const float Fl = leftSky ? (tx > 0 ? up : carry) : ppmFaceL(q, vm, dt, dx);
                        stS<MIX>(fd.out, XY * z + rowy + x, ppmFinish(q, o, VC[j], vm, Fr, Fl, dt, dx), fd.sout);
                    }
                    carry = last;
                }
            }
        }
    }
};

// y pass as a march up the columns: each thread owns the column (x, z) of
// a chunk of yc levels; the five-point stencil is a register window that
// gains one level per step (one load of the field and one of the face
// velocity per cell, issued a group of four levels ahead). The flux through
// a cell's lower face is the upper-face flux of the cell below, computed one
// step earlier (same face velocity, same three stencil values), so one field
// flux is evaluated per cell; the chunk's first cell and a cell just above
// the ground (whose lower face velocity takes the ground condition) evaluate
// it directly. Blocks are 64 (x) x 4 (z) columns of one chunk of one field.
constexpr int YC_PPM = 16;
constexpr int YG_PPM = 4;
template <bool MIX>
struct KPpmYm {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        float dt = 0.0f;
        return !ppmPassDt(nthArg<4>(a...), dt);  // pb
    }
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int yc, unsigned nyc, unsigned nbf,
                                               const PpmBatch& pb, float dtArg, float dx) {
        float dt = dtArg;
        if (!ppmPassDt(pb, dt)) return;  // the whole block
        const unsigned lin = vb.id;
        const PpmField& fd = pb.f[lin / nbf];
        const unsigned rem = lin % nbf;
        const int chunk = (int)(rem % nyc);
        const unsigned tile = rem / nyc;
        const int gx = (G.X + 63) >> 6;
        const int x = (int)(tile % (unsigned)gx) * 64 + threadIdx.x;
        const int z = (int)(tile / (unsigned)gx) * 4 + threadIdx.y;
        if (x >= G.X || z >= G.Z) return;
        auto in = [&](long i) { return ldS<MIX>(fd.in, i, fd.sin); };
        const float* __restrict__ vel = fd.vel;
        const float* __restrict__ prof = fd.prof;
        const Bounds b = fd.b;
        const bool gNeu = b.ground == BC_NEUMANN && fd.gsrc != nullptr;
        const int X = G.X, Y = G.Y;
        const long col = (long)X * Y * z + x;  // cell (x, 0, z)
        const int gh = G.gh[x + (long)X * z];
        const int y0 = chunk * yc, y1 = min(Y, y0 + yc);
        // field and unit values at level py as a stencil sees them, from the
        // field's value at the clamped level
        auto pos = [&](int py, float raw, float& qv, float& ov) {
            const int by = clampi(py, 0, Y - 1);
            const bool outp = (py < 0) || (py > Y - 1);
            const bool gnd = by <= gh;
            if (gnd) raw = gNeu ? ldS<MIX>(fd.gsrc, col + (long)X * by, fd.sg) : 0.0f;
            qv = nbv(raw, outp, gnd, prof ? prof[by] : 0.0f, b);
            ov = unitv(outp, gnd, b);
        };
        float q[5], o[5];
    #pragma unroll
        for (int s = 0; s < 4; s++) {
            const int py = y0 - 2 + s;
            pos(py, in(col + (long)X * clampi(py, 0, Y - 1)), q[s], o[s]);
        }
        // raw field values at levels g + 2 .. g + 5 and face velocities at g .. g + 3
        auto fetch = [&](int g, float* Q, float* V) {
    #pragma unroll
            for (int j = 0; j < YG_PPM; j++) {
                Q[j] = in(col + (long)X * min(g + j + 2, Y - 1));
                V[j] = vel[col + (long)X * min(g + j, Y - 1)];
            }
        };
        float vPrev = vel[col + (long)X * max(y0 - 1, 0)];
        float fPrev = 0.0f;
        auto group = [&](int g, const float* Q, const float* V) {
    #pragma unroll
            for (int j = 0; j < YG_PPM; j++) {
                const int y = g + j;
                if (y >= y1) break;
                pos(y + 2, Q[j], q[4], o[4]);
                const float vc = V[j];
                if (y > gh) {
                    // y - 1 >= 0 on every sky cell
                    const bool below = y - 1 <= gh;
                    const float vm = below ? bcv(b.ground, vc, 0.0f) : vPrev;
                    const float Fr = ppmFaceR(q, vc, dt, dx);
                    const float Fl = (y > y0 && !below) ? fPrev : ppmFaceL(q, vm, dt, dx);
                    fPrev = Fr;
                    stS<MIX>(fd.out, col + (long)X * y, ppmFinish(q, o, vc, vm, Fr, Fl, dt, dx), fd.sout);
                }
                vPrev = vc;
    #pragma unroll
                for (int s = 0; s < 4; s++) {
                    q[s] = q[s + 1];
                    o[s] = o[s + 1];
                }
            }
        };
        float Qa[YG_PPM], Va[YG_PPM], Qb[YG_PPM], Vb[YG_PPM];
        fetch(y0, Qa, Va);
        for (int g = y0; g < y1; g += 2 * YG_PPM) {
            if (g + YG_PPM < y1) fetch(g + YG_PPM, Qb, Vb);
            group(g, Qa, Va);
            if (g + YG_PPM >= y1) break;
            if (g + 2 * YG_PPM < y1) fetch(g + 2 * YG_PPM, Qa, Va);
            group(g + YG_PPM, Qb, Vb);
        }
    }
};

constexpr int ZB_PPM = 4;  // planes of the z pass whose loads go together
// z pass: the stencil runs along this thread's march; a window of planes
// zg-2 .. zg+ZB+1 (values and ground heights) stays in registers, so the
// boundary rule needs no extra loads.
template <int ZB, bool MIX>
struct KPpmZ {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        float dt = 0.0f;
        return !ppmPassDt(nthArg<3>(a...), dt);  // pb
    }
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int zc, unsigned nbf, const PpmBatch& pb,
                                               float dtArg, float dx, const float* __restrict__ cust) {
        float dt = dtArg;
        if (!ppmPassDt(pb, dt)) return;  // the whole block
        PPM_FIELD_SETUP;
        if (!inb || z0 >= z1) return;
        const int X = G.X, Z = G.Z;
        const long XY = (long)X * G.Y;
        const unsigned rc = x + X * y;
        const float pv = prof ? prof[y] : 0.0f, cv = cust[y];
        float qw[ZB + 4], vw[ZB + 1];
        int gw[ZB + 4];
    #pragma unroll
        for (int s = 0; s < 4; s++) {
            const int bz = clampi(z0 - 2 + s, 0, Z - 1);
            qw[s] = in(XY * bz + rc);
            gw[s] = (G.gh + X * bz)[(unsigned)x];
        }
        vw[0] = (vel + XY * max(z0 - 1, 0))[rc];
        // window slot m holds plane zg - 2 + m; velocity slot m plane zg - 1 + m
        auto load = [&](int z, float& q, int& g, float& v) {
            const int bz = min(z + 2, Z - 1);
            q = in(XY * bz + rc);
            g = (G.gh + X * bz)[(unsigned)x];
            v = (vel + XY * z)[rc];
        };
        // The flux through the back face of plane z is the front-face flux of
        // plane z - 1 when this thread computed that plane (same face velocity,
        // same stencil values); else it is evaluated here.
        float fPrev = 0.0f;
        auto plane = [&](int z, const float* qs, const int* gs, float vc, float vmr) {
            if (y <= gs[2]) return;
            float q[5], o[5];
    #pragma unroll
            for (int s = 0; s < 5; s++) {
                const int pz = z + s - 2;
                const bool outp = (pz < 0) || (pz > Z - 1);
                const bool gnd = y <= gs[s];
                q[s] = nbv(gnd ? graw(XY * clampi(pz, 0, Z - 1) + rc) : qs[s], outp, gnd, pv, b);
                o[s] = unitv(outp, gnd, b);
            }
            const float vm = z == 0 ? bcv(b.sides, vc, cv) : (y <= gs[1] ? bcv(b.ground, vc, cv) : vmr);
            const float Fr = ppmFaceR(q, vc, dt, dx);
            const float Fl = (z > z0 && y > gs[1]) ? fPrev : ppmFaceL(q, vm, dt, dx);
            fPrev = Fr;
            stS<MIX>(fd_.out, XY * z + rc, ppmFinish(q, o, vc, vm, Fr, Fl, dt, dx), fd_.sout);
        };
        int zg = z0;
        for (; zg + ZB <= z1; zg += ZB) {
    #pragma unroll
            for (int j = 0; j < ZB; j++) load(zg + j, qw[4 + j], gw[4 + j], vw[1 + j]);
    #pragma unroll
            for (int j = 0; j < ZB; j++) plane(zg + j, qw + j, gw + j, vw[1 + j], vw[j]);
    #pragma unroll
            for (int s = 0; s < 4; s++) {
                qw[s] = qw[s + ZB];
                gw[s] = gw[s + ZB];
            }
            vw[0] = vw[ZB];
        }
        for (; zg < z1; zg++) {
            load(zg, qw[4], gw[4], vw[1]);
            plane(zg, qw, gw, vw[1], vw[0]);
    #pragma unroll
            for (int s = 0; s < 4; s++) {
                qw[s] = qw[s + 1];
                gw[s] = gw[s + 1];
            }
            vw[0] = vw[1];
        }
    }
};

hipError_t launch_ppm_batch(EnvState& e, const PpmBatch& pb, int dir, float dt) {
    const ZGrid zg = e.rec ? zgridChunks(e, ensembleChunks(e)) : zgrid(e);
    const dim3 grid(zg.nblocks * pb.n);
    const bool m16 = e.mix && has16(pb);
    if (dir == 0) {
        // rows of four per block, chunks of zc planes: about 2048 blocks per field
        const long gy = (e.Y + 3) / 4;
        const long nzc = e.rec ? ensembleChunks(e) : std::max(1L, std::min((long)e.Z, 2048 / gy));
        const int zc = (int)((e.Z + nzc - 1) / nzc);
        const unsigned nbf = (unsigned)(gy * ((e.Z + zc - 1) / zc));
        return m16 ? klaunch<KPpmXr<true>>(e, dim3(nbf * pb.n), B3, 0, geom(e), zc, nbf, pb, dt, e.dx,
                                             e.f[SKT_DEFAULT_VELX])
                     : klaunch<KPpmXr<false>>(e, dim3(nbf * pb.n), B3, 0, geom(e), zc, nbf, pb, dt, e.dx,
                                              e.f[SKT_DEFAULT_VELX]);
    }
    if (dir == 1) {
        const unsigned nyc = (unsigned)((e.Y + YC_PPM - 1) / YC_PPM);
        const unsigned nbf = (unsigned)((e.X + 63) / 64) * (unsigned)((e.Z + 3) / 4) * nyc;
        return m16 ? klaunch<KPpmYm<true>>(e, dim3(nbf * pb.n), B3, 0, geom(e), YC_PPM, nyc, nbf, pb, dt, e.dx)
                     : klaunch<KPpmYm<false>>(e, dim3(nbf * pb.n), B3, 0, geom(e), YC_PPM, nyc, nbf, pb, dt, e.dx);
    }
    return m16 ? klaunch<KPpmZ<ZB_PPM, true>>(e, grid, B3, 0, geom(e), zg.zc, zg.nblocks, pb, dt, e.dx,
                                                e.f[SKT_DEFAULT_VELZ])
                 : klaunch<KPpmZ<ZB_PPM, false>>(e, grid, B3, 0, geom(e), zg.zc, zg.nblocks, pb, dt, e.dx,
                                                 e.f[SKT_DEFAULT_VELZ]);
}

// ---------------------------------------------------- pressure projection ---
// char conversion of a lower-triangle coefficient (Game/kernelSky.cu:762-764, quirk A1)
__device__ __forceinline__ float trunc8(float c) { return fminf(fmaxf(truncf(c), -128.0f), 127.0f); }

// Column marches (preconditioner, divergence, projection): chunks of YC_DIV
// levels, blocks of 64 (x) x 4 (z) columns of one chunk.
constexpr int YC_DIV = 8;
static unsigned divBlocks(const EnvState& e) {
    return (unsigned)((e.X + 63) / 64) * (unsigned)((e.Z + 3) / 4) * (unsigned)((e.Y + YC_DIV - 1) / YC_DIV);
}

// initAMatrix + initPrecon, Game/kernelSky.cu:1320-1430, density bounds
// {N,D,D}. The operator is not stored: the SpMV rebuilds its coefficients
// from 1/density (irho, 4 B per cell instead of 16 for a float4 operator).
// The divergence pass below writes irho and the Jacobi preconditioner
// 1/diagonal: it forms the same face sums for its harmonic densities.
struct KSetInt {
    static __device__ __forceinline__ void run(VB vb, int* a, int v) {
        if (threadIdx.x == 0) *a = v;
    }
};
// Scalar steps of the solve, run by every block of the consuming kernel
// (fixed-order reductions, identical in every block); block 0 stores the
// result. Steps follow calculatePressureProject (Game/environment.cu:930-985):
//   PRO_START: initial stop-test value; done if it is 0 (:936); tolerance
//   PRO_A0:    sigma0 = z.r (:943)
//   PRO_A:     stop test of the previous iteration (:967); then sigma1 = z.r,
//              beta = sigma1 / sigma0, sigma0 = sigma1 (:975-984)
//   PRO_B:     sigma1 (+= in reference mode: the value is not cleared before
//              the s.As product from the second iteration on), alpha
//   PRO_FINAL: stop test after the last iteration
enum { PRO_START = 0, PRO_A0, PRO_A, PRO_B, PRO_FINAL };
// A lane's share of the previous kernel's partials: elements t, t + 256, ...
// below 2048 (later ones are read in the prologue itself). Loading them
// before the kernel issues its first data loads lets both be in flight while
// the block waits; the prologue then reduces them.
struct PcgPart {
    double m[8], s[8];
};
template <int KIND>
__device__ __forceinline__ void pcgPartLoad(const double* __restrict__ pm, const double* __restrict__ ps, unsigned nm,
                                            unsigned ns, PcgPart& P) {
    const unsigned t = tid256();
    constexpr bool wantM = KIND == PRO_START || KIND == PRO_A || KIND == PRO_FINAL;
    constexpr bool wantS = KIND == PRO_A0 || KIND == PRO_A || KIND == PRO_B;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const unsigned j = t + 256u * k;
        P.m[k] = wantM && j < nm ? pm[j] : 0.0;
        P.s[k] = wantS && j < ns ? ps[j] : 0.0;
    }
}
// Under FP64 refinement (ro set; pass = the solve's index in the projection)
// PRO_START also applies the projection's shared stop value ro->target =
// tolOut * max|b|: pass 0 sets it, a later pass is done at once when its
// residual already meets it, and every pass stops at the larger of its own
// relative tolerance and the target.
template <int KIND>
__device__ void pcgPrologue(const double* __restrict__ pm, const double* __restrict__ ps, unsigned nm, unsigned ns,
                            const PcgPart& P, const PcgState* __restrict__ sin, PcgState* __restrict__ sout, int ref,
                            float tol, int it, PcgState& S, double* sh, unsigned tile, RefineStats* ro = nullptr,
                            int pass = 0, float tolOut = 0.0f) {
    const int t = tid256();
    if (t == 0) {
        if (KIND == PRO_START) {
            S = PcgState{};
        } else {
            S = *sin;
        }
    }
    __syncthreads();
    if (!S.done) {
        double vmax = 0.0, vsum = 0.0;
        if (KIND == PRO_START || KIND == PRO_A || KIND == PRO_FINAL) {
#pragma unroll
            for (int k = 0; k < 8; k++) vmax = fmax(vmax, P.m[k]);
            for (unsigned j = t + 2048; j < nm; j += 256) vmax = fmax(vmax, pm[j]);
            vmax = blockMax(vmax, sh);
        }
        if (KIND == PRO_A0 || KIND == PRO_A || KIND == PRO_B) {
#pragma unroll
            for (int k = 0; k < 8; k++) vsum += P.s[k];
            for (unsigned j = t + 2048; j < ns; j += 256) vsum += ps[j];
            vsum = blockSum(vsum, sh);
        }
        if (t == 0) {
            switch (KIND) {
                case PRO_START: {
                    const float m0 = (float)vmax;
                    S.maxr0 = m0;
                    S.maxr = m0;
                    S.iters = 0;
                    S.done = (m0 == 0.0f) ? 1 : 0;
                    S.tol = ref ? tol : tol * m0;
                    if (ro) {
                        if (pass == 0) {
                            const float tg = tolOut * m0;
                            S.tol = fmaxf(S.tol, tg);
                            if (tile == 0) ro->target = tg;
                        } else {
                            const float tg = ro->target;
                            if (m0 <= tg) S.done = 1;
                            S.tol = fmaxf(S.tol, tg);
                        }
                    }
                } break;
                case PRO_A0:
                    S.sigma0 = (double)(float)vsum;
                    S.sigma1 = 0.0;
                    break;
                case PRO_A:
                case PRO_FINAL: {
                    const float mr = (float)vmax;
                    S.maxr = mr;
                    S.iters = it;
                    if (mr <= S.tol) {
                        S.done = 1;
                    } else if (KIND == PRO_A) {
                        const float s1 = (float)vsum;
                        S.sigma1 = (double)s1;
                        S.beta = s1 / (float)S.sigma0;
                        S.sigma0 = (double)s1;
                    }
                } break;
                case PRO_B: {
                    const float s1 = ref ? (float)((double)(float)S.sigma1 + vsum) : (float)vsum;
                    S.sigma1 = (double)s1;
                    S.alpha = (float)S.sigma0 / s1;
                } break;
                default: break;
            }
        }
    }
    if (t == 0 && tile == 0) *sout = S;
    __syncthreads();
}

// calculateDivergenceGPU, Game/kernelSky.cu:780-860 (density {N,N,D}, velocity
// {C,C,D}) with the per-block maxima of the initial stop-test value
// (Game/environment.cu:933-936). p' is cleared. The pass also sets up the
// pressure solve (initAMatrix + initPrecon, Game/kernelSky.cu:1320-1430,
// density bounds {N,D,D}): the divergence's harmonic face densities 2 / s and
// the operator's face coefficients -max(1, 1.225 / (2 / s)) share the sum
// s = 1/rho + 1/rho_nb with the same neighbour rules (beyond the domain the
// cell's own density, at ground a zero face; the top face takes the sides
// condition in both), so it writes irho (1/density on sky cells, 0 in
// ground) and the Jacobi preconditioner 1/diagonal, and flags faceExact.
// A march up the columns: each thread walks a chunk of YC_DIV levels of its
// (x, z) column. The ground heights of the column and of its four neighbour
// columns are read once; the density, its reciprocal and v of the level
// below come from the step before, and the level above's density and
// reciprocal become the next step's own. Each neighbour value follows extv /
// nbOwn: a position beyond the domain clamps onto this sky cell and takes
// the sides condition, a ground position the ground condition (the level
// above a sky cell is sky); the arithmetic per cell is the reference's, with
// the correctly rounded recipRN / twoOverRN for 1/density and the harmonic
// face densities. The loads of a group of DIV_G levels are issued together
// (at clamped in-range positions, so that they are unconditional) before the
// group is computed. Blocks are 64 (x) x 4 (z) columns of one chunk, each
// writing one partial maximum.
constexpr int DIV_G = 4;
struct KDivergence {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int yc, unsigned nyc, float* __restrict__ r, float* __restrict__ pp,
                         const float* __restrict__ rho, const float* __restrict__ rold, const float* __restrict__ u,
                         const float* __restrict__ v, const float* __restrict__ w, const float* __restrict__ defU,
                         const float* __restrict__ defW, float dt, int absmode, double* part, float* __restrict__ irho,
                         float* __restrict__ pre, int* __restrict__ faceExact) {
        __shared__ double sh[4];
        const unsigned lin = vb.id;
        const int chunk = (int)(lin % nyc);
        const unsigned tile = lin / nyc;
        const int gx = (G.X + 63) >> 6;
        const int x = (int)(tile % (unsigned)gx) * 64 + (int)threadIdx.x;
        const int z = (int)(tile / (unsigned)gx) * 4 + (int)threadIdx.y;
        double m = 0.0;
        const int y0 = chunk * yc, y1 = min(G.Y, y0 + yc);
        if (x < G.X && z < G.Z && y0 < y1) {
            const int X = G.X, Y = G.Y, Z = G.Z;
            const long XY = (long)X * Y;
            const long col = XY * z + x;  // cell (x, 0, z)
            const long gc = x + (long)X * z;
            const int gh = G.gh[gc];
            const bool inL = x > 0, inR = x < X - 1, inB = z > 0, inF = z < Z - 1;
            const int ghL = inL ? G.gh[gc - 1] : 0, ghR = inR ? G.gh[gc + 1] : 0;
            const int ghB = inB ? G.gh[gc - X] : 0, ghF = inF ? G.gh[gc + X] : 0;
            const long oR = inR ? 1 : 0, oL = inL ? -1 : 0, oF = inF ? XY : 0, oB = inB ? -XY : 0;
            const Bounds bd{BC_NEUMANN, BC_NEUMANN, BC_DIRICHLET}, bv{BC_CUSTOM, BC_CUSTOM, BC_DIRICHLET};
            const bool ab = absmode || z == 0;
            float rhoB = 0.0f, vB = 0.0f;  // density and v of the level below
            if (y0 > 0) {
                rhoB = rho[col + (long)X * (y0 - 1)];
                vB = v[col + (long)X * (y0 - 1)];
            }
            float cur = rho[col + (long)X * y0];
            float icur = recipRN(cur), icPrev = recipRN(rhoB);
            bool bad = false;
            for (int yb = y0; yb < y1; yb += DIV_G) {
                float qU[DIV_G], qR[DIV_G], qL[DIV_G], qF[DIV_G], qB[DIV_G];
                float qv[DIV_G], qu[DIV_G], quL[DIV_G], qw[DIV_G], qwB[DIV_G], qo[DIV_G];
#pragma unroll
                for (int k = 0; k < DIV_G; k++) {
                    const int y = min(yb + k, y1 - 1);
                    const long i = col + (long)X * y;
                    qU[k] = rho[y + 1 < Y ? i + X : i];
                    qR[k] = rho[i + oR];
                    qL[k] = rho[i + oL];
                    qF[k] = rho[i + oF];
                    qB[k] = rho[i + oB];
                    qv[k] = v[i];
                    qu[k] = u[i];
                    quL[k] = u[i + oL];
                    qw[k] = w[i];
                    qwB[k] = w[i + oB];
                    qo[k] = rold[i];
                }
#pragma unroll
                for (int k = 0; k < DIV_G; k++) {
                    const int y = yb + k;
                    if (y >= y1) break;  // uniform: the chunk is the block's
                    const long i = col + (long)X * y;
                    pp[i] = 0.0f;
                    const float vc = qv[k];
                    const bool topS = y + 1 > Y - 1;
                    const float rU = topS ? 0.0f : qU[k];
                    // one range test for the cell: its own 1/density and every
                    // sky neighbour density it takes a reciprocal of; the cell
                    // is then computed with the fast forms throughout
                    const bool skR = inR && y > ghR, skL = inL && y > ghL, skF = inF && y > ghF, skB = inB && y > ghB;
                    const bool skD = y - 1 >= 0 && y - 1 > gh;
                    const bool fast = densOk(icur) && (topS || densOk(rU)) && (!skD || densOk(icPrev)) &&
                                      (!skR || densOk(qR[k])) && (!skL || densOk(qL[k])) && (!skF || densOk(qF[k])) &&
                                      (!skB || densOk(qB[k]));
                    float dv = 0.0f;
                    auto cell = [&](auto fastTag) __attribute__((always_inline)) {
                        constexpr bool F = decltype(fastTag)::value;
                        auto rcp = [&](float d) { return F ? recipRNfast(d) : recipRN(d); };
                        const float icU = topS ? 0.0f : rcp(rU);
                        if (y > gh) {
                            const float ur = qu[k];
                            const float ul = !inL ? bcv(bv.sides, ur, defU[y]) : (y <= ghL ? bcv(bv.ground, ur, defU[y]) : quL[k]);
                            const float uu = vc;
                            const float ud = y - 1 < 0 ? bcv(bv.sides, vc, 0.0f) : (y - 1 <= gh ? bcv(bv.ground, vc, 0.0f) : vB);
                            const float uf = qw[k];
                            const float ub = !inB ? bcv(bv.sides, uf, defW[y]) : (y <= ghB ? bcv(bv.ground, uf, defW[y]) : qwB[k]);
                            float dr = !inR ? bcv(bd.sides, cur, 1.0f) : (y <= ghR ? bcv(bd.ground, qR[k], 1.0f) : qR[k]);
                            float dl = !inL ? bcv(bd.sides, cur, 1.0f) : (y <= ghL ? bcv(bd.ground, qL[k], 1.0f) : qL[k]);
                            float du = topS ? bcv(bd.sides, cur, 1.0f) : rU;
                            float dd = y - 1 < 0 ? bcv(bd.sides, cur, 1.0f) : (y - 1 <= gh ? bcv(bd.ground, rhoB, 1.0f) : rhoB);
                            float df = !inF ? bcv(bd.sides, cur, 1.0f) : (y <= ghF ? bcv(bd.ground, qF[k], 1.0f) : qF[k]);
                            float db = !inB ? bcv(bd.sides, cur, 1.0f) : (y <= ghB ? bcv(bd.ground, qB[k], 1.0f) : qB[k]);
                            // the level below's 1/density is that cell's carried own
                            // one when it is sky (at y = 0 the cell's own; at ground
                            // dd is 0 and takes no reciprocal)
                            const float idd = skD ? icPrev : icur;
                            // each face's harmonic density 2 / s, s = 1/rho + 1/rho_nb, and
                            // the pressure operator's coefficient of the same face,
                            // -max(1, 1.225 / (2 / s)) (initAMatrix's faceCoef; a zero
                            // density, ground, gives 0 to both)
                            float cR, cL, cU, cD, cF, cB;
                            auto hf = [&](float& d, float in, float& c) {
                                if (d == 0.0f) {
                                    c = 0.0f;
                                    return;
                                }
                                const float sum = icur + in;
                                if constexpr (F) {
                                    const float t = recipRNfast(sum);
                                    d = 2.0f * t;
                                    c = faceOfRecip(t, sum);
                                } else {
                                    d = twoOverRN(sum);
                                    c = -fmaxf(1.0f, 1.225f / d);
                                }
                            };
                            hf(dr, inR ? rcp(dr) : icur, cR);
                            hf(dl, inL ? rcp(dl) : icur, cL);
                            hf(du, topS ? icur : icU, cU);
                            hf(dd, idd, cD);
                            hf(df, inF ? rcp(df) : icur, cF);
                            hf(db, inB ? rcp(db) : icur, cB);
                            // the Jacobi preconditioner: 1 / diagonal, the diagonal
                            // minus the six coefficients in initAMatrix's order
                            float wd = 0.0f;
                            wd -= cR;
                            wd -= cL;
                            wd -= cU;
                            wd -= cD;
                            wd -= cF;
                            wd -= cB;
                            irho[i] = icur;
                            pre[i] = wd == 0.0f ? 1.0f : rcp(wd);
                            bad |= !faceValueOk(icur);
                            const float mfd = ((dr * ur - dl * ul) + (du * uu - dd * ud) + (df * uf - db * ub));
                            const float dc = (cur - qo[k]) / dt;
                            dv = (mfd - dc);
                            m = fmax(m, (double)(ab ? fabsf(dv) : dv));
                        } else {
                            irho[i] = 0.0f;
                            pre[i] = 1.0f;
                        }
                        return icU;
                    };
                    const float icU = fast ? cell(std::true_type{}) : cell(std::false_type{});
                    r[i] = dv;
                    rhoB = cur;
                    vB = vc;
                    icPrev = icur;
                    cur = rU;
                    icur = icU;
                }
            }
            // the SpMV's short face form is exact when every sky 1/density is in range
            if (bad) atomicOr(faceExact, 1);
        }
        m = blockMax(m, sh);
        if (tid256() == 0) part[vb.id] = m;
    }
};
hipError_t launch_divergence(EnvState& e, float dt) {
    const int ref = e.cfg.pcg_mode == SKT_PCG_REFERENCE;
    const int absmode = ref ? e.cfg.maxabs_mode : 1;
    const unsigned nyc = (unsigned)((e.Y + YC_DIV - 1) / YC_DIV);
    hipError_t rc = klaunch<KSetInt>(e, dim3(1), dim3(64), 0, &e.sc->face_exact, 0);
    if (rc != hipSuccess) return rc;
    return klaunch<KDivergence>(e, dim3(divBlocks(e)), B3, 0, geom(e), YC_DIV, nyc, e.pcg[0], e.f[SKT_PRESSURE_PROJ],
                                e.f[SKT_DENSITY], e.f[SKT_DENSITY_OLD], e.f[SKT_VELX], e.f[SKT_VELY], e.f[SKT_VELZ],
                                e.f[SKT_DEFAULT_VELX], e.f[SKT_DEFAULT_VELZ], dt, absmode,
                                e.partials + (size_t)PART_DIV * e.maxBlocks, e.irho, e.pcg[5], &e.sc->face_exact);
}

// Prologue PRO_START, then z = M^-1 r (applyPreconditionerGPU,
// Game/kernelSky.cu:659-673) and the z.r partials (dotProductGPU, :675-717);
// under the multigrid preconditioner (jacobi 0) only the prologue.
struct KPcgStart {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int zc, const float* __restrict__ r, float* __restrict__ zz, const float* __restrict__ pre,
                        const double* __restrict__ pin, unsigned n, double* pout, const PcgState* sin, PcgState* sout,
                        int ref, float tol, RefineStats* ro, int pass, float tolOut, int jacobi) {
        ZM_SETUP_V;
        __shared__ double sh[4];
        __shared__ PcgState S;
        PcgPart P;
        // the planes are read in groups of four at positions clamped into the
        // domain (unconditional loads); the first group goes out before the
        // prologue, each later one while the group before is used
        constexpr int PG = 4;
        const int xc = min(x, G.X - 1), yc = min(y, G.Y - 1);
        float rv[PG], pv[PG];
        int gv[PG];
        auto loadGroup = [&](int zb) __attribute__((always_inline)) {
#pragma unroll
            for (int k = 0; k < PG; k++) {
                const int z = min(zb + k, z1 - 1);
                const long i = gidx(G, xc, yc, z);
                rv[k] = r[i];
                pv[k] = pre[i];
                gv[k] = G.gh[xc + (long)G.X * z];
            }
        };
        if (jacobi && z0 < z1) loadGroup(z0);
        pcgPartLoad<PRO_START>(pin, nullptr, n, n, P);
        pcgPrologue<PRO_START>(pin, nullptr, n, n, P, sin, sout, ref, tol, 0, S, sh, vb.id, ro, pass, tolOut);
        // under the multigrid preconditioner the V-cycle that follows forms z
        // and its partials: only the prologue runs
        if (S.done || !jacobi) return;
        double s = 0.0;
        for (int zb = z0; zb < z1; zb += PG) {
            float cr[PG], cp[PG];
            int cg[PG];
#pragma unroll
            for (int k = 0; k < PG; k++) {
                cr[k] = rv[k];
                cp[k] = pv[k];
                cg[k] = gv[k];
            }
            if (zb + PG < z1) loadGroup(zb + PG);
#pragma unroll
            for (int k = 0; k < PG; k++) {
                const int z = zb + k;
                if (z >= z1) break;
                if (!inb) continue;
                float zv = 0.0f;
                if (y > cg[k]) {
                    zv = cr[k] * cp[k];
                    s += (double)(zv * cr[k]);
                }
                zz[gidx(G, x, y, z)] = zv;
            }
        }
        s = blockSum(s, sh);
        if (tid256() == 0) pout[bid] = s;
    }
};
hipError_t launch_pcg_start(EnvState& e, int pass) {
    const int ref = e.cfg.pcg_mode == SKT_PCG_REFERENCE;
    const ZGrid zg = zgrid(e);
    const size_t M = e.maxBlocks;
    const bool refine = pass >= 0;
    return klaunch<KPcgStart>(e, zg.grid, B3, 0, geom(e), zg.zc, e.pcg[0], e.pcg[1], e.pcg[5], e.partials + PART_DIV * M,
                              divBlocks(e), e.partials + PART_START * M, &e.sc->st[0], &e.sc->st[1], ref,
                              refine ? e.cfg.pcg_refine_tol : e.cfg.pcg_tol, refine ? &e.sc->ref : nullptr,
                              refine ? pass : 0, e.cfg.pcg_tol, e.mg_depth > 0 ? 0 : 1);
}

// Prologue PRO_A0 / PRO_A, then the search-direction update (endIteration,
// Game/kernelSky.cu:987-1014) fused into applyAGPU (:719-778) and the s.As
// partials (:675-717): s_new = z + beta s_old (s_new = z on the first
// iteration), As = A s_new. The operator is rebuilt per cell from 1/density
// (faceI; diagonal = minus the six face coefficients, in initAMatrix's
// order; outside faces use the cell's own density, Neumann; ground faces 0).
// TRUNC selects the reference mode's char-truncated lower triangle (quirk
// A1); without it the operator is symmetric.
// s_new and 1/density of the in-plane neighbours come from the LDS tile;
// along z they stay in registers, and the back face is the front face of the
// plane before.
// The tile encodes the boundary rules, so a cell reads its four in-plane
// neighbours without tests: a ground cell holds s = 0 and 1/density = inf
// (faceI gives 0), a position just outside the domain holds s = 0 and the
// boundary cell's own 1/density (every thread reads its position clamped into
// the domain). Faces take the short division (faceFast); it is exact when
// each sky value lies in faceValueOk's range, which the divergence pass
// checks for the whole solve (faceExact set: every face takes faceI).
// s_new of the front plane, computed for this plane's front face, is the
// next plane's own s_new.
// The z loop runs four planes per trip over a ring of four plane slots (own
// data is read three planes ahead) and two halo slots, so no values move
// between registers and each load is waited for only where it is used.
// Byte offsets into the arrays count from plane zb = max(z0 - 1, 0); the
// launch keeps (chunk + 3 planes) below 4 GiB.
template <bool FIRST, bool TRUNC>
struct KPcgA {
    // a finished solve (sin: argument 12)
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        return nthArg<12>(a...)->done != 0;
    }
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int zc, const float* __restrict__ zz, const float* __restrict__ s_old,
                    float* __restrict__ s_new, const float* __restrict__ irho, float* __restrict__ As,
                    const double* __restrict__ pm, const double* __restrict__ ps, unsigned n, unsigned ns, double* pout,
                    const PcgState* sin, PcgState* sout, int it, const int* __restrict__ faceExact) {
        ZM_SETUP_V;
        __shared__ double sh[4];
        __shared__ PcgState S;
        __shared__ float tS[2][TH][TW], tI[2][TH][TW];
        const bool exact = *faceExact != 0;
        constexpr int KIND = FIRST ? PRO_A0 : PRO_A;
        PcgPart part;
        pcgPartLoad<KIND>(pm, ps, n, ns, part);
        float beta = 0.0f;
        const int tx = threadIdx.x, ty = threadIdx.y;
        const int X = G.X, Y = G.Y, Z = G.Z;
        const long XY = (long)X * Y;
        const float INF = __builtin_inff();
        auto snv = [&](float zv, float sv) { return FIRST ? zv : zv + beta * sv; };
        auto lo = [&](float c) { return TRUNC ? trunc8(c) : c; };
        // byte offsets of the lane's cell (clamped into the domain) in a plane
        // and of its column in a ground-height plane
        const int xc = min(x, X - 1), yc = min(y, Y - 1);
        const unsigned ob = 4u * (unsigned)(xc + X * yc), gb = 4u * (unsigned)xc;
        // One halo cell per lane at most: the row below the tile (wave 0), the
        // row above it (wave 1), the columns left and right of it (lanes 0-3 and
        // 4-7 of wave 2, one row each).
        const int x0 = x - tx, y0 = y - ty;
        int hx = x, hy = y0 - 1, hrow = 0, hcol = tx + 1;
        if (ty == 1) hy = y0 + 4, hrow = 5;
        if (ty == 2) hx = tx < 4 ? x0 - 1 : x0 + 64, hy = y0 + (tx & 3), hrow = 1 + (tx & 3), hcol = tx < 4 ? 0 : 65;
        const bool hasH = ty < 2 || (ty == 2 && tx < 8);
        const bool hIn = hx >= 0 && hx < X && hy >= 0 && hy < Y;
        const int hxc = min(max(hx, 0), X - 1), hyc = min(max(hy, 0), Y - 1);
        const unsigned hb = 4u * (unsigned)(hxc + X * hyc), hgb = 4u * (unsigned)hxc;
        const unsigned hbL = hasH ? hb : ob, hgbL = hasH ? hgb : gb;
        const int zb = max(z0 - 1, 0);
        const unsigned PB = 4u * (unsigned)XY, RB = 4u * (unsigned)X;
        const Rsrc rZ = arrayRsrc(zz + XY * zb), rS = arrayRsrc(s_old + XY * zb), rI = arrayRsrc(irho + XY * zb);
        const Rsrc rG = arrayRsrc(G.gh + (long)X * zb);
        const Rsrc wS = arrayRsrc(s_new + XY * zb), wA = arrayRsrc(As + XY * zb);
        auto pOff = [&](int pl) { return PB * (unsigned)(pl - zb); };
        auto gOff = [&](int pl) { return RB * (unsigned)(pl - zb); };
        struct Own {
            float z, s, i;
            int gf;  // ground height of the next plane
        };
        struct Halo {
            float z, s, i;
            int g;  // ground height of the halo cell's column
        };
        auto loadOwn = [&](int pl, Own& o) __attribute__((always_inline)) {
            const unsigned so = pOff(pl);
            o.z = bldf(rZ, ob, so);
            if (!FIRST) o.s = bldf(rS, ob, so);
            o.i = bldf(rI, ob, so);
            o.gf = bldi(rG, gb, gOff(min(pl + 1, Z - 1)));
        };
        // Every lane issues the same loads in every plane (a lane without a halo
        // cell reads its own, a prefetch past the chunk reads a clamped plane):
        // loads complete in issue order, and with no branches around them each
        // wait covers exactly the loads its values come from.
        auto loadHalo = [&](int pl, Halo& h) __attribute__((always_inline)) {
            const unsigned so = pOff(pl);
            h.z = bldf(rZ, hbL, so);
            if (!FIRST) h.s = bldf(rS, hbL, so);
            h.i = bldf(rI, hbL, so);
            h.g = bldi(rG, hgbL, gOff(pl));
        };
        // the chunk's first planes and the plane below it are read before the
        // prologue; s_new and the back face of the plane below need beta
        int ghc = bldi(rG, gb, gOff(z0)), ghb = 0;
        float bz = 0.0f, bs = 0.0f, bi = 0.0f;
        if (z0 > 0) {
            const unsigned so = pOff(z0 - 1);
            ghb = bldi(rG, gb, gOff(z0 - 1));
            bz = bldf(rZ, ob, so);
            if (!FIRST) bs = bldf(rS, ob, so);
            bi = bldf(rI, ob, so);
        }
        Own o0{0.0f, 0.0f, 0.0f, 0}, o1{0.0f, 0.0f, 0.0f, 0}, o2{0.0f, 0.0f, 0.0f, 0}, o3{0.0f, 0.0f, 0.0f, 0};
        Halo h0{0.0f, 0.0f, 0.0f, 0}, h1{0.0f, 0.0f, 0.0f, 0};
        // prefetches past the chunk read the next chunk's first planes: measured
        // faster than clamping them to the chunk's last plane
        const int zlast = Z - 1;
        loadOwn(z0, o0);
        loadOwn(min(z0 + 1, zlast), o1);
        loadOwn(min(z0 + 2, zlast), o2);
        loadHalo(z0, h0);
        pcgPrologue<KIND>(pm, ps, n, ns, part, sin, sout, 0, 0.0f, it, S, sh, vb.id);
        if (S.done) return;
        beta = S.beta;
        // plane below the chunk: s_new, and its front face (= this plane's back face)
        float snb = 0.0f, fb = 0.0f;
        bool skB = false;
        if (inb && z0 > 0 && y > ghb) {
            skB = true;
            snb = snv(bz, bs);
            if (y > ghc) fb = faceI(bi, o0.i);
        }
        bool skC = inb && y > ghc;
        float snc = skC ? snv(o0.z, o0.s) : 0.0f;  // s_new of the current plane
        double acc = 0.0;
        // plane z from oc, its front plane from of; plane z + 3 is read into o3n
        auto step = [&](int z, int buf, Own& oc, Own& of, Own& o3n, Halo& hc, Halo& hn) __attribute__((always_inline)) {
            // the halo (used next plane) goes first, so waiting for it does not
            // wait for the plane read three ahead
            loadHalo(min(z + 1, zlast), hn);
            loadOwn(min(z + 3, zlast), o3n);
            const bool inZ = z + 1 < Z;
            const bool skF = inb && inZ && y > oc.gf;
            tS[buf][ty + 1][tx + 1] = snc;
            tI[buf][ty + 1][tx + 1] = skC ? oc.i : (inb ? INF : oc.i);
            if (hasH) {
                const bool sky = hIn && hy > hc.g;
                tS[buf][hrow][hcol] = sky ? snv(hc.z, hc.s) : 0.0f;
                tI[buf][hrow][hcol] = sky ? hc.i : (hIn ? INF : hc.i);
            }
            __syncthreads();
            const float snf = skF ? snv(of.z, of.s) : 0.0f;
            float fcur = 0.0f;
            if (skC) {
                const float ic = oc.i;
                const float iR = tI[buf][ty + 1][tx + 2], iL = tI[buf][ty + 1][tx];
                const float iU = tI[buf][ty + 2][tx + 1], iD = tI[buf][ty][tx + 1];
                const float sR = tS[buf][ty + 1][tx + 2], sL = tS[buf][ty + 1][tx];
                const float sU = tS[buf][ty + 2][tx + 1], sD = tS[buf][ty][tx + 1];
                const float iF = inZ ? (skF ? of.i : INF) : ic;
                auto face = [&](float in) {
                    // computed before the select, so no branch skips it
                    const float v = faceFast(ic + in);
                    return isinf(in) ? 0.0f : v;
                };
                float r, l, u, d, f;
                if (!exact) {
                    r = face(iR), l = face(iL), u = face(iU), d = face(iD), f = face(iF);
                } else {
                    r = faceI(ic, iR), l = faceI(ic, iL), u = faceI(ic, iU), d = faceI(ic, iD), f = faceI(ic, iF);
                }
                float bb = skB ? fb : 0.0f;
                if (z == 0) bb = faceI(ic, ic);
                float w = 0.0f;
                w -= r;
                w -= l;
                w -= u;
                w -= d;
                w -= f;
                w -= bb;
                const float aL = x == 0 ? 0.0f : lo(l), aD = lo(d), aB = skB ? lo(bb) : 0.0f;
                const float o = w * snc + ((aL * sL + aD * sD + aB * snb + r * sR + u * sU + f * snf));
                fcur = f;
                bstf(wS, ob, pOff(z), snc);
                bstf(wA, ob, pOff(z), o);
                acc += (double)(o * snc);
            } else if (inb) {
                bstf(wS, ob, pOff(z), 0.0f);
                bstf(wA, ob, pOff(z), 0.0f);
            }
            snb = snc;
            snc = snf;
            fb = fcur;
            skB = skC;
            skC = skF;
            ghc = oc.gf;
        };
        for (int zg = z0; zg < z1; zg += 4) {
            step(zg, 0, o0, o1, o3, h0, h1);
            if (zg + 1 < z1) step(zg + 1, 1, o1, o2, o0, h1, h0);
            if (zg + 2 < z1) step(zg + 2, 0, o2, o3, o1, h0, h1);
            if (zg + 3 < z1) step(zg + 3, 1, o3, o0, o2, h1, h0);
        }
        acc = blockSum(acc, sh);
        if (tid256() == 0) pout[bid] = acc;
    }
};
hipError_t launch_pcg_A(EnvState& e, int it) {
    float* s_old = e.pcg[2 + ((it + 1) & 1)];
    float* s_new = e.pcg[2 + (it & 1)];
    const ZGrid zg = e.gA;
    // byte offsets from the chunk's base plane stay below 4 GiB
    if ((double)(zg.zc + 3) * (double)e.X * (double)e.Y * 4.0 >= 4294967296.0) return hipErrorInvalidValue;
    const size_t M = e.maxBlocks;
    // the first iteration reduces pcg_start's partials, later ones pcg_B's;
    // with the multigrid preconditioner the z.r partials are the V-cycle's
    // (e.mg_dot_n of them)
    const double* pm = e.partials + PART_BMAX * M;
    const double* ps = it == 0 ? e.partials + PART_START * M : e.partials + PART_BSUM * M;
    const unsigned np = it == 0 ? e.gZ.nblocks : e.gB.nblocks;
    const unsigned ns = e.mg_depth > 0 ? e.mg_dot_n : np;
    const bool trunc = e.cfg.pcg_mode == SKT_PCG_REFERENCE;
    auto go = [&](auto body) {
        return klaunch<decltype(body)>(e, zg.grid, B3, 0, geom(e), zg.zc, e.pcg[1], s_old, s_new, e.irho, e.pcg[4],
                                       pm, ps, np, ns, e.partials + PART_A * M, &e.sc->st[1], &e.sc->st[0], it,
                                       &e.sc->face_exact);
    };
    if (it == 0) return trunc ? go(KPcgA<true, true>{}) : go(KPcgA<true, false>{});
    return trunc ? go(KPcgA<false, true>{}) : go(KPcgA<false, false>{});
}

// Blocks of a kernel the device holds at once. A diffusion sweep's grid per
// field is one such round and the pressure operator's grid at most one (a
// second, partly filled round would nearly double a launch's time).
template <class K>
static long resident_blocks(K kernel, int device, int cus) {
    int nb = 0;
    if (cus <= 0 || hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) != hipSuccess) return 0;
    return (long)nb * cus;
}
ResidentBlocks resident_blocks(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) cus = 0;
    return ResidentBlocks{resident_blocks(&k_solo<KPcgA<false, true>, Geom, int, float*, float*, float*, float*, float*, const double*,
                                          const double*, unsigned, unsigned, double*, PcgState*, PcgState*, int, int*>,
                                          device, cus),
                          resident_blocks(&k_solo<KDiffuse<false>, Geom, int, unsigned, DiffBatch, float, float*, float*,
                                                  float*>,
                                          device, cus),
                          cus};
}

// Prologue PRO_B, then updatePandDiv (Game/kernelSky.cu:955-985) + the
// stop-test partials (Game/environment.cu:963-971) + applyPreconditionerGPU
// + z.r partials: p' += alpha s, r -= alpha As, z = M^-1 r.
// Elementwise, so a block walks whole rows (x fastest) of the Y*Z rows:
// tile b takes rows [b rpb, (b + 1) rpb), V cells per lane step (V = 4
// when X is a multiple of 4: 16-byte accesses). Lanes advance by 256 groups
// and carry their (column, y, z) along. Ground cells keep p and r (the same
// bits are stored back) and get z = 0. The partials of pcg_A and then the
// lane's first two groups are read before the prologue reduces the partials;
// after that each group's data is read two groups ahead (two register slots,
// the loop handles one group per slot per trip). Under the multigrid
// preconditioner (jacobi 0) the V-cycle that follows forms z and its z.r
// partials, so the pass reads no preconditioner and writes neither.
template <int V>
struct KPcgB {
    template <class... T>
    static __device__ __forceinline__ bool idle(const T&... a) {
        return nthArg<12>(a...)->done != 0;  // sin
    }
    static __device__ __forceinline__ void run(VB vb, const Geom& G, unsigned rpb, float* __restrict__ pp, float* __restrict__ r, const float* __restrict__ s,
                    const float* __restrict__ As, const float* __restrict__ pre, float* __restrict__ zz,
                    const double* __restrict__ pin, unsigned n, double* pmax, double* psum, const PcgState* sin,
                    PcgState* sout, int ref, int absmode, int jacobi) {
        typedef float __attribute__((ext_vector_type(V))) fv;
        typedef int __attribute__((ext_vector_type(V))) iv;
        __shared__ double sh[4];
        __shared__ PcgState S;
        PcgPart part;
        pcgPartLoad<PRO_B>(nullptr, pin, n, n, part);
        const int X = G.X, Y = G.Y, q = G.X / V;
        const long rows = (long)Y * G.Z;
        const long r0 = (long)vb.id * rpb, r1 = std::min(rows, r0 + (long)rpb);
        const int t = tid256();
        const long gend = r1 * q;
        struct Grp {
            fv P, R, S, A, Q;
        };
        auto load = [&](long gk, Grp& d) __attribute__((always_inline)) {
            if (gk < gend) {
                d.P = ((const fv*)pp)[gk];
                d.R = ((const fv*)r)[gk];
                d.S = ((const fv*)s)[gk];
                d.A = ((const fv*)As)[gk];
                if (jacobi) d.Q = ((const fv*)pre)[gk];
            }
        };
        long g = r0 * q + t;
        Grp d0, d1;
        load(g, d0);
        load(g + 256, d1);
        pcgPrologue<PRO_B>(nullptr, pin, n, n, part, sin, sout, ref, 0.0f, 0, S, sh, vb.id);
        if (S.done) return;
        const float alpha = S.alpha;
        float mf = 0.0f;
        double dz = 0.0;
        if (r0 < r1) {
            const long row = g / q;
            int gx = (int)(g - row * q), y = (int)(row % Y), z = (int)(row / Y);
            // a step of 256 groups: dr whole rows and dgx groups
            const int dr = 256 / q, dgx = 256 % q;
            auto group = [&](long gk, const Grp& d) __attribute__((always_inline)) {
                const iv gh = *(const iv*)(G.gh + gx * V + (long)X * z);
                fv Po = d.P, Ro = d.R, Zo;
                const bool ab = absmode || z == 0;
    #pragma unroll
                for (int c = 0; c < V; c++) {
                    float zv = 0.0f;
                    if (y > gh[c]) {
                        Po[c] = d.P[c] + alpha * d.S[c];
                        const float rv = d.R[c] - alpha * d.A[c];
                        Ro[c] = rv;
                        mf = fmaxf(mf, ab ? fabsf(rv) : rv);
                        if (jacobi) {
                            zv = rv * d.Q[c];
                            dz += (double)(zv * rv);
                        }
                    }
                    Zo[c] = zv;
                }
                ((fv*)pp)[gk] = Po;
                ((fv*)r)[gk] = Ro;
                if (jacobi) ((fv*)zz)[gk] = Zo;
                gx += dgx;
                y += dr;
                if (gx >= q) gx -= q, ++y;
                while (y >= Y) y -= Y, ++z;
            };
            for (; g < gend; g += 512) {
                group(g, d0);
                load(g + 512, d0);
                if (g + 256 < gend) {
                    group(g + 256, d1);
                    load(g + 768, d1);
                }
            }
        }
        const double m = blockMax((double)mf, sh);
        if (jacobi) dz = blockSum(dz, sh);
        if (t == 0) {
            pmax[vb.id] = m;
            if (jacobi) psum[vb.id] = dz;
        }
    }
};
hipError_t launch_pcg_B(EnvState& e, int it) {
    float* s_new = e.pcg[2 + (it & 1)];
    const int ref = e.cfg.pcg_mode == SKT_PCG_REFERENCE;
    const int absmode = ref ? e.cfg.maxabs_mode : 1;
    const size_t M = e.maxBlocks;
    auto go = [&](auto body) {
        return klaunch<decltype(body)>(e, dim3(e.gB.nblocks), B3, 0, geom(e), e.gB.rpb, e.f[SKT_PRESSURE_PROJ],
                                       e.pcg[0], s_new, e.pcg[4], e.pcg[5], e.pcg[1], e.partials + PART_A * M,
                                       e.gA.nblocks, e.partials + PART_BMAX * M, e.partials + PART_BSUM * M,
                                       &e.sc->st[0], &e.sc->st[1], ref, absmode, e.mg_depth > 0 ? 0 : 1);
    };
    return e.X % 4 == 0 ? go(KPcgB<4>{}) : go(KPcgB<1>{});
}

// Stop test after the last iteration (one block): slot 1 -> slot 0.
struct KPcgFinal {
    static __device__ __forceinline__ void run(VB vb, const double* __restrict__ pm, unsigned n, const PcgState* sin, PcgState* sout, int it) {
        __shared__ double sh[4];
        __shared__ PcgState S;
        PcgPart P;
        pcgPartLoad<PRO_FINAL>(pm, nullptr, n, n, P);
        pcgPrologue<PRO_FINAL>(pm, nullptr, n, n, P, sin, sout, 0, 0.0f, it, S, sh, vb.id);
    }
};
hipError_t launch_pcg_final(EnvState& e, int it) {
    return klaunch<KPcgFinal>(e, dim3(1), B3, 0, e.partials + (size_t)PART_BMAX * e.maxBlocks, e.gB.nblocks,
                              &e.sc->st[1], &e.sc->st[0], it);
}

// ------------------------------------------------- FP64 residual refinement ---
// The solves of one projection sum into x64 in FP64; each refinement pass
// solves A d = b - A x64 in FP32 with the residual formed in FP64. A is the
// FP32 solver's operator exactly (the face coefficients faceI of the two
// cells' 1/density and the diagonal summed in float, in the SpMV's order, with
// its boundary rules); only its products and sums are carried in FP64.
//
// b (the divergence) is kept before the first solve overwrites r.
struct KCopyN {
    static __device__ __forceinline__ void run(VB vb, const float* __restrict__ a, float* __restrict__ d, long n) {
        for (long i = vb.id * 256L + tid256(); i < n; i += (long)vb.n * 256) d[i] = a[i];
    }
};
// x64 = (first ? 0 : x64) + p' (the last solve's correction); p' = 0 for the
// next solve or, after the last one, float(x64). Tile 0 totals the
// iterations of the solve just finished (its final state is in st[0]).
struct KRefineAccum {
    static __device__ __forceinline__ void run(VB vb, long n, float* __restrict__ pp, double* __restrict__ x64,
                                               int first, int last, Scalars* sc) {
        for (long i = vb.id * 256L + tid256(); i < n; i += (long)vb.n * 256) {
            const double v = (first ? 0.0 : x64[i]) + (double)pp[i];
            x64[i] = v;
            pp[i] = last ? (float)v : 0.0f;
        }
        if (vb.id == 0 && tid256() == 0) {
            const PcgState& st = sc->st[0];
            RefineStats& R = sc->ref;
            if (first) {
                R.iters_total = st.iters;
                R.passes = 1;
                R.rhs64 = (double)st.maxr0;
            } else {
                R.iters_total += st.iters;
                R.passes += st.iters > 0 ? 1 : 0;
            }
        }
    }
};
// r = b - A x64 in FP64 over the divergence kernel's column march, written
// as float into r (when r is set) and its per-block max |r| into part, in the
// divergence's partial layout (so the next solve's start reads it as it reads
// the divergence's).
struct KRefineResidual {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int yc, unsigned nyc, const double* __restrict__ x64,
                                               const float* __restrict__ irho, const float* __restrict__ b,
                                               float* __restrict__ r, double* part) {
        __shared__ double sh[4];
        const unsigned lin = vb.id;
        const int chunk = (int)(lin % nyc);
        const unsigned tile = lin / nyc;
        const int gx = (G.X + 63) >> 6;
        const int x = (int)(tile % (unsigned)gx) * 64 + (int)threadIdx.x;
        const int z = (int)(tile / (unsigned)gx) * 4 + (int)threadIdx.y;
        double m = 0.0;
        if (x < G.X && z < G.Z) {
            const int X = G.X, Y = G.Y, Z = G.Z;
            const long XY = (long)X * Y;
            const long col = XY * z + x;
            const long gc = x + (long)X * z;
            const int gh = G.gh[gc];
            const bool inL = x > 0, inR = x < X - 1, inB = z > 0, inF = z < Z - 1;
            const int ghL = inL ? G.gh[gc - 1] : 0, ghR = inR ? G.gh[gc + 1] : 0;
            const int ghB = inB ? G.gh[gc - X] : 0, ghF = inF ? G.gh[gc + X] : 0;
            const int y0 = chunk * yc, y1 = min(Y, y0 + yc);
            for (int y = y0; y < y1; y++) {
                const long i = col + (long)X * y;
                double rv = 0.0;
                if (y > gh) {
                    const OpRow o = opRow(irho, i, x, y, z, X, Y, Z, XY, gh, ghL, ghR, ghB, ghF);
                    // neighbours at clamped offsets (unconditional loads),
                    // added in the same order where they are sky
                    const double xL = x64[i - (inL ? 1 : 0)], xD = x64[i - (y > 0 ? X : 0)];
                    const double xB = x64[i - (inB ? XY : 0)], xR = x64[i + (inR ? 1 : 0)];
                    const double xU = x64[i + (y + 1 < Y ? X : 0)], xF = x64[i + (inF ? XY : 0)];
                    double ax = (double)o.w * x64[i];
                    ax = o.sL ? ax + (double)o.fl * xL : ax;
                    ax = o.sD ? ax + (double)o.fd * xD : ax;
                    ax = o.sB ? ax + (double)o.fb * xB : ax;
                    ax = o.sR ? ax + (double)o.fr * xR : ax;
                    ax = o.sU ? ax + (double)o.fu * xU : ax;
                    ax = o.sF ? ax + (double)o.ff * xF : ax;
                    rv = (double)b[i] - ax;
                    m = fmax(m, fabs(rv));
                }
                if (r) r[i] = (float)rv;
            }
        }
        m = blockMax(m, sh);
        if (tid256() == 0) part[vb.id] = m;
    }
};
// The final residual's maximum over the partials into the refinement totals.
struct KRefineReport {
    static __device__ __forceinline__ void run(VB vb, const double* __restrict__ part, unsigned n, Scalars* sc) {
        __shared__ double sh[4];
        (void)vb;
        double m = 0.0;
        for (unsigned j = tid256(); j < n; j += 256) m = fmax(m, part[j]);
        m = blockMax(m, sh);
        if (tid256() == 0) sc->ref.res64 = m;
    }
};
hipError_t launch_refine_keep_rhs(EnvState& e) {
    return klaunch<KCopyN>(e, grid1(e.N), dim3(256), 0, (const float*)e.pcg[0], e.rhs, e.N);
}
hipError_t launch_refine_accum(EnvState& e, int first, int last) {
    return klaunch<KRefineAccum>(e, grid1(e.N), dim3(256), 0, e.N, e.f[SKT_PRESSURE_PROJ], e.x64, first, last, e.sc);
}
hipError_t launch_refine_residual(EnvState& e, int write_r) {
    const unsigned nyc = (unsigned)((e.Y + YC_DIV - 1) / YC_DIV);
    return klaunch<KRefineResidual>(e, dim3(divBlocks(e)), B3, 0, geom(e), YC_DIV, nyc, (const double*)e.x64,
                                    (const float*)e.irho, (const float*)e.rhs, write_r ? e.pcg[0] : nullptr,
                                    e.partials + (size_t)PART_DIV * e.maxBlocks);
}
hipError_t launch_refine_report(EnvState& e) {
    return klaunch<KRefineReport>(e, dim3(1), dim3(256), 0, (const double*)(e.partials + (size_t)PART_DIV * e.maxBlocks),
                                  (unsigned)divBlocks(e), e.sc);
}

// updatePressure (Game/kernelSky.cu:1016-1030) fused with the initDensity that
// follows it (Game/environment.cu:878-879). Each thread takes ZPT cells
// along z; their loads are unconditional (a ground cell's values are read
// and dropped), so they issue together.
struct KUpdatePressureDensity {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, float* __restrict__ p, const float* __restrict__ pp, float* __restrict__ rho,
                                      const float* __restrict__ th, const float* __restrict__ qv,
                                      const float* __restrict__ gP) {
        constexpr int zc = ZPT;
        ZM_SETUP_V;
        (void)bid;
        if (!inb || z0 >= z1) return;
        float lp[ZPT], lpp[ZPT], t[ZPT], q[ZPT], g[ZPT];
        int gh[ZPT];
#pragma unroll
        for (int k = 0; k < ZPT; k++) {
            const int z = min(z0 + k, z1 - 1);
            const long i = gidx(G, x, y, z);
            lp[k] = p[i];
            lpp[k] = pp[i];
            t[k] = th[i];
            q[k] = qv[i];
            g[k] = gP[x + (long)G.X * z];
            gh[k] = G.gh[x + G.X * z];
        }
#pragma unroll
        for (int k = 0; k < ZPT; k++) {
            const int z = z0 + k;
            if (z >= z1 || y <= gh[k]) continue;
            const long i = gidx(G, x, y, z);
            const float pn = lp[k] - lpp[k] / 100.0f;
            p[i] = pn;
            const float T = t[k] * powNear1(pn / g[k], kc::Rsd / kc::Cpd);
            const float Tv = T * (0.608f * q[k] + 1);
            rho[i] = pn * 100 / (kc::Rsd * Tv);
        }
    }
};
hipError_t launch_update_pressure_density(EnvState& e) {
    return klaunch<KUpdatePressureDensity>(e, dim3(gridZpt(e)), B3, 0, geom(e), e.f[SKT_PRESSURE], e.f[SKT_PRESSURE_PROJ],
                                           e.f[SKT_DENSITY], e.f[SKT_POTTEMP], e.f[SKT_QV], e.f[SKT_GROUND_P]);
}

// applyPresProjGPU, Game/kernelSky.cu:862-908 (pressure {D,N,N}, density
// {N,N,D}), optionally followed per cell by calculateNewPressure.
// A march up the columns like the divergence: each thread walks a chunk of
// YC_DIV levels of its (x, z) column, reads the ground heights of the +x and
// +z columns once and carries p' and the density of the level above into the
// next step, where they are the cell's own. Each neighbour value follows
// nbOwn (beyond the domain: the sides condition on the cell's own value;
// ground: the ground condition), with the same arithmetic per cell.
struct KApplyProjection {
    static __device__ __forceinline__ void run(VB vb, const Geom& G, int yc, unsigned nyc, const float* __restrict__ pp, const float* __restrict__ rho,
                               float* __restrict__ u, float* __restrict__ v, float* __restrict__ w, float* __restrict__ p,
                               const float* __restrict__ th, const float* __restrict__ qv, const float* __restrict__ gP,
                               int newp) {
        const unsigned lin = vb.id;
        const int chunk = (int)(lin % nyc);
        const unsigned tile = lin / nyc;
        const int gx = (G.X + 63) >> 6;
        const int x = (int)(tile % (unsigned)gx) * 64 + (int)threadIdx.x;
        const int z = (int)(tile / (unsigned)gx) * 4 + (int)threadIdx.y;
        if (x >= G.X || z >= G.Z) return;
        const int X = G.X, Y = G.Y, Z = G.Z;
        const long XY = (long)X * Y;
        const long col = XY * z + x;  // cell (x, 0, z)
        const long gc = x + (long)X * z;
        const int gh = G.gh[gc];
        const bool inR = x < X - 1, inF = z < Z - 1;
        const int ghR = inR ? G.gh[gc + 1] : 0, ghF = inF ? G.gh[gc + X] : 0;
        const Bounds bp{BC_DIRICHLET, BC_NEUMANN, BC_NEUMANN}, bd{BC_NEUMANN, BC_NEUMANN, BC_DIRICHLET};
        const float gps = gP[gc];
        const int y0 = max(chunk * yc, gh + 1), y1 = min(Y, chunk * yc + yc);
        if (y0 >= y1) return;
        // p' and density of the cell, carried with 1/density from the level
        // below (correctly rounded reciprocals, the reference's bits). The
        // loads of a group of DIV_G levels are issued together (at clamped
        // in-range positions, so that they are unconditional) before the
        // group is computed.
        const long oR = inR ? 1 : 0, oF = inF ? XY : 0;
        float cp = pp[col + (long)X * y0], cd = rho[col + (long)X * y0];
        float icd = recipRN(cd);
        for (int yb = y0; yb < y1; yb += DIV_G) {
            float qpU[DIV_G], qdU[DIV_G], qpR[DIV_G], qpF[DIV_G], qdR[DIV_G], qdF[DIV_G];
            float lu[DIV_G], lv[DIV_G], lw[DIV_G], qp[DIV_G], qt[DIV_G], qq[DIV_G];
#pragma unroll
            for (int k = 0; k < DIV_G; k++) {
                const int y = min(yb + k, y1 - 1);
                const long i = col + (long)X * y;
                const long iu = y + 1 <= Y - 1 ? i + X : i;
                qpU[k] = pp[iu];
                qdU[k] = rho[iu];
                qpR[k] = pp[i + oR];
                qpF[k] = pp[i + oF];
                qdR[k] = rho[i + oR];
                qdF[k] = rho[i + oF];
                lu[k] = u[i];
                lv[k] = v[i];
                lw[k] = w[i];
                if (newp) {
                    qp[k] = p[i];
                    qt[k] = th[i];
                    qq[k] = qv[i];
                }
            }
#pragma unroll
            for (int k = 0; k < DIV_G; k++) {
                const int y = yb + k;
                if (y >= y1) break;
                const long i = col + (long)X * y;
                const bool up = y + 1 <= Y - 1;  // the level above a sky cell is sky
                const float ppU = up ? qpU[k] : 0.0f, rhoU = up ? qdU[k] : 0.0f;
                const float pr = !inR ? bcv(bp.sides, cp, 0.0f) : (y <= ghR ? bcv(bp.ground, cp, 0.0f) : qpR[k]);
                const float pu = !up ? bcv(bp.sides, cp, 0.0f) : ppU;
                const float pf = !inF ? bcv(bp.sides, cp, 0.0f) : (y <= ghF ? bcv(bp.ground, cp, 0.0f) : qpF[k]);
                float dr = !inR ? bcv(bd.sides, cd, 0.0f) : (y <= ghR ? bcv(bd.ground, cd, 0.0f) : qdR[k]);
                float du = !up ? bcv(bd.sides, cd, 0.0f) : rhoU;
                float df = !inF ? bcv(bd.sides, cd, 0.0f) : (y <= ghF ? bcv(bd.ground, cd, 0.0f) : qdF[k]);
                // one range test for the cell (its own 1/density and every sky
                // neighbour density it takes a reciprocal of), then the fast
                // forms throughout
                const bool fast = densOk(icd) && (!up || densOk(rhoU)) && (!(inR && y > ghR) || densOk(qdR[k])) &&
                                  (!(inF && y > ghF) || densOk(qdF[k]));
                auto harm = [&](auto fastTag) __attribute__((always_inline)) {
                    constexpr bool F = decltype(fastTag)::value;
                    auto rcp = [&](float d) { return F ? recipRNfast(d) : recipRN(d); };
                    auto two = [&](float d) { return F ? 2.0f * recipRNfast(d) : twoOverRN(d); };
                    const float iU = up ? rcp(rhoU) : 0.0f;
                    dr = dr == 0.0f ? cd : two(icd + (inR ? rcp(dr) : icd));
                    du = du == 0.0f ? cd : two(icd + (up ? iU : icd));
                    df = df == 0.0f ? cd : two(icd + (inF ? rcp(df) : icd));
                    return iU;
                };
                const float icU = fast ? harm(std::true_type{}) : harm(std::false_type{});
                u[i] = lu[k] + (pr - cp) / dr;
                v[i] = lv[k] + (pu - cp) / du;
                w[i] = lw[k] + (pf - cp) / df;
                if (newp) {
                    const float T = qt[k] * powNear1(qp[k] / gps, kc::Rsd / kc::Cpd);
                    const float Tv = T * (0.608f * qq[k] + 1);
                    p[i] = cd * kc::Rsd * Tv / 100.0f;
                }
                cp = ppU;
                cd = rhoU;
                icd = icU;
            }
        }
    }
};
hipError_t launch_apply_projection(EnvState& e, int new_pressure) {
    const unsigned nyc = (unsigned)((e.Y + YC_DIV - 1) / YC_DIV);
    return klaunch<KApplyProjection>(e, dim3(divBlocks(e)), B3, 0, geom(e), YC_DIV, nyc, e.f[SKT_PRESSURE_PROJ],
                                     e.f[SKT_DENSITY], e.f[SKT_VELX], e.f[SKT_VELY], e.f[SKT_VELZ], e.f[SKT_PRESSURE],
                                     e.f[SKT_POTTEMP], e.f[SKT_QV], e.f[SKT_GROUND_P], new_pressure);
}

}  // namespace skt
