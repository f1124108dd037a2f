// 3D Cartesian MUSCL-Hancock + HLLC step, its first-order fallback, the
// chemistry half-step, CFL maxima and conserved totals.
//
// One fused kernel does the reference's whole `_hydro` second-order pass
// (python/cfdSolver/torchSolver3D.py:122-203): ghost cells by index maps,
// limited slopes, the half-step predictor with its positivity test, HLLC on
// every face, the conservative update, the species transport of `_advect`
// (:246-256), the outward boundary-flux sums of `Simulation.step`
// (torchSession.py:498-508) and the CFL maxima of `_dt` (:356-360).
//
// Layout: state is structure-of-arrays, field-major, x fastest, exactly the
// reference's (5, nz, ny, nx) C-order array. A block owns a tile of
// (BX-2) x (BY-2) columns plus a one-cell halo ring of threads and marches
// up a chunk of z planes: x neighbours travel by lane shuffles (a block row
// is one 64-lane wave, or half of one for BX = 32), y neighbours through
// LDS, z neighbours stay in registers. Each HLLC face is evaluated once per
// tile; the update of plane k-1 completes when plane k's predictor has
// produced the z face between them.
#include "fv_device.h"
#include "fv_launch.h"

#include <string.h>

namespace FV_NS {
using namespace fv;
using fvl::Grid;
using fvl::Step;

struct G3 {
    int nx, ny, nz, bcx, bcy, bcz;
};

__device__ __forceinline__ long cell(int x, int y, int z, int nx, int ny) {
    return ((long)z * ny + y) * nx + x;
}

// 5 components of ghost-mapped cell (x, y, z); reflective ghosts negate the
// momentum normal to that wall. The energy row comes from Ue (U itself for
// a whole state).
__device__ __forceinline__ void load5(const real* __restrict__ U, const real* __restrict__ Ue, long N, const G3& g,
                                      int x, int y, int z, real* a) {
    long i = cell(ghost_src(x, g.nx, g.bcx), ghost_src(y, g.ny, g.bcy), ghost_src(z, g.nz, g.bcz), g.nx, g.ny);
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = U[c * N + i];
    a[4] = Ue[4 * N + i];
    if (ghost_flip(x, g.nx, g.bcx)) a[1] = -a[1];
    if (ghost_flip(y, g.ny, g.bcy)) a[2] = -a[2];
    if (ghost_flip(z, g.nz, g.bcz)) a[3] = -a[3];
}

__device__ __forceinline__ real load1(const real* __restrict__ q, const G3& g, int x, int y, int z) {
    return q[cell(ghost_src(x, g.nx, g.bcx), ghost_src(y, g.ny, g.bcy), ghost_src(z, g.nz, g.bcz), g.nx, g.ny)];
}

// Fold one block's partial results: maxima and sums over the block's
// threads, written to partials[block * NSLOT + slot] by thread 0. Passes
// without outward sums (SUMS false: chemistry, CFL passes) or without a
// source sum (SRC false) skip those wave reductions and store +0, which is
// what reducing their all-zero inputs gives.
template <int NT, bool SUMS = true, bool SRC = true>
__device__ void block_partials(double (&acc)[5], double (&mx)[3], double src, double* partials, int* flags,
                               int bad, long block = -1) {
    __shared__ double red[NT / 64 > 0 ? NT / 64 : 1][9];
    __shared__ int sbad;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    if (tid == 0) sbad = 0;
    __syncthreads();
    if (bad) atomicOr(&sbad, bad);
    double v[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) v[i] = wave_max(mx[i]);
#pragma unroll
    for (int i = 0; i < 5; ++i) v[3 + i] = SUMS ? wave_sum(acc[i]) : 0.0;
    v[8] = SRC ? wave_sum(src) : 0.0;
    const int w = tid / 64, nw = (NT + 63) / 64;
    if ((tid & 63) == 0) {
#pragma unroll
        for (int i = 0; i < 9; ++i) red[w][i] = v[i];
    }
    __syncthreads();
    if (tid == 0) {
        const long b =
            block >= 0 ? block : blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
        double* p = partials + b * fvl::NSLOT;
        for (int i = 0; i < 9; ++i) {
            double r = red[0][i];
            if ((i < 3) | (SUMS & (i < 8)) | (SRC & (i == 8)))
                for (int k = 1; k < nw; ++k) r = i < 3 ? fmax(r, red[k][i]) : r + red[k][i];
            p[i < 3 ? i : (i < 8 ? fvl::S_OUT + (i - 3) : fvl::S_SRC)] = r;
        }
        if (sbad) atomicOr(flags, sbad);
    }
}

// |u_m| + a of a committed 3D state (torchSession.py:356-360) in the storage
// precision. Every 3D CFL maximum goes through this one formula, whether the
// update, the chemistry half-step or a standalone pass finds it, so a
// state's dt does not depend on which pass saw it last (a resumed session
// steps on exactly). Contraction is off so it rounds the same wherever it
// is inlined.
__device__ __forceinline__ void speeds3(const real* o, real g, real (&v)[3]) {
#pragma clang fp contract(off)
    const real ir = crcp(o[0]);
    const real p = (g - (real)1) * (o[4] - ((real)0.5 * (o[1] * o[1] + o[2] * o[2] + o[3] * o[3])) * ir);
    const real a = csqrt((g * p) * ir);
#pragma unroll
    for (int m = 0; m < 3; ++m) v[m] = fabs(o[m + 1] * ir) + a;
}

// Maxima of speeds3 kept in the storage precision (widening to double
// afterwards is exact, so this equals reducing in double); a NaN sticks.
__device__ __forceinline__ void wave3r(const real* o, real g, real (&mx)[3]) {
    real v[3];
    speeds3(o, g, v);
#pragma unroll
    for (int m = 0; m < 3; ++m) mx[m] = (v[m] != v[m] || v[m] > mx[m]) ? v[m] : mx[m];
}

// Maxima of speeds3 reduced in double.
__device__ __forceinline__ void wave3(const real* o, double g, double (&mx)[3]) {
    real v[3];
    speeds3(o, (real)g, v);
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        const double w = (double)v[m];
        mx[m] = (w != w || w > mx[m]) ? w : mx[m];
    }
}

// Waves per SIMD the register allocation of a step kernel aims for. A
// 256-thread FP32 tile with at most one transported fraction fits three
// blocks per CU at <= 168 VGPRs, which hides the z-march's load and barrier
// latency better than two; 256-thread FP64 tiles and FP32 tiles carrying two
// fractions are held to two (<= 256 VGPRs, no accumulation registers) so two
// blocks share a CU instead of one. Larger blocks run one per CU: the
// 768-thread FP32 tile (32x24) then gets three waves per SIMD (<= 168
// VGPRs; with two fractions 6 VGPRs spill, still faster than 32x8 at two
// waves), 512-thread tiles two.
constexpr int occupancy_of(int threads, int nq = 0) {
    return threads == 256 ? (sizeof(real) == 4 && nq < 2 ? 3 : 2) : 1;
}

// LDS footprint of one tile: the A plane with a two-cell halo, the y face
// states and the y fluxes (separate buffers, so a plane needs two barriers:
// A plane written, y face states written; the y fluxes written after the
// second are read after the next plane's first).
template <int BX, int BY, int NQ_>
struct Tile3 {
    static constexpr int NQ = NQ_;
    static constexpr int PW = BX + 2, PH = BY + 2;  // plane: x in [-2, TX+1], y in [-2, TY+1]
    static constexpr int NYS = 6 + NQ;              // y-face payload: state, gamma, fractions
    static constexpr int NFY = 5 + NQ;
    static constexpr int PLANE = 5 * PW * PH;
    static constexpr int FLUX = NFY * BX * BY;
    static constexpr int BUF = PLANE;
    static constexpr int YS = NYS * BY * BX;
    // FP32 tiles stage their global reads through LDS a plane ahead: each
    // thread's own cell (5 fields, then lambda or gamma and the fractions:
    // OWN rows) and the ring cells (5 fields), double-buffered by plane
    // parity, the ring padded to whole waves.
    static constexpr bool STAGED = sizeof(real) == sizeof(float);
    static constexpr int OWN = NQ_ == 0 ? 5 : 6 + NQ_;
    static constexpr int RING = (2 * PH + 2 * BX + 63) / 64 * 64;
    static constexpr int DMA = STAGED ? 2 * (OWN * BX * BY + 5 * RING) : 1;
};

using fvl::Coef;
using fvl::coef_of;

// One tile (tbx, tby) of one z chunk tbz; `pblk` is the index of this
// block's slot in s.partials.
template <int LIM, int BX, int BY, int NQ_>
__device__ __forceinline__ void step3d_tile(const Step& s, const G3& g, int tbx, int tby, int tbz, long pblk,
                                            real* sbuf, real* sybuf, real* sfbuf, double* sacc, real* sdma) {
    using T = Tile3<BX, BY, NQ_>;
    constexpr int TX = BX - 2, TY = BY - 2;
    constexpr int NQ = T::NQ;
    constexpr int PW = T::PW, PH = T::PH;
    real(*sA)[PH][PW] = reinterpret_cast<real(*)[PH][PW]>(sbuf);
    real(*sF)[BY][BX] = reinterpret_cast<real(*)[BY][BX]>(sfbuf);
    real(*sY)[BY][BX] = reinterpret_cast<real(*)[BY][BX]>(sybuf);

    const int lane = threadIdx.x, row = threadIdx.y;
    const int lx = lane - 1, ly = row - 1;
    const int bx0 = tbx * TX, by0 = tby * TY;
    const int x = bx0 + lx, y = by0 + ly;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long N = (long)nx * ny * nz;
    const int z0 = tbz * s.zchunk, z1 = min(z0 + s.zchunk, nz);
    const real* __restrict__ U = (const real*)s.U;
    const real* __restrict__ Ue = s.Ue ? (const real*)s.Ue : U;
    real* __restrict__ Uo = (real*)s.Uout;
    const real* __restrict__ lamp = (const real*)s.lam;
    const int nq = NQ;  // the launcher picks the instance whose s.nq equals NQ

    const bool in_active = x >= -1 && x <= nx && y >= -1 && y <= ny;
    const bool owns = lx >= 0 && lx < TX && ly >= 0 && ly < TY && x < nx && y < ny;
    const bool xface = lx >= 0 && ly >= 0 && ly < TY && x <= nx && y < ny;
    const bool yface = lx >= 0 && lx < TX && ly >= 0 && x < nx && y <= ny;
    const Coef K = coef_of(s);
    const real c0 = (real)K.c[0], c1 = (real)K.c[1], c2 = (real)K.c[2];
    const real hc[3] = {(real)K.hc[0], (real)K.hc[1], (real)K.hc[2]};
    const double w0 = K.w[0], w1 = K.w[1], w2 = K.w[2];
    const real gconst = (real)s.gamma;

    // Ring of the LDS plane not covered by the block's own columns.
    const int tid = lane + BX * row;
    const int nring = 2 * PH + 2 * BX;
    int rx = 0, ry = 0;
    if (tid < 2 * PH) {
        rx = (tid & 1) ? TX + 1 : -2;
        ry = tid / 2 - 2;
    } else if (tid < nring) {
        int t = tid - 2 * PH;
        rx = t / 2 - 1;
        ry = (t & 1) ? TY + 1 : -2;
    }

    real Ap[5], Ac[5], An[5];
    real gp = gconst, gc = gconst, gn = gconst;
    real qp[2] = {0, 0}, qc[2] = {0, 0}, qn[2] = {0, 0};
    real zr[5] = {0, 0, 0, 0, 0}, Fzp[5] = {0, 0, 0, 0, 0}, sfzp[2] = {0, 0};
    real R[5] = {0, 0, 0, 0, 0}, rq[2] = {0, 0};
    // The y part of a plane's partial update is applied after the next
    // plane's first barrier, from the y fluxes left in sF (one barrier fewer
    // per plane).
    bool pend = false;
    // Outward boundary sums, per thread in plane order: kept in the thread's
    // own LDS slot (only boundary faces touch it) instead of ten registers.
    constexpr int NT = BX * BY;
#pragma unroll
    for (int c = 0; c < 5; ++c) sacc[c * NT + lane + BX * row] = 0.0;
    double mx[3];
    real mxr[3] = {0, 0, 0};
    int bad = 0;

    load5(U, Ue, N, g, x, y, z0 - 2, Ap);
    load5(U, Ue, N, g, x, y, z0 - 1, Ac);
    // FP64 tiles load their ring cell one plane ahead into registers (after
    // the slopes, so the load overlaps the plane's arithmetic): 2-6% faster
    // there. FP32 tiles have no registers to spare and instead copy each
    // thread's next-but-one cell (with its lambda or gamma and fractions)
    // and the next ring cell from global memory straight into LDS
    // (global_load_lds, no registers), issued right after a plane's first
    // barrier: the barrier after it waits for the copy (a pending LDS
    // write), so its latency hides behind the slopes and the predictor, and
    // the next plane starts from LDS instead of waiting on global loads at
    // its first barrier (256^3 Sod step 7-10% faster, reactive batch 7%,
    // same results bitwise). Staging the state alone while lambda and the
    // fractions stayed register loads spilled and was 6% slower there.
    constexpr bool kRingAhead = sizeof(real) == sizeof(double);
    constexpr bool kDma = T::STAGED;
    constexpr int RING = T::RING, OWN = T::OWN;
    real* sown = sdma;                        // [2][OWN][NT]: own cell, by plane parity
    real* sring = sdma + 2 * OWN * BX * BY;  // [2][5][RING]: ring cell, by plane parity
    const int wbase = tid & ~63;        // first thread of this wave: its LDS rows start there
    const int ox = ghost_src(x, nx, g.bcx), oy = ghost_src(y, ny, g.bcy);
    const int qx = ghost_src(bx0 + rx, nx, g.bcx), qy = ghost_src(by0 + ry, ny, g.bcy);
    // Copies of plane kk's own cells and plane kr's ring cells into LDS
    // (each lane's value lands at its wave's row base + 4 x lane).
    auto stage = [&](int kk, int kr) {
        const long io = ((long)ghost_src(kk, nz, g.bcz) * ny + oy) * nx + ox;
        real* d = sown + (kk & 1) * OWN * BX * BY + wbase;
#pragma unroll
        for (int c = 0; c < 4; ++c) __builtin_amdgcn_global_load_lds(U + c * N + io, d + c * BX * BY, 4, 0, 0);
        __builtin_amdgcn_global_load_lds(Ue + 4 * N + io, d + 4 * BX * BY, 4, 0, 0);
        if constexpr (OWN > 5) {
            if (lamp) __builtin_amdgcn_global_load_lds(lamp + io, d + 5 * BX * BY, 4, 0, 0);
            for (int j = 0; j < NQ; ++j)
                if (j < nq)
                    __builtin_amdgcn_global_load_lds((const real*)s.q[j] + io, d + (6 + j) * BX * BY, 4, 0, 0);
        }
        if (tid < nring) {
            const long ir = ((long)ghost_src(kr, nz, g.bcz) * ny + qy) * nx + qx;
            real* e = sring + (kr & 1) * 5 * RING + wbase;
#pragma unroll
            for (int c = 0; c < 4; ++c) __builtin_amdgcn_global_load_lds(U + c * N + ir, e + c * RING, 4, 0, 0);
            __builtin_amdgcn_global_load_lds(Ue + 4 * N + ir, e + 4 * RING, 4, 0, 0);
        }
    };
    real rg[5] = {0, 0, 0, 0, 0};
    if (kRingAhead && tid < nring) load5(U, Ue, N, g, bx0 + rx, by0 + ry, z0 - 1, rg);
    if constexpr (kDma) stage(z0, z0 - 1);
    const bool direct = s.lam_is_gamma != 0;
    if (lamp) {
        gp = gamma_from(load1(lamp, g, x, y, z0 - 2), direct);
        gc = gamma_from(load1(lamp, g, x, y, z0 - 1), direct);
    }
    for (int j = 0; j < NQ; ++j) if (j < nq) {
        qp[j] = load1((const real*)s.q[j], g, x, y, z0 - 2);
        qc[j] = load1((const real*)s.q[j], g, x, y, z0 - 1);
    }

    for (int k = z0 - 1; k <= z1; ++k) {
        if constexpr (kDma) {
            // Plane k+1, staged in LDS during the previous plane (or before
            // the loop); the same sign flips as load5.
            const real* so = sown + ((k + 1) & 1) * OWN * BX * BY + tid;
#pragma unroll
            for (int c = 0; c < 5; ++c) An[c] = so[c * BX * BY];
            if (ghost_flip(x, nx, g.bcx)) An[1] = -An[1];
            if (ghost_flip(y, ny, g.bcy)) An[2] = -An[2];
            if (ghost_flip(k + 1, nz, g.bcz)) An[3] = -An[3];
            if constexpr (OWN > 5) {
                if (lamp) gn = gamma_from(so[5 * BX * BY], direct);
                for (int j = 0; j < NQ; ++j) if (j < nq) qn[j] = so[(6 + j) * BX * BY];
            } else {
                if (lamp) gn = gamma_from(load1(lamp, g, x, y, k + 1), direct);
            }
        } else {
            load5(U, Ue, N, g, x, y, k + 1, An);
            if (lamp) gn = gamma_from(load1(lamp, g, x, y, k + 1), direct);
            for (int j = 0; j < NQ; ++j) if (j < nq) qn[j] = load1((const real*)s.q[j], g, x, y, k + 1);
        }

        // No barrier needed before rewriting sA: its last readers finished
        // before the previous plane's second barrier.
#pragma unroll
        for (int c = 0; c < 5; ++c) sA[c][ly + 2][lx + 2] = Ac[c];
        if (tid < nring) {
            if constexpr (kDma) {
                const real* sr = sring + (k & 1) * 5 * RING + tid;
#pragma unroll
                for (int c = 0; c < 5; ++c) rg[c] = sr[c * RING];
                if (ghost_flip(bx0 + rx, nx, g.bcx)) rg[1] = -rg[1];
                if (ghost_flip(by0 + ry, ny, g.bcy)) rg[2] = -rg[2];
                if (ghost_flip(k, nz, g.bcz)) rg[3] = -rg[3];
            } else if (!kRingAhead) {
                load5(U, Ue, N, g, bx0 + rx, by0 + ry, k, rg);
            }
#pragma unroll
            for (int c = 0; c < 5; ++c) sA[c][ry + 2][rx + 2] = rg[c];
        }
        __syncthreads();
        // The next plane's staging: its LDS rows were last read before this
        // barrier (a plane ago, by the same lanes).
        if constexpr (kDma)
            if (k < z1) stage(k + 2, k + 1);

        // y part of the previous plane's partial update: its y fluxes were
        // stored in sF before the barrier above, and sF is next written after
        // this plane's second barrier.
        if (pend) {
#pragma unroll
            for (int c = 0; c < 5; ++c) R[c] = R[c] - c1 * (sF[c][row + 1][lane] - sF[c][row][lane]);
            for (int j = 0; j < NQ; ++j)
                if (j < nq) rq[j] = rq[j] - c1 * (sF[5 + j][row + 1][lane] - sF[5 + j][row][lane]);
        }

        real sx[5], sy[5], sz[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            sx[c] = limited<LIM, 3>(Ac[c] - sA[c][ly + 2][lx + 1], sA[c][ly + 2][lx + 3] - Ac[c]);
            sy[c] = limited<LIM, 3>(Ac[c] - sA[c][ly + 1][lx + 2], sA[c][ly + 3][lx + 2] - Ac[c]);
            sz[c] = limited<LIM, 3>(Ac[c] - Ap[c], An[c] - Ac[c]);
        }
        if (kRingAhead && k < z1 && tid < nring) load5(U, Ue, N, g, bx0 + rx, by0 + ry, k + 1, rg);
        // Half-step predictor (torchSolver3D.py:140-151).
        real P[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) P[c] = Ac[c];
        bool flat[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const real* sl = m == 0 ? sx : (m == 1 ? sy : sz);
            flat[m] = (sl[0] == (real)0) & (sl[1] == (real)0) & (sl[2] == (real)0) & (sl[3] == (real)0) &
                      (sl[4] == (real)0);
        }
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const real* sl = m == 0 ? sx : (m == 1 ? sy : sz);
            // Zero slopes give equal half states, equal fluxes and an exact
            // zero change: skip the two flux evaluations.
            if (!flat[m]) {
                real up[5], um[5], fp[5], fm[5];
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    up[c] = Ac[c] + (real)0.5 * sl[c];
                    um[c] = Ac[c] - (real)0.5 * sl[c];
                }
                flux3(up, gc, m + 1, fp);
                flux3(um, gc, m + 1, fm);
#pragma unroll
                for (int c = 0; c < 5; ++c) P[c] = P[c] - hc[m] * (fp[c] - fm[c]);
            }
        }
        // Face states: P -+ s/2 (torchSolver3D.py:178-180).
        real xl[5], xr[5], yl[5], yr[5], zl[5], zn[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            real hx = (real)0.5 * sx[c], hy = (real)0.5 * sy[c], hz = (real)0.5 * sz[c];
            xl[c] = P[c] - hx; xr[c] = P[c] + hx;
            yl[c] = P[c] - hy; yr[c] = P[c] + hy;
            zl[c] = P[c] - hz; zn[c] = P[c] + hz;
        }
        // Positivity of every active predicted face state (:152-159).
        if (in_active && k >= -1 && k <= nz) {
            // Both face states of a flat direction are P itself.
            bool ok = true;
            if (flat[0] & flat[1] & flat[2]) {
                ok = physical5(P);
            } else {
                ok = physical5(xl) & physical5(xr) & physical5(yl) & physical5(yr) & physical5(zl) & physical5(zn);
            }
            if (!ok) bad |= fvl::F_PRED_BAD;
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) sY[c][row][lane] = yr[c];
        sY[5][row][lane] = gc;
        for (int j = 0; j < NQ; ++j) sY[6 + j][row][lane] = qc[j];
        __syncthreads();  // sY visible

        const bool plane_out = k >= z0 && k < z1;
        // x face x-1/2 from the left neighbour's right state.
        real lxr[5], Fx[5] = {0, 0, 0, 0, 0}, sfx[2] = {0, 0};
#pragma unroll
        for (int c = 0; c < 5; ++c) lxr[c] = shfl_up1(xr[c]);
        real lg = shfl_up1(gc);
        real lq[2] = {0, 0};
        for (int j = 0; j < NQ; ++j) lq[j] = shfl_up1(qc[j]);
        if (plane_out && xface) {
            hllc_cart5<true>(lxr, xl, lg, gc, 0, Fx);
            for (int j = 0; j < NQ; ++j) if (j < nq) sfx[j] = Fx[0] * (Fx[0] >= (real)0 ? lq[j] : qc[j]);
        }
        // y face y-1/2 from the row below.
        real Fy[5] = {0, 0, 0, 0, 0}, sfy[2] = {0, 0};
        if (plane_out && yface) {
            real below[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) below[c] = sY[c][row - 1][lane];
            hllc_cart5<true>(below, yl, sY[5][row - 1][lane], gc, 1, Fy);
            for (int j = 0; j < NQ; ++j) if (j < nq) sfy[j] = Fy[0] * (Fy[0] >= (real)0 ? sY[6 + j][row - 1][lane] : qc[j]);
        }
        // z face k-1/2 from the carried state of plane k-1.
        real Fz[5] = {0, 0, 0, 0, 0}, sfz[2] = {0, 0};
        const bool zface = owns && k >= z0;
        if (zface) {
            hllc_cart5<true>(zr, zl, gp, gc, 2, Fz);
            for (int j = 0; j < NQ; ++j) if (j < nq) sfz[j] = Fz[0] * (Fz[0] >= (real)0 ? qp[j] : qc[j]);
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) sF[c][row][lane] = Fy[c];
        for (int j = 0; j < NQ; ++j) sF[5 + j][row][lane] = sfy[j];

        real FxR[5], sfxR[2] = {0, 0};
#pragma unroll
        for (int c = 0; c < 5; ++c) FxR[c] = shfl_down1(Fx[c]);
        for (int j = 0; j < NQ; ++j) sfxR[j] = shfl_down1(sfx[j]);

        // Complete plane k-1 with the z faces (torchSolver3D.py:192-202).
        if (owns && k - 1 >= z0) {
            real o[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) o[c] = R[c] - c2 * (Fz[c] - Fzp[c]);
            const long i = cell(x, y, k - 1, nx, ny);
#pragma unroll
            for (int c = 0; c < 5; ++c) Uo[c * N + i] = o[c];
            if (!physical5(o)) bad |= fvl::F_UPD_BAD;
            if (s.want_cfl) wave3r(o, gconst, mxr);
            for (int j = 0; j < NQ; ++j) if (j < nq) {
                real v = rq[j] - c2 * (sfz[j] - sfzp[j]);
                real f = v / o[0];
                if (f < -kSpecTol || f > (real)1 + kSpecTol || f != f) bad |= fvl::F_SPEC_BAD;
                ((real*)s.qout[j])[i] = rmin(rmax(f, (real)0), (real)1);
            }
        }
        // Partial update of plane k with the x faces; the y faces follow
        // after the next plane's first barrier.
        pend = owns && plane_out;
        if (pend) {
#pragma unroll
            for (int c = 0; c < 5; ++c) R[c] = Ac[c] - c0 * (FxR[c] - Fx[c]);
            for (int j = 0; j < NQ; ++j)
                if (j < nq) rq[j] = Ac[0] * qc[j] - c0 * (sfxR[j] - sfx[j]);
        }
        // Outward boundary fluxes and optional face dumps.
        if (plane_out && xface) {
            if (x == 0 || x == nx) {
                double sgn = x == 0 ? -w0 : w0;
#pragma unroll
                for (int c = 0; c < 5; ++c) sacc[c * NT + tid] += sgn * (double)Fx[c];
            }
            if (s.F[0]) {
                long f = ((long)k * ny + y) * (nx + 1) + x, M = (long)nz * ny * (nx + 1);
#pragma unroll
                for (int c = 0; c < 5; ++c) ((real*)s.F[0])[c * M + f] = Fx[c];
            }
        }
        if (plane_out && yface) {
            if (y == 0 || y == ny) {
                double sgn = y == 0 ? -w1 : w1;
#pragma unroll
                for (int c = 0; c < 5; ++c) sacc[c * NT + tid] += sgn * (double)Fy[c];
            }
            if (s.F[1]) {
                long f = ((long)k * (ny + 1) + y) * nx + x, M = (long)nz * (ny + 1) * nx;
#pragma unroll
                for (int c = 0; c < 5; ++c) ((real*)s.F[1])[c * M + f] = Fy[c];
            }
        }
        if (zface) {
            if (k == 0 || k == nz) {
                double sgn = k == 0 ? -w2 : w2;
#pragma unroll
                for (int c = 0; c < 5; ++c) sacc[c * NT + tid] += sgn * (double)Fz[c];
            }
            if (s.F[2]) {
                long f = ((long)k * ny + y) * nx + x, M = (long)(nz + 1) * ny * nx;
#pragma unroll
                for (int c = 0; c < 5; ++c) ((real*)s.F[2])[c * M + f] = Fz[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            Ap[c] = Ac[c];
            Ac[c] = An[c];
            zr[c] = zn[c];
            Fzp[c] = Fz[c];
        }
        gp = gc;
        gc = gn;
        for (int j = 0; j < NQ; ++j) {
            qp[j] = qc[j];
            qc[j] = qn[j];
            sfzp[j] = sfz[j];
        }
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) mx[m] = (double)mxr[m];
    double acc[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) acc[c] = sacc[c * NT + tid];
    block_partials<BX * BY>(acc, mx, 0.0, s.partials, s.flags, bad, pblk);
}

template <int LIM, int BX, int BY, int NQ>
__global__ void __launch_bounds__(BX* BY, occupancy_of(BX* BY, NQ)) k_step3d(Step s, G3 g) {
    using T = Tile3<BX, BY, NQ>;
    __shared__ real sbuf[T::BUF];
    __shared__ real sy[T::YS];
    __shared__ real sf[T::FLUX];
    __shared__ double sacc[5 * BX * BY];
    __shared__ real sdma[T::DMA];
    const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    step3d_tile<LIM, BX, BY, NQ>(s, g, blockIdx.x, blockIdx.y, blockIdx.z, b, sbuf, sy, sf, sacc, sdma);
}

// Work-list form: blocks of many instances of different sizes and
// boundaries (one limiter per launch) in one launch. Block b belongs to the
// instance i with first[i] <= b < first[i+1]; its tile follows from that
// instance's tile counts. Instances that are idle this step (coef active flag
// clear) only publish empty partials.
template <int LIM, int BX, int BY, int NQ>
__global__ void __launch_bounds__(BX* BY, occupancy_of(BX* BY, NQ)) k_batch3d(const fvl::BatchItem* items, const int* first, int n) {
    using T = Tile3<BX, BY, NQ>;
    __shared__ real sbuf[T::BUF];
    __shared__ real sy[T::YS];
    __shared__ real sf[T::FLUX];
    __shared__ double sacc[5 * BX * BY];
    __shared__ real sdma[T::DMA];
    const int b = blockIdx.x;
    const int lo = fvl::owner(first, n, b);
    const fvl::BatchItem& it = fvl::item_at(items, lo);
    const Step& s = it.step;
    const G3 g{it.grid.nx, it.grid.ny, it.grid.nz, it.grid.bc[0], it.grid.bc[1], it.grid.bc[2]};
    const int local = b - first[lo];
    const int tbx = local % it.tiles[0], tby = (local / it.tiles[0]) % it.tiles[1];
    const int tbz = local / (it.tiles[0] * it.tiles[1]);
    if (s.coef && s.coef[9] == 0.0) {
        double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
        block_partials<BX * BY, false, false>(acc, mx, 0.0, s.partials, s.flags, 0, local);
        return;
    }
    step3d_tile<LIM, BX, BY, NQ>(s, g, tbx, tby, tbz, local, sbuf, sy, sf, sacc, sdma);
}

// First-order LLF pass over the whole grid, run only when the second-order
// pass flagged a failure (torchSolver3D.py:161-176, 192-202). The flag test
// happens on the device so the launch can sit in a captured graph.
__device__ __forceinline__ void llf_face(const real* L, const real* R, real gl, real gr, int m, real* F) {
    real pl = (gl - (real)1) * (L[4] - ((real)0.5 * (L[1] * L[1] + L[2] * L[2] + L[3] * L[3])) / L[0]);
    real pr = (gr - (real)1) * (R[4] - ((real)0.5 * (R[1] * R[1] + R[2] * R[2] + R[3] * R[3])) / R[0]);
    real speed = nmax(fabs(L[m] / L[0]) + sqrt((gl * pl) / L[0]), fabs(R[m] / R[0]) + sqrt((gr * pr) / R[0]));
    real fl[5], fr[5];
    flux3(L, gl, m, fl);
    flux3(R, gr, m, fr);
#pragma unroll
    for (int c = 0; c < 5; ++c) F[c] = (real)0.5 * (fl[c] + fr[c] - speed * (R[c] - L[c]));
}

// Cells start, start+stride, ... of one instance; partial slot pblk.
__device__ __forceinline__ void llf3d_cells(const Step& s, const G3& g, long start, long stride, long pblk) {
    double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
    int bad = 0;
    const int need = *s.flags & (fvl::F_PRED_BAD | fvl::F_UPD_BAD);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long N = (long)nx * ny * nz;
    if (!need) return;  // the host reads these partials only after a fallback
    const Coef K = coef_of(s);
    if (!K.active) return;
    for (long i = start; i < N; i += stride) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((long)nx * ny));
        const real* U = (const real*)s.U;
        const real* Ue = s.Ue ? (const real*)s.Ue : U;
        const real* lamp = (const real*)s.lam;
        real Ac[5], o[5], rqv[2];
        load5(U, Ue, N, g, x, y, z, Ac);
        const bool direct = s.lam_is_gamma != 0;
        const real gcen = lamp ? gamma_from(load1(lamp, g, x, y, z), direct) : (real)s.gamma;
        real qcen[2] = {0, 0};
        for (int j = 0; j < 2; ++j) if (j < s.nq) {
            qcen[j] = load1((const real*)s.q[j], g, x, y, z);
            rqv[j] = Ac[0] * qcen[j];
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) o[c] = Ac[c];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const int dx = m == 0, dy = m == 1, dz = m == 2;
            real Lo[5], Hi[5], Fl[5], Fh[5];
            load5(U, Ue, N, g, x - dx, y - dy, z - dz, Lo);
            load5(U, Ue, N, g, x + dx, y + dy, z + dz, Hi);
            real glo = lamp ? gamma_from(load1(lamp, g, x - dx, y - dy, z - dz), direct) : (real)s.gamma;
            real ghi = lamp ? gamma_from(load1(lamp, g, x + dx, y + dy, z + dz), direct) : (real)s.gamma;
            llf_face(Lo, Ac, glo, gcen, m + 1, Fl);
            llf_face(Ac, Hi, gcen, ghi, m + 1, Fh);
            const real cm = (real)K.c[m];
#pragma unroll
            for (int c = 0; c < 5; ++c) o[c] = o[c] - cm * (Fh[c] - Fl[c]);
            for (int j = 0; j < 2; ++j) if (j < s.nq) {
                const real* q = (const real*)s.q[j];
                real qlo = load1(q, g, x - dx, y - dy, z - dz), qhi = load1(q, g, x + dx, y + dy, z + dz);
                real sl = Fl[0] * (Fl[0] >= (real)0 ? qlo : qcen[j]);
                real sh = Fh[0] * (Fh[0] >= (real)0 ? qcen[j] : qhi);
                rqv[j] = rqv[j] - cm * (sh - sl);
            }
            const int pos = m == 0 ? x : (m == 1 ? y : z), n = m == 0 ? nx : (m == 1 ? ny : nz);
            if (pos == 0)
                for (int c = 0; c < 5; ++c) acc[c] -= K.w[m] * (double)Fl[c];
            if (pos == n - 1)
                for (int c = 0; c < 5; ++c) acc[c] += K.w[m] * (double)Fh[c];
            if (s.F[m]) {
                const int ex = dx, ey = dy, ez = dz;
                const long M = (long)(nz + ez) * (ny + ey) * (nx + ex);
                long f = ((long)z * (ny + ey) + y) * (nx + ex) + x;
                long fh = ((long)(z + ez) * (ny + ey) + (y + ey)) * (nx + ex) + (x + ex);
                for (int c = 0; c < 5; ++c) ((real*)s.F[m])[c * M + f] = Fl[c];
                if (pos == n - 1)
                    for (int c = 0; c < 5; ++c) ((real*)s.F[m])[c * M + fh] = Fh[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) ((real*)s.Uout)[c * N + i] = o[c];
        if (!physical5(o)) bad |= fvl::F_LLF_BAD;
        if (s.want_cfl) wave3(o, (double)s.gamma, mx);
        for (int j = 0; j < 2; ++j) if (j < s.nq) {
            real f = rqv[j] / o[0];
            if (f < -kSpecTol || f > (real)1 + kSpecTol || f != f) bad |= fvl::F_SPEC_BAD_LLF;
            ((real*)s.qout[j])[i] = rmin(rmax(f, (real)0), (real)1);
        }
    }
    block_partials<256>(acc, mx, 0.0, s.partials, s.flags, bad, pblk);
}

__global__ void __launch_bounds__(256) k_llf3d(Step s, G3 g) {
    llf3d_cells(s, g, blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x, blockIdx.x);
}

// Work-list form: instance i owns blocks first[i] .. first[i+1]-1.
__global__ void __launch_bounds__(256) k_batch_llf3d(const fvl::BatchItem* items, const int* first, int n) {
    const int b = blockIdx.x;
    const int lo = fvl::owner(first, n, b);
    const fvl::BatchItem& it = fvl::item_at(items, lo);
    const G3 g{it.grid.nx, it.grid.ny, it.grid.nz, it.grid.bc[0], it.grid.bc[1], it.grid.bc[2]};
    const int local = b - first[lo], nblk = first[lo + 1] - first[lo];
    llf3d_cells(it.step, g, local * 256L + threadIdx.x, nblk * 256L, it.llf_first + local);
}

// ---------------------------------------------------------------------------
// Chemistry half-step (torchSolver3D.py:225-243 / torchSolver2D.py:474-509 /
// torchSession.py:365-369 on top of torchSolver.py:708-789). Always double.
// ---------------------------------------------------------------------------
struct Kin {
    int model;
    double q, ea, a;
};

// Division and exp of the kinetics. The FP64 plain method keeps IEEE
// division and the library exp (the reference's arithmetic to a few ulp).
// The FP32 form keeps double arithmetic for the temperature and the stiff
// Arrhenius exponent -Ea/(R T), but divides by a Newton-refined hardware
// reciprocal (r' = r + r (1 - b r), twice, then one residual correction of
// the quotient) and evaluates exp(x) as 2^n 2^f with x log2(e) = n + f split
// in double: 2^f (0 <= f < 1) by the hardware exp2 and 2^n exactly by
// ldexp. Its relative error is about 1e-7, the level of the FP32 storage,
// and does not grow with the exponent's magnitude: enough for the rate
// A exp(-Ea/RT), which enters multiplied. The advance factor exp(-rate t),
// which enters as 1 - exp(-rate t), takes kexp_fine instead.
#ifdef FV_FAST
__device__ __forceinline__ double kdiv(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(e, r, r);
    e = fma(-b, r, 1.0);
    r = fma(e, r, r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ double kexp(double x) {
    if (x != x) return x;
    const double y = fmin(fmax(x * 1.4426950408889634, -1100.0), 1100.0);
    const double n = floor(y);
    return ldexp((double)__builtin_amdgcn_exp2f((float)(y - n)), (int)n);
}
// exp(x) to about an ulp, for the advance factor exp(-rate t): the heat
// release takes 1 - exp(-x), where kexp's ~1e-7 would swamp a weakly reacting
// cell's share (exp(-1e-12) would come out as 1: no heat at all). x = n ln2
// + r with |r| <= ln2/2 (ln2 as a 32-bit head, whose products with n are
// exact, plus a tail), exp(r) by its Taylor series through r^10/10! (the rest
// is below 2.3e-13 relative, and vanishes with r, so a tiny x keeps every
// digit of 1 - exp(x)), scaled by 2^n.
__device__ __forceinline__ double kexp_fine(double x) {
    if (x != x) return x;
    // Weakly reacting cells, most of a grid: |x| < 2^-10 takes the series
    // through x^4/24 directly (the rest is below 1e-17), so a wave of them
    // skips the reduction and the long polynomial.
    if (fabs(x) < 0x1p-10) return fma(fma(fma(fma(x, 1.0 / 24.0, 1.0 / 6.0), x, 0.5), x, 1.0), x, 1.0);
    x = fmin(fmax(x, -746.0), 710.0);
    const double n = rint(x * 1.4426950408889634);
    double r = fma(-n, 0.6931471803691238, x);
    r = fma(-n, 1.9082149292705877e-10, r);
    double p = 2.755731922398589e-07;
    p = fma(p, r, 2.7557319223985893e-06);
    p = fma(p, r, 2.48015873015873e-05);
    p = fma(p, r, 0.0001984126984126984);
    p = fma(p, r, 0.001388888888888889);
    p = fma(p, r, 0.008333333333333333);
    p = fma(p, r, 0.041666666666666664);
    p = fma(p, r, 0.16666666666666666);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return ldexp(p, (int)n);
}
// 1/b refined like kdiv's reciprocal; kquot(a, b, ir) = a / b through it.
__device__ __forceinline__ double krcp(double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(e, r, r);
    e = fma(-b, r, 1.0);
    return fma(e, r, r);
}
__device__ __forceinline__ double kquot(double a, double, double ir) { return a * ir; }
#else
__device__ __forceinline__ double kdiv(double a, double b) { return a / b; }
__device__ __forceinline__ double kexp(double x) { return exp(x); }
__device__ __forceinline__ double kexp_fine(double x) { return exp(x); }
__device__ __forceinline__ double krcp(double b) { return 1.0 / b; }
__device__ __forceinline__ double kquot(double a, double b, double) { return a / b; }
#endif

// Rate constant A exp(-Ea / (8.314 T)) of a cell with density rho (floored),
// velocity u = n1 / rho, reduced energy n2 and fraction lam.
__device__ __forceinline__ double rate_at(const Kin& k, double lam, double rho, double u, double n2) {
#ifdef FV_FAST
    // gamma - 1 = 1 / (3.5 - lam), so with neither floor active (p > 1e-12,
    // T > 1e-12) the exponent -Ea rho R d / (8.314 e M), with e the internal
    // energy, d = 3.5 - lam, M the molar mass and R = 8314.46, takes one
    // division instead of the three of gamma, T and the exponent; M is linear
    // in lam (18.016 - 6.00533.. lam). Cells at a floor take the literal
    // sequence below.
    {
        const double d = 3.5 - lam;
        const double e = n2 - (0.5 * rho) * (u * u);
        const double num = e * fma(lam, (2.0 / 3.0) * 2.016 + (1.0 / 3.0) * 32.0 - 18.016, 18.016);
        const double den = rho * d;
        if (e > 1e-12 * d && num > (1e-12 * 8314.46) * den)
            return k.a * kexp(kdiv(-k.ea * (8314.46 / 8.314) * den, num));
    }
#endif
    double g = kdiv(4.5 - lam, 3.5 - lam);
    double p = fmax((n2 - (0.5 * rho) * (u * u)) * (g - 1.0), 1e-12);
    double rho2 = fmax(rho, 1e-12);
    double molar = lam * (2.0 / 3.0) * 2.016 + lam * (1.0 / 3.0) * 32.0 + (1.0 - lam) * 18.016;
    double t = fmax(kdiv(p * molar, rho2 * 8314.46), 1e-12);
    return k.a * kexp(kdiv(-k.ea, 8.314 * t));
}

__device__ __forceinline__ void advance(const Kin& k, double lam0, double beta0, double n0, double n2, double rate,
                                        double dur, double& lam, double& beta, double& e) {
    if (k.model == 1) {
        double active = fmax(0.0, dur - kdiv(beta0, fmax(rate, 1e-300)));
        lam = lam0 * kexp_fine(-5.0 * rate * active);
        beta = fmin(fmax(beta0 - rate * dur, 0.0), 1.0);
    } else {
        lam = lam0 * kexp_fine(-rate * dur);
        beta = 0.0;
    }
    e = n2 + k.q * n0 * (lam0 - lam);
}

#ifdef FV_FAST
// Change of the sound speed a = sqrt(g (g - 1) e / rho), g = 1 + 1/(3.5 - lam),
// when the chemistry adds de to the internal energy e and moves the fraction
// from lo to ln, free of cancellation: with go, gn the two gammas,
// d(a^2) = [(gn - go)(gn + go - 1) e + gn (gn - 1) de] / rho where
// gn - go = (ln - lo) / ((3.5 - ln)(3.5 - lo)), and da = d(a^2) / (a + a').
// Exactly zero when the cell did not react. ir = 1/rho.
__device__ __forceinline__ double heated_sound_speed_change(double rho, double e, double de, double lo, double ln,
                                                            double ir) {
    const double dlo = 3.5 - lo, dln = 3.5 - ln;
    const double go = 1.0 + krcp(dlo), gn = 1.0 + krcp(dln);
    const double dg = kdiv(ln - lo, dln * dlo);
    const double da2 = ((dg * (gn + go - 1.0)) * e + (gn * (gn - 1.0)) * de) * ir;
    const double a2 = (go * (go - 1.0)) * e * ir;
    const float ao = __builtin_amdgcn_sqrtf((float)a2), an = __builtin_amdgcn_sqrtf((float)fmax(a2 + da2, 0.0));
    return da2 == 0.0 ? 0.0 : (double)((float)da2 / (ao + an));
}
#endif

// Chemistry of cells start, start + stride, ... with half-step duration
// `dt`; partial slot pblk of r.partials. Each thread walks several cells
// (react_blocks), so a block folds its partials once per 256 x kReactCells
// cells; the source sum runs per thread in cell order, then over the block.
// The next cell's inputs are loaded before this cell's kinetics, so their
// latency overlaps the arithmetic; outputs are written from registers.
template <int DIM>
__device__ __forceinline__ void react_cells_dim(const fvl::React& r, const Grid& g, double dt, long start,
                                                long stride, long pblk) {
    constexpr int NF = DIM + 2;
    const long N = (long)g.nx * g.ny * g.nz;
    // No restrict: an energy-only update may write its input buffers in
    // place (each cell's outputs are stored after its inputs were read).
    const real* U = (const real*)r.U;
    const real* L = (const real*)r.lam;
    const real* B = (const real*)r.beta;
    const real* __restrict__ S = (const real*)r.Ustart;
    const real* __restrict__ O = (const real*)r.Uold;
    real* Uo = (real*)r.Uout;
    real* Lo = (real*)r.lamout;
    real* Bo = (real*)r.betaout;
    const bool all_rows = !r.energy_only;
    const Kin k{r.model, r.q_heat, r.activation_energy, r.pre_exponential};
    double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0}, src = 0;
    int bad = 0;
    real nu[NF], nl = 0, nb = 1, ns = 0, no = 0;
    auto fetch = [&](long i) {
#pragma unroll
        for (int c = 0; c < NF; ++c) nu[c] = U[c * N + i];
        nl = L[i];
        if (B) nb = B[i];
        if (S) {
            ns = S[(NF - 1) * N + i];
            no = O[(NF - 1) * N + i];
        }
    };
    if (start < N) fetch(start);
    for (long i = start; i < N; i += stride) {
        real cu[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) cu[c] = nu[c];
        const real cl = nl, cb = nb, cs = ns, co = no;
        if (i + stride < N) fetch(i + stride);
        double u[NF];
#pragma unroll
        for (int c = 0; c < NF; ++c) u[c] = (double)cu[c];
        const double E = u[NF - 1];
        // Every quotient by the density goes through one reciprocal in the
        // FP32 form (a density at or below the floor fails `ok` anyway); the
        // FP64 plain method divides each time, as the reference does.
        const double rho = fmax(u[0], 1e-12), ir = krcp(rho);
        double n2;
        if (DIM == 3) n2 = E - kquot(0.5 * (u[2] * u[2] + u[3] * u[3]), u[0], ir);
        else if (DIM == 2) n2 = E - kquot(0.5 * (u[2] * u[2]), u[0], ir);
        else n2 = E;
        const double lam0 = (double)cl;
        const double beta0 = B ? (double)cb : 1.0;
        bool ok = isfinite(u[0]) && isfinite(u[1]) && isfinite(n2) && u[0] > 1e-12 &&
                  n2 - kquot(0.5 * (u[1] * u[1]), u[0], ir) > 1e-12;
        ok = ok && isfinite(lam0) && lam0 >= 0.0 && lam0 <= 1.0 && isfinite(beta0) && beta0 >= 0.0 && beta0 <= 1.0;
        if (!ok) bad |= fvl::F_REACT_BAD;
        const double vel = kquot(u[1], rho, ir);
        double lm, bm, em, lam, beta, e;
        advance(k, lam0, beta0, u[0], n2, rate_at(k, lam0, rho, vel, n2), 0.5 * dt, lm, bm, em);
        advance(k, lam0, beta0, u[0], n2, rate_at(k, lm, rho, vel, em), dt, lam, beta, e);
        const real En = (real)(E + (e - n2));
        const real ln = (real)lam;
        if (all_rows) {
#pragma unroll
            for (int c = 0; c < NF - 1; ++c) Uo[c * N + i] = cu[c];
        }
        Uo[(NF - 1) * N + i] = En;
        Lo[i] = ln;
        if (Bo) Bo[i] = (real)beta;
        // (old - start + state - hydro) of the energy row (torchSession.py:513-515).
        if (S) src += (((double)co - (double)cs) + (double)En) - E;
        double s2 = 0;
#pragma unroll
        for (int c = 1; c < NF - 1; ++c) s2 = c == 1 ? u[c] * u[c] : s2 + u[c] * u[c];
        if (!(u[0] > 1e-12 && (double)En - kquot(0.5 * s2, u[0], ir) > 1e-12 && isfinite((double)En)))
            bad |= fvl::F_STATE_BAD;
#ifdef FV_FAST
        // Attempts outside 3D Cartesian sessions: the increments go to the
        // CFL pass over the committed state (the same check as the 3D maxima
        // below).
        if (r.increments) {
            if (r.increment_kind == fvl::INC_HEAT) {
                ((real*)r.increments)[i] = (real)(e - n2);
                ((real*)r.increments)[N + i] = (real)(lam - lam0);
            } else {
                ((real*)r.increments)[i] = (real)heated_sound_speed_change(
                    u[0], n2 - kquot(0.5 * (u[1] * u[1]), u[0], ir), e - n2, lam0, lam, ir);
            }
        }
#endif
        if (DIM == 3 && r.want_cfl) {
            // The stored state and fraction, as a standalone CFL pass reads them.
            real so[5];
#pragma unroll
            for (int c = 0; c < 4; ++c) so[c] = cu[c];
            so[4] = En;
            const double ls = (double)ln;
#ifdef FV_FAST
            if (r.attempt) {
                // The attempt is accepted when the heated state's speeds do
                // not exceed the committed state's (which set dt) by more
                // than 1e-12 relative. Stored in FP32, the heated energy
                // loses any heating below half an ulp (~6e-8), which the
                // FP64 method sees and rejects. So report the committed
                // cell's stored-state speeds plus the exact increment of
                // its sound speed: with no chemistry the increment is zero
                // and the maxima are those that set dt.
                so[4] = cu[4];
                real v[3];
                speeds3(so, (real)((4.5 - lam0) / (3.5 - lam0)), v);
                const double da = heated_sound_speed_change(u[0], n2 - kquot(0.5 * (u[1] * u[1]), u[0], ir),
                                                            e - n2, lam0, lam, ir);
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const double w = (double)v[m] + da;
                    mx[m] = (w != w || w > mx[m]) ? w : mx[m];
                }
            } else
#endif
            wave3(so, (4.5 - ls) / (3.5 - ls), mx);
        }
    }
    block_partials<256, false, true>(acc, mx, src, r.partials, r.flags, bad, pblk);
}

__device__ __forceinline__ void react_cells(const fvl::React& r, const Grid& g, double dt, long start, long stride,
                                            long pblk) {
    if (g.dim == 3) react_cells_dim<3>(r, g, dt, start, stride, pblk);
    else if (g.dim == 2) react_cells_dim<2>(r, g, dt, start, stride, pblk);
    else react_cells_dim<1>(r, g, dt, start, stride, pblk);
}

__global__ void __launch_bounds__(256) k_react(fvl::React r, Grid g) {
    react_cells(r, g, r.dt, blockIdx.x * (long)blockDim.x + threadIdx.x, gridDim.x * (long)blockDim.x, blockIdx.x);
}

// Work-list form over reactive instances: phase 1 is an attempt of the first
// half-step (while the chemistry gate is open and dt not yet accepted),
// phase 2 the second half-step (once the hydro gate is open).
__global__ void __launch_bounds__(256) k_batch_react(const fvl::BatchItem* items, const int* first, int n,
                                                     int phase) {
    const int i = fvl::owner(first, n, blockIdx.x);
    const fvl::BatchItem& it = fvl::item_at(items, i);
    const double* k = it.step.coef;
    // Phase 1 runs until this step's dt is accepted; phase 2 once it is.
    if (phase == 1 ? (k[fvl::K_CHEM] == 0.0 || k[fvl::K_HYDRO] != 0.0) : k[fvl::K_HYDRO] == 0.0) return;
    const int local = blockIdx.x - first[i], nblk = first[i + 1] - first[i];
    // Batched reactive instances are 3D.
    react_cells_dim<3>(phase == 1 ? it.r1 : it.r2, it.grid, k[fvl::K_HALF], local * 256L + threadIdx.x, nblk * 256L,
                       (phase == 1 ? it.r1_first : it.r2_first) + local);
}

// CFL maxima of a stored 3D state (torchSession.py:356-360).
__global__ void __launch_bounds__(256) k_cfl3d(const real* U, const real* lam, double gamma, Grid g,
                                               double* partials, int direct) {
    const long N = (long)g.nx * g.ny * g.nz;
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
    if (i < N) {
        real o[5];
        for (int c = 0; c < 5; ++c) o[c] = U[c * N + i];
        double gg = lam ? (direct ? (double)lam[i] : (4.5 - (double)lam[i]) / (3.5 - (double)lam[i])) : gamma;
        wave3(o, gg, mx);
    }
    block_partials<256, false, false>(acc, mx, 0.0, partials, nullptr, 0);
}

// Conserved totals (5 row sums) for drift reporting.
__global__ void __launch_bounds__(256) k_totals(const real* U, Grid g, double* partials, int* flags) {
    const int nf = g.dim + 2;
    const long N = (long)g.nx * g.ny * g.nz;
    double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < N; i += (long)gridDim.x * blockDim.x)
        for (int c = 0; c < nf; ++c) acc[c] += (double)U[c * N + i];
    block_partials<256>(acc, mx, 0.0, partials, flags, 0);
}

// Stand-alone HLLC over face arrays, Cartesian component order
// (hllc_flux_grid, torchFlux.py:6-77): the drop-in for the CPU flux workers.
__global__ void __launch_bounds__(256) k_hllc_faces(int nv, int dir, long m, const real* l, const real* r,
                                                    const real* gl, const real* gr, real* out) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= m) return;
    real a[5], b[5], f[5];
    for (int c = 0; c < nv; ++c) {
        a[c] = l[c * m + i];
        b[c] = r[c * m + i];
    }
    if (nv == 5) hllc_cart5(a, b, gl[i], gr[i], dir, f);
    else hllc_cart4(a, b, gl[i], gr[i], dir, f);
    for (int c = 0; c < nv; ++c) out[c * m + i] = f[c];
}

__global__ void k_convert_in(const double* src, real* dst, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = (real)src[i];
}
__global__ void k_convert_out(const real* src, double* dst, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = (double)src[i];
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static G3 g3(const Grid& g) { return G3{g.nx, g.ny, g.nz, g.bc[0], g.bc[1], g.bc[2]}; }

int step3d_blocks(const Grid& g, const fvl::Launch& l, int zchunk) {
    int tx = l.block_x - 2, ty = l.block_y - 2;
    return ((g.nx + tx - 1) / tx) * ((g.ny + ty - 1) / ty) * ((g.nz + zchunk - 1) / zchunk);
}

template <int BX, int BY, int NQ>
static hipError_t launch_lim(const Grid& g, const Step& s, hipStream_t st) {
    dim3 grid((g.nx + BX - 3) / (BX - 2), (g.ny + BY - 3) / (BY - 2), (g.nz + s.zchunk - 1) / s.zchunk);
    dim3 block(BX, BY);
    G3 q = g3(g);
    switch (s.limiter) {
        case VANLEER: k_step3d<VANLEER, BX, BY, NQ><<<grid, block, 0, st>>>(s, q); break;
        case ALBADA: k_step3d<ALBADA, BX, BY, NQ><<<grid, block, 0, st>>>(s, q); break;
        case MINBEE: k_step3d<MINBEE, BX, BY, NQ><<<grid, block, 0, st>>>(s, q); break;
        case SUPERBEE: k_step3d<SUPERBEE, BX, BY, NQ><<<grid, block, 0, st>>>(s, q); break;
        default: k_step3d<UNOS, BX, BY, NQ><<<grid, block, 0, st>>>(s, q); break;
    }
    return hipGetLastError();
}

template <int NQ>
static hipError_t launch_shape(const Grid& g, const Step& s, const fvl::Launch& l) {
    if (l.block_x == 64 && l.block_y == 4) return launch_lim<64, 4, NQ>(g, s, l.stream);
    if (l.block_x == 64 && l.block_y == 8) return launch_lim<64, 8, NQ>(g, s, l.stream);
    if (l.block_x == 32 && l.block_y == 8) return launch_lim<32, 8, NQ>(g, s, l.stream);
    if (l.block_x == 32 && l.block_y == 16) return launch_lim<32, 16, NQ>(g, s, l.stream);
    if (l.block_x == 32 && l.block_y == 24) {
        // The 768-thread tile exists in the FP32 form only.
        if constexpr (sizeof(real) == sizeof(float)) return launch_lim<32, 24, NQ>(g, s, l.stream);
    }
    return hipErrorInvalidValue;
}

hipError_t step3d(const Grid& g, const Step& s, const fvl::Launch& l) {
    if (s.nq == 2) return launch_shape<2>(g, s, l);
    return s.nq == 1 ? launch_shape<1>(g, s, l) : launch_shape<0>(g, s, l);
}

// Blocks of the first-order pass: it runs only after a failed second-order
// pass and otherwise every block just reads the flag word and exits, which
// in a batch's replay costs time per block; 256 grid-striding blocks per
// instance keep that small.
int llf3d_blocks(const Grid& g) {
    long b = ((long)g.nx * g.ny * g.nz + 255) / 256;
    return (int)(b < 256 ? b : 256);
}

hipError_t llf3d(const Grid& g, const Step& s, hipStream_t st) {
    k_llf3d<<<llf3d_blocks(g), 256, 0, st>>>(s, g3(g));
    return hipGetLastError();
}

template <int BX, int BY, int NQ>
static void launch_batch3d(const fvl::BatchItem* items, const int* first, int n, int blocks, int limiter,
                           hipStream_t st) {
    const dim3 block(BX, BY);
    switch (limiter) {
        case VANLEER: k_batch3d<VANLEER, BX, BY, NQ><<<blocks, block, 0, st>>>(items, first, n); break;
        case ALBADA: k_batch3d<ALBADA, BX, BY, NQ><<<blocks, block, 0, st>>>(items, first, n); break;
        case MINBEE: k_batch3d<MINBEE, BX, BY, NQ><<<blocks, block, 0, st>>>(items, first, n); break;
        case SUPERBEE: k_batch3d<SUPERBEE, BX, BY, NQ><<<blocks, block, 0, st>>>(items, first, n); break;
        default: k_batch3d<UNOS, BX, BY, NQ><<<blocks, block, 0, st>>>(items, first, n); break;
    }
}

template <int NQ>
static hipError_t batch_shape(const fvl::BatchItem* items, const int* first, int n, int blocks, int bx, int by,
                              int limiter, hipStream_t st) {
    if (bx == 64 && by == 4) launch_batch3d<64, 4, NQ>(items, first, n, blocks, limiter, st);
    else if (bx == 64 && by == 8) launch_batch3d<64, 8, NQ>(items, first, n, blocks, limiter, st);
    else if (bx == 32 && by == 8) launch_batch3d<32, 8, NQ>(items, first, n, blocks, limiter, st);
    else if (bx == 32 && by == 16) launch_batch3d<32, 16, NQ>(items, first, n, blocks, limiter, st);
    else if (bx == 32 && by == 24) {
        // The 768-thread tile exists in the FP32 form only.
        if constexpr (sizeof(real) == sizeof(float)) launch_batch3d<32, 24, NQ>(items, first, n, blocks, limiter, st);
        else return hipErrorInvalidValue;
    } else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t batch3d(const fvl::BatchItem* items, const int* first, int n, int blocks, int bx, int by, int limiter,
                   int nq, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    if (nq == 2) return batch_shape<2>(items, first, n, blocks, bx, by, limiter, st);
    return nq == 1 ? batch_shape<1>(items, first, n, blocks, bx, by, limiter, st)
                   : batch_shape<0>(items, first, n, blocks, bx, by, limiter, st);
}

hipError_t batch_llf3d(const fvl::BatchItem* items, const int* first, int n, int blocks, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    k_batch_llf3d<<<blocks, 256, 0, st>>>(items, first, n);
    return hipGetLastError();
}

// CFL maxima of each instance's current state (the batch's first step).
__global__ void __launch_bounds__(256) k_batch_cfl3d(const fvl::BatchItem* items, const int* first, int n) {
    const int b = blockIdx.x;
    const int lo = fvl::owner(first, n, b);
    const fvl::BatchItem& it = fvl::item_at(items, lo);
    const int local = b - first[lo], nblk = first[lo + 1] - first[lo];
    const long N = (long)it.grid.nx * it.grid.ny * it.grid.nz;
    // A reactive instance's committed state is the first half-step's input.
    const real* U = (const real*)(it.reactive ? it.r1.U : it.step.U);
    const real* lam = it.reactive ? (const real*)it.r1.lam : nullptr;
    double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
    for (long i = local * 256L + threadIdx.x; i < N; i += nblk * 256L) {
        real o[5];
        for (int c = 0; c < 5; ++c) o[c] = U[c * N + i];
        const double gg = lam ? (4.5 - (double)lam[i]) / (3.5 - (double)lam[i]) : it.step.gamma;
        wave3(o, gg, mx);
    }
    block_partials<256, false, false>(acc, mx, 0.0, it.step.partials, nullptr, 0, it.cfl_first + local);
}

hipError_t batch_cfl3d(const fvl::BatchItem* items, const int* first, int n, int blocks, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    k_batch_cfl3d<<<blocks, 256, 0, st>>>(items, first, n);
    return hipGetLastError();
}

// Cells per thread of the chemistry passes.
constexpr long kReactCells = 8;
int react_blocks(const Grid& g) {
    return (int)(((long)g.nx * g.ny * g.nz + 256 * kReactCells - 1) / (256 * kReactCells));
}

hipError_t react(const Grid& g, const fvl::React& r, hipStream_t st) {
    k_react<<<react_blocks(g), 256, 0, st>>>(r, g);
    return hipGetLastError();
}

hipError_t batch_react(const fvl::BatchItem* items, const int* first, int n, int blocks, int phase,
                       hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    k_batch_react<<<blocks, 256, 0, st>>>(items, first, n, phase);
    return hipGetLastError();
}

int cfl_blocks(const Grid& g) {
    if (g.dim == 2) return cfl2d_blocks(g);
    if (g.dim == 1) return cfl1d_blocks(g);
    return (int)(((long)g.nx * g.ny * g.nz + 255) / 256);
}

// `increments` (FP32 form, 2D and 1D, else null): the chemistry's per-cell
// increments of a reactive attempt; the maxima are then those of U (the
// committed state) raised by the exact change the heating makes to them.
hipError_t cfl_cells(const Grid& g, const void* U, const void* lam, double gamma, double* partials, int direct,
                     const void* increments, hipStream_t st) {
#ifndef FV_FAST
    if (increments) return hipErrorInvalidValue;
#endif
    if (g.dim == 2) return cfl2d(g, U, lam, gamma, partials, direct, increments, st);
    if (g.dim == 3 && increments) return hipErrorInvalidValue;
    if (g.dim == 1) {
        Step s;
        memset(&s, 0, sizeof(s));
        s.U = U;
        s.lam = lam;
        s.lam_is_gamma = direct;
        s.gamma = gamma;
        s.wavespeed = g.wavespeed;
        s.partials = partials;
        s.heat = increments;
        return cfl1d(g, s, st);
    }
    k_cfl3d<<<cfl_blocks(g), 256, 0, st>>>((const real*)U, (const real*)lam, gamma, g, partials, direct);
    return hipGetLastError();
}

hipError_t hllc_faces(int nv, int dir, long m, const void* l, const void* r, const void* gl, const void* gr,
                      void* out, hipStream_t st) {
    if (m <= 0) return hipSuccess;
    k_hllc_faces<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(nv, dir, m, (const real*)l, (const real*)r,
                                                              (const real*)gl, (const real*)gr, (real*)out);
    return hipGetLastError();
}

hipError_t totals(const Grid& g, const void* U, double* partials, hipStream_t st) {
    k_totals<<<256, 256, 0, st>>>((const real*)U, g, partials, nullptr);
    return hipGetLastError();
}

hipError_t convert_in(const double* src, void* dst, long n, hipStream_t st) {
    k_convert_in<<<1024, 256, 0, st>>>(src, (real*)dst, n);
    return hipGetLastError();
}
hipError_t convert_out(const void* src, double* dst, long n, hipStream_t st) {
    k_convert_out<<<1024, 256, 0, st>>>((const real*)src, dst, n);
    return hipGetLastError();
}

}  // namespace FV_NS
