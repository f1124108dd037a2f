// 2D Cartesian MUSCL-Hancock + HLLC step (python/cfdSolver/torchSolver2D.py:
// 274-466), its first-order LLF fallback (:419-462), the species transport
// of advect_scalar_2d (:659-667) and the padded-state CFL of
// compute_cfl_dt_2d (:512-530, torchSession.py:346-355).
//
// The 2D scheme is not the 3D one restricted: predictor fluxes floor rho and
// p, the predicted state is repaired rather than triggering a fallback,
// VanLeer/Albada are gated on r > 0, Unos lacks the (1+2r)/3 bound, and the
// y axis accepts an Inflow boundary (fixed inlet state below, open above).
//
// One wave owns 62 columns plus a halo lane on each side and marches up a
// chunk of rows: x neighbours by lane shuffles, y neighbours in registers.
#include "fv_device.h"
#include "fv_launch.h"

namespace FV_NS {
using namespace fv;
using fvl::Grid;
using fvl::Step;

struct G2 {
    int nx, ny, bcx, bcy;
    real inlet[4];
};

__device__ __forceinline__ void load4(const real* __restrict__ U, long N, const G2& g, int x, int y, real* a) {
    if (y < 0 && g.bcy == BC_INFLOW) {
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] = g.inlet[c];
        return;
    }
    long i = (long)ghost_src(y, g.ny, g.bcy) * g.nx + ghost_src(x, g.nx, g.bcx);
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = U[c * N + i];
    if (ghost_flip(x, g.nx, g.bcx)) a[1] = -a[1];
    if (ghost_flip(y, g.ny, g.bcy)) a[2] = -a[2];
}

// Ghost-mapped scalar; below an Inflow boundary it takes `inlet`.
__device__ __forceinline__ real load1s(const real* __restrict__ q, const G2& g, int x, int y, real inlet) {
    if (y < 0 && g.bcy == BC_INFLOW) return inlet;
    return q[(long)ghost_src(y, g.ny, g.bcy) * g.nx + ghost_src(x, g.nx, g.bcx)];
}

__device__ __forceinline__ real gamma2(const Step& s, const G2& g, int x, int y) {
    if (y < 0 && g.bcy == BC_INFLOW) return (real)1.4;
    return s.lam ? gamma_from(load1s((const real*)s.lam, g, x, y, (real)1), s.lam_is_gamma != 0) : (real)s.gamma;
}

// Floored predictor flux (torchSolver2D.py:326-352); dir 0 x, 1 y.
__device__ __forceinline__ void flux2f(const real* u, real g, int dir, real* f) {
    real rho = rmax(u[0], (real)1e-12);
    real vx = u[1] / rho, vy = u[2] / rho;
    real p = rmax((g - (real)1) * (u[3] - ((real)0.5 * rho) * (vx * vx + vy * vy)), (real)1e-12);
    if (dir == 0) {
        f[0] = rho * vx;
        f[1] = rho * (vx * vx) + p;
        f[2] = (rho * vx) * vy;
        f[3] = vx * (u[3] + p);
    } else {
        f[0] = rho * vy;
        f[1] = (rho * vx) * vy;
        f[2] = rho * (vy * vy) + p;
        f[3] = vy * (u[3] + p);
    }
}

template <int NT>
__device__ void partials2(double (&acc)[5], double (&mx)[3], double* partials, int* flags, int bad,
                          long block = -1) {
    __shared__ double red[NT / 64][8];
    __shared__ int sbad;
    const int tid = threadIdx.x;
    if (tid == 0) sbad = 0;
    __syncthreads();
    if (bad) atomicOr(&sbad, bad);
    double v[8];
    for (int i = 0; i < 3; ++i) v[i] = wave_max(mx[i]);
    for (int i = 0; i < 5; ++i) v[3 + i] = wave_sum(acc[i]);
    if ((tid & 63) == 0)
        for (int i = 0; i < 8; ++i) red[tid / 64][i] = v[i];
    __syncthreads();
    if (tid == 0) {
        double* p = partials + (block >= 0 ? block : blockIdx.x + (long)gridDim.x * blockIdx.y) * fvl::NSLOT;
        for (int i = 0; i < 8; ++i) {
            double r = red[0][i];
            for (int k = 1; k < NT / 64; ++k) r = i < 3 ? fmax(r, red[k][i]) : r + red[k][i];
            p[i < 3 ? i : fvl::S_OUT + (i - 3)] = r;
        }
        p[fvl::S_SRC] = 0.0;
        if (sbad) atomicOr(flags, sbad);
    }
}

// |u|+a and |v|+a of a 2D state with the reference's floors (the CFL rule of
// compute_cfl_dt_2d, torchSolver2D.py:512-530), accumulated as maxima.
__device__ __forceinline__ void wave2(const real* u, real g, double (&mx)[3]);

// One 62-column strip (tbx) of one row chunk (tby); partial slot pblk.
template <int LIM, bool SPEC>
__device__ __forceinline__ void step2d_strip(const Step& s, const G2& g, int ychunk, int tbx, int tby, long pblk) {
    constexpr int TX = 62;
    const int lane = threadIdx.x, lx = lane - 1;
    const int x = tbx * TX + lx;
    const int nx = g.nx, ny = g.ny;
    const long N = (long)nx * ny;
    const int y0 = tby * ychunk, y1 = min(y0 + ychunk, ny);
    const real* __restrict__ U = (const real*)s.U;
    real* __restrict__ Uo = (real*)s.Uout;
    const int nq = SPEC ? s.nq : 0;
    const bool owns = lx >= 0 && lx < TX && x < nx;
    const bool xface = lx >= 0 && x <= nx;
    const fvl::Coef K = fvl::coef_of(s);
    const real c0 = (real)K.c[0], c1 = (real)K.c[1], h0 = (real)K.hc[0], h1 = (real)K.hc[1];

    real Ap[4], Ac[4], An[4], gc, gn;
    real qc[2] = {0, 0}, qp[2] = {0, 0}, qn[2] = {0, 0};
    real yt[4] = {0, 0, 0, 0}, Fyp[4] = {0, 0, 0, 0}, sfyp[2] = {0, 0}, R[4] = {0, 0, 0, 0}, rq[2] = {0, 0};
    double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
    int bad = 0;
    load4(U, N, g, x, y0 - 2, Ap);
    load4(U, N, g, x, y0 - 1, Ac);
    gc = gamma2(s, g, x, y0 - 1);
    for (int j = 0; j < nq; ++j) {
        qp[j] = load1s((const real*)s.q[j], g, x, y0 - 2, (real)1);
        qc[j] = load1s((const real*)s.q[j], g, x, y0 - 1, (real)1);
    }
    real gp = gc;
    for (int k = y0 - 1; k <= y1; ++k) {
        load4(U, N, g, x, k + 1, An);
        gn = gamma2(s, g, x, k + 1);
        for (int j = 0; j < nq; ++j) qn[j] = load1s((const real*)s.q[j], g, x, k + 1, (real)1);
        // x neighbours of row k (halo lanes fetch the missing one).
        real Al[4], Ar[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            Al[c] = shfl_up1(Ac[c]);
            Ar[c] = shfl_down1(Ac[c]);
        }
        if (lane == 0) load4(U, N, g, x - 1, k, Al);
        if (lane == 63) load4(U, N, g, x + 1, k, Ar);
        real dxs[4], dys[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            dxs[c] = limited<LIM, 2>(Ac[c] - Al[c], Ar[c] - Ac[c]);
            dys[c] = limited<LIM, 2>(Ac[c] - Ap[c], An[c] - Ac[c]);
        }
        // Zero slopes in a direction give equal half states and an exact
        // zero change: its two flux evaluations are skipped.
        real ub[4], a[4], b[4], fa[4], fb[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) ub[c] = Ac[c];
        if (!((dxs[0] == (real)0) & (dxs[1] == (real)0) & (dxs[2] == (real)0) & (dxs[3] == (real)0))) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                a[c] = Ac[c] + (real)0.5 * dxs[c];
                b[c] = Ac[c] - (real)0.5 * dxs[c];
            }
            flux2f(a, gc, 0, fa);
            flux2f(b, gc, 0, fb);
#pragma unroll
            for (int c = 0; c < 4; ++c) ub[c] = Ac[c] - h0 * (fa[c] - fb[c]);
        }
        if (!((dys[0] == (real)0) & (dys[1] == (real)0) & (dys[2] == (real)0) & (dys[3] == (real)0))) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                a[c] = Ac[c] + (real)0.5 * dys[c];
                b[c] = Ac[c] - (real)0.5 * dys[c];
            }
            flux2f(a, gc, 1, fa);
            flux2f(b, gc, 1, fb);
#pragma unroll
            for (int c = 0; c < 4; ++c) ub[c] = ub[c] - h1 * (fa[c] - fb[c]);
        }
        // Positivity repair (torchSolver2D.py:367-377).
        ub[0] = rmax(ub[0], (real)1e-12);
        real kin = ((real)0.5 * (ub[1] * ub[1] + ub[2] * ub[2])) / ub[0];
        real pbar = (gc - (real)1) * (ub[3] - kin);
        if (pbar < (real)1e-12) ub[3] = (real)1e-12 / (gc - (real)1) + kin;
        real xl[4], xr[4], yl[4], ytn[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            real hx = (real)0.5 * dxs[c], hy = (real)0.5 * dys[c];
            xl[c] = ub[c] - hx; xr[c] = ub[c] + hx;
            yl[c] = ub[c] - hy; ytn[c] = ub[c] + hy;
        }
        const bool row_out = k >= y0 && k < y1;
        real lxr[4], Fx[4] = {0, 0, 0, 0}, sfx[2] = {0, 0};
#pragma unroll
        for (int c = 0; c < 4; ++c) lxr[c] = shfl_up1(xr[c]);
        real lg = shfl_up1(gc), lq[2] = {0, 0};
        for (int j = 0; j < (SPEC ? 2 : 0); ++j) lq[j] = shfl_up1(qc[j]);
        if (row_out && xface) {
            hllc_cart4(lxr, xl, lg, gc, 0, Fx);
            for (int j = 0; j < nq; ++j) sfx[j] = Fx[0] * (Fx[0] >= (real)0 ? lq[j] : qc[j]);
        }
        real Fy[4] = {0, 0, 0, 0}, sfy[2] = {0, 0};
        const bool yface = owns && k >= y0;
        if (yface) {
            hllc_cart4(yt, yl, gp, gc, 1, Fy);
            for (int j = 0; j < nq; ++j) sfy[j] = Fy[0] * (Fy[0] >= (real)0 ? qp[j] : qc[j]);
        }
        real FxR[4], sfxR[2] = {0, 0};
#pragma unroll
        for (int c = 0; c < 4; ++c) FxR[c] = shfl_down1(Fx[c]);
        for (int j = 0; j < (SPEC ? 2 : 0); ++j) sfxR[j] = shfl_down1(sfx[j]);
        if (owns && k - 1 >= y0) {
            real o[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) o[c] = R[c] - c1 * (Fy[c] - Fyp[c]);
            const long i = (long)(k - 1) * nx + x;
#pragma unroll
            for (int c = 0; c < 4; ++c) Uo[c * N + i] = o[c];
            if (!physical4(o)) bad |= fvl::F_UPD_BAD;
            if (s.want_cfl) wave2(o, gp, mx);  // gp: gamma of row k-1
            for (int j = 0; j < nq; ++j) {
                real v = rq[j] - c1 * (sfy[j] - sfyp[j]);
                ((real*)s.qout[j])[i] = rmin(rmax(v / rmax(o[0], (real)1e-12), (real)0), (real)1);
            }
        }
        if (owns && row_out) {
#pragma unroll
            for (int c = 0; c < 4; ++c) R[c] = Ac[c] - c0 * (FxR[c] - Fx[c]);
            for (int j = 0; j < nq; ++j) rq[j] = Ac[0] * qc[j] - c0 * (sfxR[j] - sfx[j]);
        }
        if (row_out && xface) {
            if (x == 0 || x == nx) {
                double sg = x == 0 ? -K.w[0] : K.w[0];
                for (int c = 0; c < 4; ++c) acc[c] += sg * (double)Fx[c];
            }
            if (s.F[0]) {
                long f = (long)k * (nx + 1) + x, M = (long)ny * (nx + 1);
                for (int c = 0; c < 4; ++c) ((real*)s.F[0])[c * M + f] = Fx[c];
            }
        }
        if (yface) {
            if (k == 0 || k == ny) {
                double sg = k == 0 ? -K.w[1] : K.w[1];
                for (int c = 0; c < 4; ++c) acc[c] += sg * (double)Fy[c];
            }
            if (s.F[1]) {
                long f = (long)k * nx + x, M = (long)(ny + 1) * nx;
                for (int c = 0; c < 4; ++c) ((real*)s.F[1])[c * M + f] = Fy[c];
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            Ap[c] = Ac[c];
            Ac[c] = An[c];
            yt[c] = ytn[c];
            Fyp[c] = Fy[c];
        }
        gp = gc;
        gc = gn;
        for (int j = 0; j < nq; ++j) {
            qp[j] = qc[j];
            qc[j] = qn[j];
            sfyp[j] = sfy[j];
        }
    }
    partials2<64>(acc, mx, s.partials, s.flags, bad, pblk);
}

template <int LIM, bool SPEC>
__global__ void __launch_bounds__(64) k_step2d(Step s, G2 g, int ychunk) {
    step2d_strip<LIM, SPEC>(s, g, ychunk, blockIdx.x, blockIdx.y, blockIdx.x + (long)gridDim.x * blockIdx.y);
}

__device__ __forceinline__ G2 g2_of(const Grid& g) {
    G2 q;
    q.nx = g.nx;
    q.ny = g.ny;
    q.bcx = g.bc[0];
    q.bcy = g.bc[1];
    for (int c = 0; c < 4; ++c) q.inlet[c] = (real)g.inlet[c];
    return q;
}

// Work-list form over 2D instances (strips x row chunks per instance), one
// limiter per launch.
template <int LIM>
__global__ void __launch_bounds__(64) k_batch2d(const fvl::BatchItem* items, const int* first, int n) {
    const int i = fvl::owner(first, n, blockIdx.x);
    const fvl::BatchItem& it = fvl::item_at(items, i);
    const int local = blockIdx.x - first[i];
    const G2 g = g2_of(it.grid);
    const Step& s = it.step;
    if (s.coef && s.coef[9] == 0.0) {
        double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
        partials2<64>(acc, mx, s.partials, s.flags, 0, local);
        return;
    }
    step2d_strip<LIM, false>(s, g, it.tiles[2], local % it.tiles[0], local / it.tiles[0], local);
}

__device__ __forceinline__ void wave2(const real* u, real g, double (&mx)[3]) {
    const Prim2 p = prim2(u, g);
    const double vx = (double)(fabs(p.ux) + p.a), vy = (double)(fabs(p.vy) + p.a);
    mx[0] = (vx != vx || vx > mx[0]) ? vx : mx[0];
    mx[1] = (vy != vy || vy > mx[1]) ? vy : mx[1];
}

__device__ __forceinline__ void llf2_face(const real* L, const real* R, real gl, real gr, int dir, real* F) {
    Prim2 pl = prim2(L, gl), pr = prim2(R, gr);
    real fl[4], fr[4];
    real vl = dir == 0 ? pl.ux : pl.vy, vr = dir == 0 ? pr.ux : pr.vy;
    fl[0] = L[1 + dir];
    fl[1] = L[1] * vl + (dir == 0 ? pl.p : (real)0);
    fl[2] = L[2] * vl + (dir == 1 ? pl.p : (real)0);
    fl[3] = vl * (L[3] + pl.p);
    fr[0] = R[1 + dir];
    fr[1] = R[1] * vr + (dir == 0 ? pr.p : (real)0);
    fr[2] = R[2] * vr + (dir == 1 ? pr.p : (real)0);
    fr[3] = vr * (R[3] + pr.p);
    real sp = nmax(fabs(vl) + pl.a, fabs(vr) + pr.a);
#pragma unroll
    for (int c = 0; c < 4; ++c) F[c] = (real)0.5 * (fl[c] + fr[c] - sp * (R[c] - L[c]));
}

// First-order pass over cells start, start+stride, ...; partial slot pblk.
__device__ __forceinline__ void llf2d_cells(const Step& s, const G2& g, long start, long stride, long pblk) {
    double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
    int bad = 0;
    const bool need = (*s.flags & fvl::F_UPD_BAD) != 0;
    if (!need) return;  // partials of the second-order pass stay in use
    const int nx = g.nx, ny = g.ny;
    const long N = (long)nx * ny;
    const fvl::Coef K = fvl::coef_of(s);
    if (!K.active) return;
    for (long i = start; i < N; i += stride) {
        const int x = (int)(i % nx), y = (int)(i / nx);
        const real* U = (const real*)s.U;
        real Ac[4], o[4], rqv[2], qc[2] = {0, 0};
        load4(U, N, g, x, y, Ac);
        const real gcen = gamma2(s, g, x, y);
        for (int j = 0; j < s.nq; ++j) {
            qc[j] = load1s((const real*)s.q[j], g, x, y, (real)1);
            rqv[j] = Ac[0] * qc[j];
        }
        for (int c = 0; c < 4; ++c) o[c] = Ac[c];
        for (int m = 0; m < 2; ++m) {
            const int dx = m == 0, dy = m == 1;
            real Lo[4], Hi[4], Fl[4], Fh[4];
            load4(U, N, g, x - dx, y - dy, Lo);
            load4(U, N, g, x + dx, y + dy, Hi);
            llf2_face(Lo, Ac, gamma2(s, g, x - dx, y - dy), gcen, m, Fl);
            llf2_face(Ac, Hi, gcen, gamma2(s, g, x + dx, y + dy), m, Fh);
            const real cm = (real)K.c[m];
            for (int c = 0; c < 4; ++c) o[c] = o[c] - cm * (Fh[c] - Fl[c]);
            for (int j = 0; j < s.nq; ++j) {
                const real* q = (const real*)s.q[j];
                real qlo = load1s(q, g, x - dx, y - dy, (real)1), qhi = load1s(q, g, x + dx, y + dy, (real)1);
                real sl = Fl[0] * (Fl[0] >= (real)0 ? qlo : qc[j]);
                real sh = Fh[0] * (Fh[0] >= (real)0 ? qc[j] : qhi);
                rqv[j] = rqv[j] - cm * (sh - sl);
            }
            const int pos = m == 0 ? x : y, n = m == 0 ? nx : ny;
            if (pos == 0)
                for (int c = 0; c < 4; ++c) acc[c] -= K.w[m] * (double)Fl[c];
            if (pos == n - 1)
                for (int c = 0; c < 4; ++c) acc[c] += K.w[m] * (double)Fh[c];
            if (s.F[m]) {
                const long M = (long)(ny + dy) * (nx + dx);
                long f = (long)y * (nx + dx) + x, fh = (long)(y + dy) * (nx + dx) + (x + dx);
                for (int c = 0; c < 4; ++c) ((real*)s.F[m])[c * M + f] = Fl[c];
                if (pos == n - 1)
                    for (int c = 0; c < 4; ++c) ((real*)s.F[m])[c * M + fh] = Fh[c];
            }
        }
        for (int c = 0; c < 4; ++c) ((real*)s.Uout)[c * N + i] = o[c];
        if (!physical4(o)) bad |= fvl::F_LLF_BAD;
        if (s.want_cfl) wave2(o, gcen, mx);
        for (int j = 0; j < s.nq; ++j)
            ((real*)s.qout[j])[i] = rmin(rmax(rqv[j] / rmax(o[0], (real)1e-12), (real)0), (real)1);
    }
    partials2<256>(acc, mx, s.partials, s.flags, bad, pblk);
}

__global__ void __launch_bounds__(256) k_llf2d(Step s, G2 g) {
    llf2d_cells(s, g, blockIdx.x * (long)blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_batch_llf2d(const fvl::BatchItem* items, const int* first, int n) {
    const int i = fvl::owner(first, n, blockIdx.x);
    const fvl::BatchItem& it = fvl::item_at(items, i);
    const int local = blockIdx.x - first[i], nblk = first[i + 1] - first[i];
    llf2d_cells(it.step, g2_of(it.grid), local * 256L + threadIdx.x, nblk * 256L, it.llf_first + local);
}

// CFL maxima of each 2D instance's state (interior; the inlet ghost state of
// an Inflow boundary is added by the batch post kernel).
__global__ void __launch_bounds__(256) k_batch_cfl2d(const fvl::BatchItem* items, const int* first, int n) {
    const int i = fvl::owner(first, n, blockIdx.x);
    const fvl::BatchItem& it = fvl::item_at(items, i);
    const int local = blockIdx.x - first[i], nblk = first[i + 1] - first[i];
    const long N = (long)it.grid.nx * it.grid.ny;
    const real* U = (const real*)it.step.U;
    double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
    for (long c = local * 256L + threadIdx.x; c < N; c += nblk * 256L) {
        real u[4];
        for (int k = 0; k < 4; ++k) u[k] = U[k * N + c];
        wave2(u, (real)it.step.gamma, mx);
    }
    partials2<256>(acc, mx, it.step.partials, nullptr, 0, it.cfl_first + local);
}

// CFL maxima over the padded 2D state: interior cells and, with an Inflow
// boundary, the inlet ghost state at gamma 1.4.
// With `dspeed`, each cell's speeds plus its sound-speed increment (the
// inflow state has none): the FP32 form's check of a reactive attempt, over
// the committed state whose own maxima set dt.
__global__ void __launch_bounds__(256) k_cfl2d(const real* U, const real* lam, double gamma, G2 g,
                                               double* partials, int direct, const real* dspeed) {
    const long N = (long)g.nx * g.ny;
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    double acc[5] = {0, 0, 0, 0, 0}, mx[3] = {0, 0, 0};
    if (i < N || (i == N && g.bcy == BC_INFLOW)) {
        real u[4];
        real gg;
        if (i < N) {
            for (int c = 0; c < 4; ++c) u[c] = U[c * N + i];
            gg = lam ? gamma_from(lam[i], direct != 0) : (real)gamma;
        } else {
            for (int c = 0; c < 4; ++c) u[c] = g.inlet[c];
            gg = (real)1.4;
        }
        Prim2 p = prim2(u, gg);
        mx[0] = (double)(fabs(p.ux) + p.a);
        mx[1] = (double)(fabs(p.vy) + p.a);
        if (dspeed && i < N) {
            mx[0] += (double)dspeed[i];
            mx[1] += (double)dspeed[i];
        }
    }
    partials2<256>(acc, mx, partials, nullptr, 0);
}

static G2 g2(const Grid& g) {
    G2 q;
    q.nx = g.nx;
    q.ny = g.ny;
    q.bcx = g.bc[0];
    q.bcy = g.bc[1];
    for (int c = 0; c < 4; ++c) q.inlet[c] = (real)g.inlet[c];
    return q;
}

static int ychunk_of(const Grid& g) {
    int tiles = (g.nx + 61) / 62;
    int chunk = 8;
    while (chunk < 64 && (long)tiles * ((g.ny + chunk - 1) / chunk) > 8192) chunk *= 2;
    return chunk;
}

int step2d_blocks(const Grid& g, const fvl::Launch&) {
    int ch = ychunk_of(g);
    return ((g.nx + 61) / 62) * ((g.ny + ch - 1) / ch);
}

int strip2d_rows(const Grid& g) { return ychunk_of(g); }

hipError_t batch2d(const fvl::BatchItem* items, const int* first, int n, int blocks, int limiter, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    switch (limiter) {
        case VANLEER: k_batch2d<VANLEER><<<blocks, 64, 0, st>>>(items, first, n); break;
        case ALBADA: k_batch2d<ALBADA><<<blocks, 64, 0, st>>>(items, first, n); break;
        case MINBEE: k_batch2d<MINBEE><<<blocks, 64, 0, st>>>(items, first, n); break;
        case SUPERBEE: k_batch2d<SUPERBEE><<<blocks, 64, 0, st>>>(items, first, n); break;
        default: k_batch2d<UNOS><<<blocks, 64, 0, st>>>(items, first, n); break;
    }
    return hipGetLastError();
}

hipError_t batch_llf2d(const fvl::BatchItem* items, const int* first, int n, int blocks, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    k_batch_llf2d<<<blocks, 256, 0, st>>>(items, first, n);
    return hipGetLastError();
}

hipError_t batch_cfl2d(const fvl::BatchItem* items, const int* first, int n, int blocks, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    k_batch_cfl2d<<<blocks, 256, 0, st>>>(items, first, n);
    return hipGetLastError();
}

template <bool SPEC>
static void launch2(const Grid& g, const Step& s, hipStream_t st) {
    int ch = ychunk_of(g);
    dim3 grid((g.nx + 61) / 62, (g.ny + ch - 1) / ch);
    G2 q = g2(g);
    switch (s.limiter) {
        case VANLEER: k_step2d<VANLEER, SPEC><<<grid, 64, 0, st>>>(s, q, ch); break;
        case ALBADA: k_step2d<ALBADA, SPEC><<<grid, 64, 0, st>>>(s, q, ch); break;
        case MINBEE: k_step2d<MINBEE, SPEC><<<grid, 64, 0, st>>>(s, q, ch); break;
        case SUPERBEE: k_step2d<SUPERBEE, SPEC><<<grid, 64, 0, st>>>(s, q, ch); break;
        default: k_step2d<UNOS, SPEC><<<grid, 64, 0, st>>>(s, q, ch); break;
    }
}

hipError_t step2d(const Grid& g, const Step& s, const fvl::Launch& l) {
    if (s.nq > 0) launch2<true>(g, s, l.stream);
    else launch2<false>(g, s, l.stream);
    return hipGetLastError();
}

// The first-order pass checks the update flag on the device, so it can be
// queued unconditionally; its partials are read only when it ran.
// At most 256 grid-striding blocks: the pass is idle unless a second-order
// update failed (see llf3d_blocks).
int llf2d_blocks(const Grid& g) {
    long b = ((long)g.nx * g.ny + 255) / 256;
    return (int)(b < 256 ? b : 256);
}

hipError_t llf2d(const Grid& g, const Step& s, hipStream_t st) {
    k_llf2d<<<llf2d_blocks(g), 256, 0, st>>>(s, g2(g));
    return hipGetLastError();
}

hipError_t cfl2d(const Grid& g, const void* U, const void* lam, double gamma, double* partials, int direct,
                 const void* dspeed, hipStream_t st) {
    k_cfl2d<<<(unsigned)(((long)g.nx * g.ny + 256) / 256), 256, 0, st>>>(
        (const real*)U, (const real*)lam, gamma, g2(g), partials, direct, (const real*)dspeed);
    return hipGetLastError();
}

int cfl2d_blocks(const Grid& g) { return (int)(((long)g.nx * g.ny + 256) / 256); }

}  // namespace FV_NS
