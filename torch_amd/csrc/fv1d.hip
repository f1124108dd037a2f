// Legacy 1D MUSCL-Hancock scheme of the session's dimension-1 path
// (python/cfdSolver/torchSession.py:413-445): PVRS or enthalpy wave speeds
// from cell averages (torchSolver.py:283-404), sound-speed-controlled
// Sweby-bounded limiting and half-step sub-cell states (:201-275, 412-508),
// "Variante 1"/"Variante_2" HLLC on the predicted interfaces (:534-676), the
// conservative update with its LLF fallback (:1054-1071), upwind species
// (:1125-1132) and the two-bound CFL (:684-689, 1049-1051).
//
// States are (3, n) [rho, rho*u, E] with one ghost cell per side. The 1D
// problems are small (n <= 8192 in the workloads), so each stage is its own
// elementwise kernel over a scratch buffer rather than one fused pass.
#include "fv_device.h"
#include "fv_launch.h"

namespace FV_NS {
using namespace fv;
using fvl::Grid;
using fvl::Step;

struct P1 {
    real rho, v, p;
};

__device__ __forceinline__ P1 prim1(const real* u, real g) {
    P1 q;
    q.rho = rmax(u[0], (real)1e-12);
    q.v = u[1] / q.rho;
    q.p = rmax((u[2] - ((real)0.5 * q.rho) * (q.v * q.v)) * (g - (real)1), (real)1e-12);
    return q;
}

__device__ __forceinline__ real sound1(const P1& q, real g) {
    return sqrt((g * rmax(q.p, (real)0)) / rmax(q.rho, (real)1e-12));
}

__device__ __forceinline__ void eflux1(const real* u, real g, real* f) {
    P1 q = prim1(u, g);
    f[0] = q.rho * q.v;
    f[1] = q.rho * (q.v * q.v) + q.p;
    f[2] = q.v * (u[2] + q.p);
}

// Padded cell j in [0, n+1] (physical j-1) with the 1D ghosts of bc_creator.
__device__ __forceinline__ void cell1(const real* U, int n, int bc, int j, real* a) {
    const int i = j - 1;
    const int s = ghost_src(i, n, bc);
#pragma unroll
    for (int c = 0; c < 3; ++c) a[c] = U[c * (long)n + s];
    if (ghost_flip(i, n, bc)) a[1] = -a[1];
}

__device__ __forceinline__ real gamma1(const Step& s, int n, int bc, int j) {
    if (!s.lam) return (real)s.gamma;
    return gamma_from(((const real*)s.lam)[ghost_src(j - 1, n, bc)], s.lam_is_gamma != 0);
}

__device__ __forceinline__ real zero_bad(real x) { return isfinite(x) ? x : (real)0; }

// Wave speeds S_L, S*, S_R at face f (between padded cells f and f+1).
__device__ void waves1(const real* L, const real* R, real gl, real gr, int mode, real* S) {
    P1 a = prim1(L, gl), b = prim1(R, gr);
    real rl = rmax(a.rho, (real)1e-12), rr = rmax(b.rho, (real)1e-12);
    real al = sound1(a, gl), ar = sound1(b, gr);
    real sl, sr;
    if (mode == 0) {
        real ps = nmax((real)0, (real)0.5 * (a.p + b.p) -
                                    (real)0.5 * (b.v - a.v) * ((real)0.5 * (b.rho + a.rho)) * ((real)0.5 * (al + ar)));
        real ql = (real)1, qr = (real)1;
        if (ps > a.p) ql = sqrt((real)1 + (gl + (real)1) / ((real)2 * gl) * (ps / a.p - (real)1));
        if (ps > b.p) qr = sqrt((real)1 + (gr + (real)1) / ((real)2 * gr) * (ps / b.p - (real)1));
        sl = a.v - al * ql;
        sr = b.v + ar * qr;
    } else {
        real elt = rl * ((real)0.5 * (a.v * a.v) + a.p / ((gl - (real)1) * rl));
        real ert = rr * ((real)0.5 * (b.v * b.v) + b.p / ((gr - (real)1) * rr));
        real sql = sqrt(rl), sqr = sqrt(rr);
        real den = rmax(sql + sqr, (real)1e-12);
        real ua = (sql * a.v + sqr * b.v) / den;
        real ha = (sql * ((elt + a.p) / rl) + sqr * ((ert + b.p) / rr)) / den;
        real ga = (real)0.5 * (gl + gr);
        real aa = sqrt(rmax((ga - (real)1) * (ha - (real)0.5 * (ua * ua)), (real)1e-10));
        sl = nmin(a.v - al, ua - aa);
        sr = nmax(b.v + ar, ua + aa);
    }
    real den = guard14(rl * (sl - a.v) - rr * (sr - b.v));
    real ss = (b.p - a.p + (rl * a.v) * (sl - a.v) - (rr * b.v) * (sr - b.v)) / den;
    ss = nmin(nmax(ss, sl), sr);
    S[0] = sl;
    S[1] = ss;
    S[2] = sr;
}

// Sweby-bounded limited slope of one component (torchSolver.py:201-275).
template <int LIM>
__device__ __forceinline__ real slope1(real delta, real dup, real ddown, real w) {
    real r = zero_bad(dup / ddown);
    real den = (real)1 - w + ((real)1 + w) * r;
    real eps_r = zero_bad((real)2 / (fabs(den) < (real)1e-12 ? (real)1e-12 : den));
    real eps = 0;
    if (LIM == VANLEER) {
        if (r > 0) eps = nmin(zero_bad(((real)2 * r) / ((real)1 + r)), eps_r);
    } else if (LIM == ALBADA) {
        if (r > 0) eps = nmin(zero_bad((r * ((real)1 + r)) / ((real)1 + r * r)), eps_r);
    } else if (LIM == MINBEE) {
        if (r > 0 && r < (real)1) eps = r;
        if (r >= (real)1) eps = nmin((real)1, eps_r);
    } else if (LIM == SUPERBEE) {
        if (r > 0)
            eps = nmin(zero_bad(nmax((real)0, nmax(nmin((real)2 * r, (real)1), nmin(r, (real)2)))), eps_r);
    } else {
        eps = nmin(nmax((real)0, nmin(nmin((real)2, (real)2 * r), ((real)2 + r) / (real)3)), eps_r);
    }
    eps = nmax((real)0, zero_bad(eps));
    real sl = eps * delta;
    real bound = (real)2 * nmin(fabs(dup), fabs(ddown));
    real sgn = ddown > 0 ? (real)1 : (ddown < 0 ? (real)-1 : (real)0);
    return dup * ddown > 0 ? sgn * nmin(fabs(sl), bound) : (real)0;
}

struct Scratch1 {
    real* S;   // 3 x (n+1) wave speeds
    real* SL;  // 3 x n left sub-cell states
    real* SR;  // 3 x n right sub-cell states
    real* F;   // 3 x (n+1) interface fluxes
};

static Scratch1 split(void* scratch, int n) {
    real* p = (real*)scratch;
    Scratch1 s;
    s.S = p;
    s.SL = s.S + 3 * (n + 1);
    s.SR = s.SL + 3 * n;
    s.F = s.SR + 3 * n;
    return s;
}

__global__ void k_waves1(Step s, int n, int bc, Scratch1 w) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > n) return;
    real L[3], R[3], S[3];
    cell1((const real*)s.U, n, bc, f, L);
    cell1((const real*)s.U, n, bc, f + 1, R);
    waves1(L, R, gamma1(s, n, bc, f), gamma1(s, n, bc, f + 1), s.wavespeed, S);
    for (int c = 0; c < 3; ++c) w.S[c * (n + 1) + f] = S[c];
}

template <int LIM>
__global__ void k_recon1(Step s, int n, int bc, Scratch1 w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    real lo[3], in[3], hi[3];
    cell1((const real*)s.U, n, bc, i, lo);
    cell1((const real*)s.U, n, bc, i + 1, in);
    cell1((const real*)s.U, n, bc, i + 2, hi);
    const real g = gamma1(s, n, bc, i + 1);
    real wv = 0;
    if (s.controller == 0) {
        P1 q = prim1(in, g);
        real mach = zero_bad(fabs(q.v) / rmax(sound1(q, g), (real)1e-12));
        wv = nmin(nmax(((real)1 - mach) / ((real)1 + mach), (real)-1), (real)1);
    }
    real ul[3], ur[3], fl[3], fr[3];
    for (int c = 0; c < 3; ++c) {
        real d = slope1<LIM>((real)0.5 * (hi[c] - lo[c]), in[c] - lo[c], hi[c] - in[c], wv);
        ul[c] = in[c] - (real)0.5 * d;
        ur[c] = in[c] + (real)0.5 * d;
    }
    eflux1(ur, g, fr);
    eflux1(ul, g, fl);
    // -(dt / (2 dx)); 0.5*dt/dx rounds to the same double.
    const real k = -(real)s.hc[0];
    for (int side = 0; side < 2; ++side) {
        real st[3];
        for (int c = 0; c < 3; ++c) st[c] = (side == 0 ? ul[c] : ur[c]) + k * (fr[c] - fl[c]);
        real p = (g - (real)1) * (st[2] - ((real)0.5 * (st[1] * st[1])) / rmax(st[0], (real)1e-12));
        bool bad = st[0] <= (real)1e-12 || p <= (real)1e-12 || !isfinite(st[0]) || !isfinite(st[1]) ||
                   !isfinite(st[2]);
        real* dst = side == 0 ? w.SL : w.SR;
        for (int c = 0; c < 3; ++c) dst[c * n + i] = bad ? in[c] : st[c];
    }
}

__device__ void hllc1(const real* ul, const real* ur, real gl, real gr, const real* S, int variant, real* out) {
    P1 a = prim1(ul, gl), b = prim1(ur, gr);
    real rl = rmax(a.rho, (real)1e-12), rr = rmax(b.rho, (real)1e-12);
    real el = rl * ((real)0.5 * (a.v * a.v) + a.p / ((gl - (real)1) * rl));
    real er = rr * ((real)0.5 * (b.v * b.v) + b.p / ((gr - (real)1) * rr));
    real fl[3], fr[3];
    eflux1(ul, gl, fl);
    eflux1(ur, gr, fr);
    const real sl = S[0], sr = S[2];
    real den = guard14(rl * (sl - a.v) - rr * (sr - b.v));
    real ss = (b.p - a.p + (rl * a.v) * (sl - a.v) - (rr * b.v) * (sr - b.v)) / den;
    ss = nmin(nmax(ss, sl), sr);
    const real eps = (real)1e-12;
    real fsl[3], fsr[3];
    if (variant == 0) {
        bool okl = fabs(sl - ss) > eps, okr = fabs(sr - ss) > eps;
        real dl = okl ? sl - ss : -eps, dr = okr ? sr - ss : eps;
        real dlu = fabs(sl - a.v) < eps ? -eps : sl - a.v;
        real dru = fabs(sr - b.v) < eps ? eps : sr - b.v;
        real tl = el / rl + (ss - a.v) * (ss + (a.p / (rl * dlu)));
        real tr = er / rr + (ss - b.v) * (ss + (b.p / (rr * dru)));
        real kl = rl * ((sl - a.v) / dl), kr = rr * ((sr - b.v) / dr);
        real usl[3] = {kl * (real)1, kl * ss, kl * tl}, usr[3] = {kr * (real)1, kr * ss, kr * tr};
        for (int c = 0; c < 3; ++c) {
            real vl = okl ? usl[c] : ul[c], vr = okr ? usr[c] : ur[c];
            fsl[c] = fl[c] + sl * (vl - ul[c]);
            fsr[c] = fr[c] + sr * (vr - ur[c]);
        }
    } else {
        real ds[3] = {(real)0, (real)1, ss};
        real dsl = sl - ss + (real)1e-30, dsr = sr - ss + (real)1e-30;
        real sgl = dsl > 0 ? (real)1 : (dsl < 0 ? (real)-1 : (real)0);
        real sgr = dsr > 0 ? (real)1 : (dsr < 0 ? (real)-1 : (real)0);
        real dl = fabs(sl - ss) < eps ? sgl * eps : sl - ss;
        real dr = fabs(sr - ss) < eps ? sgr * eps : sr - ss;
        for (int c = 0; c < 3; ++c) {
            fsl[c] = (ss * (sl * ul[c] - fl[c]) + sl * (a.p + rl * (sl - a.v) * (ss - a.v)) * ds[c]) / dl;
            fsr[c] = (ss * (sr * ur[c] - fr[c]) + sr * (b.p + rr * (sr - b.v) * (ss - b.v)) * ds[c]) / dr;
        }
    }
    for (int c = 0; c < 3; ++c) {
        real o = 0;
        if (sl > 0) o = fl[c];
        if (sl <= 0 && ss > 0) o = fsl[c];
        if (ss <= 0 && sr >= 0) o = fsr[c];
        if (sr < 0) o = fr[c];
        out[c] = o;
    }
}

__global__ void k_flux1(Step s, int n, int bc, Scratch1 w) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > n) return;
    real ul[3], ur[3], S[3], F[3];
    // Sub-cell ghosts (bc_creator applied to the sub-cell array).
    if (f == 0) {
        const bool wrap = bc == BC_PERIODIC;
        for (int c = 0; c < 3; ++c) ul[c] = wrap ? w.SR[c * n + n - 1] : w.SL[c * n];
        if (bc == BC_REFLECTIVE) ul[1] = -ul[1];
    } else {
        for (int c = 0; c < 3; ++c) ul[c] = w.SR[c * n + f - 1];
    }
    if (f == n) {
        const bool wrap = bc == BC_PERIODIC;
        for (int c = 0; c < 3; ++c) ur[c] = wrap ? w.SL[c * n] : w.SR[c * n + n - 1];
        if (bc == BC_REFLECTIVE) ur[1] = -ur[1];
    } else {
        for (int c = 0; c < 3; ++c) ur[c] = w.SL[c * n + f];
    }
    for (int c = 0; c < 3; ++c) S[c] = w.S[c * (n + 1) + f];
    hllc1(ul, ur, gamma1(s, n, bc, f), gamma1(s, n, bc, f + 1), S, s.variant, F);
    for (int c = 0; c < 3; ++c) w.F[c * (n + 1) + f] = F[c];
}

// LLF interface flux on cell averages, run only after a failed update.
__global__ void k_llf_flux1(Step s, int n, int bc, Scratch1 w) {
    if (!(*s.flags & fvl::F_UPD_BAD)) return;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f > n) return;
    real L[3], R[3], fl[3], fr[3];
    cell1((const real*)s.U, n, bc, f, L);
    cell1((const real*)s.U, n, bc, f + 1, R);
    real gl = gamma1(s, n, bc, f), gr = gamma1(s, n, bc, f + 1);
    P1 a = prim1(L, gl), b = prim1(R, gr);
    real al = nmax(fabs(a.v) + sound1(a, gl), fabs(b.v) + sound1(b, gr));
    eflux1(L, gl, fl);
    eflux1(R, gr, fr);
    for (int c = 0; c < 3; ++c) w.F[c * (n + 1) + f] = (real)0.5 * (fl[c] + fr[c] - al * (R[c] - L[c]));
}

// Conservative update; `second` selects the pass (second order or LLF).
__global__ void k_update1(Step s, int n, int bc, Scratch1 w, int second) {
    const int flag_in = *s.flags;
    if (!second && !(flag_in & fvl::F_UPD_BAD)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    real in[3], o[3];
    cell1((const real*)s.U, n, bc, i + 1, in);
    const real k = -(real)s.c[0];
    for (int c = 0; c < 3; ++c) o[c] = in[c] + k * (w.F[c * (n + 1) + i + 1] - w.F[c * (n + 1) + i]);
    for (int c = 0; c < 3; ++c) ((real*)s.Uout)[c * (long)n + i] = o[c];
    bool ok = isfinite(o[0]) && isfinite(o[1]) && isfinite(o[2]) && o[0] > (real)1e-12 &&
              o[2] - ((real)0.5 * (o[1] * o[1])) / o[0] > (real)1e-12;
    if (!ok) atomicOr(s.flags, second ? fvl::F_UPD_BAD : fvl::F_LLF_BAD);
}

__global__ void k_species1(Step s, int n, int bc, Scratch1 w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const real* U = (const real*)s.U;
    const real rn = rmax(((const real*)s.Uout)[i], (real)1e-12);
    for (int j = 0; j < s.nq; ++j) {
        const real* q = (const real*)s.q[j];
        real qm = q[ghost_src(i - 1, n, bc == BC_PERIODIC ? BC_PERIODIC : BC_OPEN)], qc = q[i];
        real qp = q[ghost_src(i + 1, n, bc == BC_PERIODIC ? BC_PERIODIC : BC_OPEN)];
        real f0 = w.F[i], f1 = w.F[i + 1];
        real sl = f0 * (f0 >= 0 ? qm : qc), sh = f1 * (f1 >= 0 ? qc : qp);
        real rq = U[i] * qc - (real)s.c[0] * (sh - sl);
        ((real*)s.qout[j])[i] = rmin(rmax(rq / rn, (real)0), (real)1);
    }
}

__global__ void k_bound1(Step s, int n, Scratch1 w) {
    if (threadIdx.x != 0) return;
    double* p = s.partials;
    for (int k = 0; k < fvl::NSLOT; ++k) p[k] = 0.0;
    for (int c = 0; c < 3; ++c)
        p[fvl::S_OUT + c] = s.w[0] * (double)(w.F[c * (n + 1) + n] - w.F[c * (n + 1)]);
}

// CFL maxima: max |S| over faces (slot 0), max |u|+a over padded cells (slot 1).
__global__ void k_cfl1(Step s, int n, int bc, double* partials) {
    __shared__ double red[4][2];
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    double m0 = 0, m1 = 0;
    if (f <= n) {
        real L[3], R[3], S[3];
        cell1((const real*)s.U, n, bc, f, L);
        cell1((const real*)s.U, n, bc, f + 1, R);
        real gl = gamma1(s, n, bc, f), gr = gamma1(s, n, bc, f + 1);
        waves1(L, R, gl, gr, s.wavespeed, S);
        for (int c = 0; c < 3; ++c) m0 = fmax(m0, (double)fabs(S[c]));
        P1 a = prim1(L, gl);
        m1 = (double)(fabs(a.v) + sound1(a, gl));
        if (f == n) {
            P1 b = prim1(R, gr);
            m1 = fmax(m1, (double)(fabs(b.v) + sound1(b, gr)));
        }
    }
    m0 = wave_max(m0);
    m1 = wave_max(m1);
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x / 64][0] = m0;
        red[threadIdx.x / 64][1] = m1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* p = partials + blockIdx.x * (long)fvl::NSLOT;
        for (int k = 0; k < fvl::NSLOT; ++k) p[k] = 0.0;
        for (int k = 0; k < 4; ++k) {
            p[0] = fmax(p[0], red[k][0]);
            p[1] = fmax(p[1], red[k][1]);
        }
    }
}

size_t step1d_scratch(const Grid& g) { return sizeof(real) * (size_t)(3 * (g.nx + 1) * 2 + 6 * g.nx + 64); }
int step1d_blocks(const Grid&) { return 1; }
int cfl1d_blocks(const Grid& g) { return (g.nx + 1 + 255) / 256; }

hipError_t step1d(const Grid& g, const Step& s, void* scratch, hipStream_t st) {
    const int n = g.nx, bc = g.bc[0];
    Scratch1 w = split(scratch, n);
    const unsigned bf = (n + 1 + 255) / 256, bcn = (n + 255) / 256;
    k_waves1<<<bf, 256, 0, st>>>(s, n, bc, w);
    switch (s.limiter) {
        case VANLEER: k_recon1<VANLEER><<<bcn, 256, 0, st>>>(s, n, bc, w); break;
        case ALBADA: k_recon1<ALBADA><<<bcn, 256, 0, st>>>(s, n, bc, w); break;
        case MINBEE: k_recon1<MINBEE><<<bcn, 256, 0, st>>>(s, n, bc, w); break;
        case SUPERBEE: k_recon1<SUPERBEE><<<bcn, 256, 0, st>>>(s, n, bc, w); break;
        default: k_recon1<UNOS><<<bcn, 256, 0, st>>>(s, n, bc, w); break;
    }
    k_flux1<<<bf, 256, 0, st>>>(s, n, bc, w);
    k_update1<<<bcn, 256, 0, st>>>(s, n, bc, w, 1);
    k_llf_flux1<<<bf, 256, 0, st>>>(s, n, bc, w);
    k_update1<<<bcn, 256, 0, st>>>(s, n, bc, w, 0);
    if (s.nq > 0) k_species1<<<bcn, 256, 0, st>>>(s, n, bc, w);
    k_bound1<<<1, 64, 0, st>>>(s, n, w);
    return hipGetLastError();
}

hipError_t cfl1d(const Grid& g, const Step& s, hipStream_t st) {
    k_cfl1<<<cfl1d_blocks(g), 256, 0, st>>>(s, g.nx, g.bc[0], s.partials);
    return hipGetLastError();
}

}  // namespace FV_NS
