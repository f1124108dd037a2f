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

// Primitive state, sound speed and face wave speeds in the arithmetic type
// T of the caller: the storage precision, or double for the exact change a
// reactive attempt's heating makes to the CFL maxima (FP32 form).
template <typename T>
struct P1T {
    T rho, v, p;
};
using P1 = P1T<real>;

template <typename T>
__device__ __forceinline__ P1T<T> prim1(const T* u, T g) {
    P1T<T> q;
    q.rho = fmax(u[0], (T)1e-12);
    q.v = u[1] / q.rho;
    q.p = fmax((u[2] - ((T)0.5 * q.rho) * (q.v * q.v)) * (g - (T)1), (T)1e-12);
    return q;
}

template <typename T>
__device__ __forceinline__ T sound1(const P1T<T>& q, T g) {
    return sqrt((g * fmax(q.p, (T)0)) / fmax(q.rho, (T)1e-12));
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
template <typename T>
__device__ void waves1(const T* L, const T* R, T gl, T gr, int mode, T* S) {
    P1T<T> a = prim1(L, gl), b = prim1(R, gr);
    T rl = fmax(a.rho, (T)1e-12), rr = fmax(b.rho, (T)1e-12);
    T al = sound1(a, gl), ar = sound1(b, gr);
    T sl, sr;
    if (mode == 0) {
        T ps = nmax((T)0, (T)0.5 * (a.p + b.p) - (T)0.5 * (b.v - a.v) * ((T)0.5 * (b.rho + a.rho)) * ((T)0.5 * (al + ar)));
        T ql = (T)1, qr = (T)1;
        if (ps > a.p) ql = sqrt((T)1 + (gl + (T)1) / ((T)2 * gl) * (ps / a.p - (T)1));
        if (ps > b.p) qr = sqrt((T)1 + (gr + (T)1) / ((T)2 * gr) * (ps / b.p - (T)1));
        sl = a.v - al * ql;
        sr = b.v + ar * qr;
    } else {
        T elt = rl * ((T)0.5 * (a.v * a.v) + a.p / ((gl - (T)1) * rl));
        T ert = rr * ((T)0.5 * (b.v * b.v) + b.p / ((gr - (T)1) * rr));
        T sql = sqrt(rl), sqr = sqrt(rr);
        T den = fmax(sql + sqr, (T)1e-12);
        T ua = (sql * a.v + sqr * b.v) / den;
        T ha = (sql * ((elt + a.p) / rl) + sqr * ((ert + b.p) / rr)) / den;
        T ga = (T)0.5 * (gl + gr);
        T aa = sqrt(fmax((ga - (T)1) * (ha - (T)0.5 * (ua * ua)), (T)1e-10));
        sl = nmin(a.v - al, ua - aa);
        sr = nmax(b.v + ar, ua + aa);
    }
    T den = rl * (sl - a.v) - rr * (sr - b.v);
    den = fabs(den) < (T)1e-14 ? (T)1e-14 : den;
    T ss = (b.p - a.p + (rl * a.v) * (sl - a.v) - (rr * b.v) * (sr - b.v)) / den;
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


// Sub-cell states (left, right) of cell `src` from the cell averages:
// limited central slope, half-step predictor, states that fail the
// positivity test replaced by the cell average (torchSolver.py:447-491).
template <int LIM>
__device__ __forceinline__ void recon1(const Step& s, int n, int bc, int src, real hc, real* sl, real* sr) {
    const real* U = (const real*)s.U;
    real lo[3], in[3], hi[3];
    cell1(U, n, bc, src, lo);
    cell1(U, n, bc, src + 1, in);
    cell1(U, n, bc, src + 2, hi);
    const real g = gamma1(s, n, bc, src + 1);
    real wv = 0;
    if (s.controller == 0) {
        P1 q = prim1(in, g);
        real mach = zero_bad(fabs(q.v) / rmax(sound1(q, g), (real)1e-12));
        wv = nmin(nmax(((real)1 - mach) / ((real)1 + mach), (real)-1), (real)1);
    }
    real ul[3], ur[3], fl[3], fr[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        real d = slope1<LIM>((real)0.5 * (hi[c] - lo[c]), in[c] - lo[c], hi[c] - in[c], wv);
        ul[c] = in[c] - (real)0.5 * d;
        ur[c] = in[c] + (real)0.5 * d;
    }
    eflux1(ur, g, fr);
    eflux1(ul, g, fl);
    // -(dt / (2 dx)); 0.5*dt/dx rounds to the same double.
    const real k = -hc;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        real st[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) st[c] = (side == 0 ? ul[c] : ur[c]) + k * (fr[c] - fl[c]);
        real p = (g - (real)1) * (st[2] - ((real)0.5 * (st[1] * st[1])) / rmax(st[0], (real)1e-12));
        bool bad = st[0] <= (real)1e-12 || p <= (real)1e-12 || !isfinite(st[0]) || !isfinite(st[1]) ||
                   !isfinite(st[2]);
        real* dst = side == 0 ? sl : sr;
#pragma unroll
        for (int c = 0; c < 3; ++c) dst[c] = bad ? in[c] : st[c];
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

// One block of 256 threads updates T1 cells [i0, i0+T1): thread t builds the
// sub-cell pair of padded cell i0-1+t, then the interface flux of face i0+t,
// then the update of cell i0+t. The interface fluxes also go to the
// instance's flux array (3 x (n+1)), read afterwards by the species and
// boundary-integral passes.
constexpr int T1 = 254;

struct Lds1 {
    real sl[3][256], sr[3][256], f[3][256];
};

__device__ __forceinline__ void store_face(const Step& s, real* Fs, int n, int f, const real* F, Lds1& m, int t) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        m.f[c][t] = F[c];
        Fs[c * (n + 1) + f] = F[c];
    }
    if (s.F[0])
        for (int c = 0; c < 3; ++c) ((real*)s.F[0])[c * (n + 1) + f] = F[c];
}

// Conservative update of cell i0+t from the block's face fluxes; `pass`
// names the flag a nonphysical result raises (conservative_update_1d,
// torchSolver.py:1054-1071).
__device__ __forceinline__ void update1(const Step& s, int n, int bc, int i0, real c0, const Lds1& m, int pass) {
    const int t = threadIdx.x, i = i0 + t;
    if (t >= T1 || i >= n) return;
    real in[3], o[3];
    cell1((const real*)s.U, n, bc, i + 1, in);
    const real k = -c0;
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = in[c] + k * (m.f[c][t + 1] - m.f[c][t]);
#pragma unroll
    for (int c = 0; c < 3; ++c) ((real*)s.Uout)[c * (long)n + i] = o[c];
    bool ok = isfinite(o[0]) && isfinite(o[1]) && isfinite(o[2]) && o[0] > (real)1e-12 &&
              o[2] - ((real)0.5 * (o[1] * o[1])) / o[0] > (real)1e-12;
    if (!ok) atomicOr(s.flags, pass);
}

template <int LIM>
__device__ void step1d_tile(const Step& s, int n, int bc, int tile, real* Fs, Lds1& m) {
    const fvl::Coef K = fvl::coef_of(s);
    const int t = threadIdx.x, i0 = tile * T1, j = i0 - 1 + t;
    if (j <= n) {
        // Sub-cell ghosts are bc_creator applied to the sub-cell arrays:
        // periodic wraps, open copies the adjacent sub-cell, reflective copies
        // it with the momentum negated.
        const bool wrap = bc == BC_PERIODIC, mirror = bc == BC_REFLECTIVE;
        const int src = j < 0 ? (wrap ? n - 1 : 0) : (j >= n ? (wrap ? 0 : n - 1) : j);
        real a[3], b[3];
        recon1<LIM>(s, n, bc, src, (real)K.hc[0], a, b);
        if (j < 0 && !wrap) {
            for (int c = 0; c < 3; ++c) b[c] = a[c];
            if (mirror) b[1] = -b[1];
        }
        if (j >= n && !wrap) {
            for (int c = 0; c < 3; ++c) a[c] = b[c];
            if (mirror) a[1] = -a[1];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            m.sl[c][t] = a[c];
            m.sr[c][t] = b[c];
        }
    }
    __syncthreads();
    const int f = i0 + t;
    if (t <= T1 && f <= n) {
        const real* U = (const real*)s.U;
        real ul[3], ur[3], L[3], R[3], S[3], F[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            ul[c] = m.sr[c][t];
            ur[c] = m.sl[c][t + 1];
        }
        cell1(U, n, bc, f, L);
        cell1(U, n, bc, f + 1, R);
        const real gl = gamma1(s, n, bc, f), gr = gamma1(s, n, bc, f + 1);
        waves1(L, R, gl, gr, s.wavespeed, S);
        hllc1(ul, ur, gl, gr, S, s.variant, F);
        store_face(s, Fs, n, f, F, m, t);
    }
    __syncthreads();
    update1(s, n, bc, i0, (real)K.c[0], m, fvl::F_UPD_BAD);
}

// LLF pass on cell averages, run only after a failed second-order update.
__device__ void llf1d_tile(const Step& s, int n, int bc, int tile, real* Fs, Lds1& m) {
    if (!(*s.flags & fvl::F_UPD_BAD)) return;  // the same answer for every thread of the block
    const fvl::Coef K = fvl::coef_of(s);
    const int t = threadIdx.x, i0 = tile * T1, f = i0 + t;
    if (t <= T1 && f <= n) {
        const real* U = (const real*)s.U;
        real L[3], R[3], fl[3], fr[3], F[3];
        cell1(U, n, bc, f, L);
        cell1(U, n, bc, f + 1, R);
        real gl = gamma1(s, n, bc, f), gr = gamma1(s, n, bc, f + 1);
        P1 a = prim1(L, gl), b = prim1(R, gr);
        real al = nmax(fabs(a.v) + sound1(a, gl), fabs(b.v) + sound1(b, gr));
        eflux1(L, gl, fl);
        eflux1(R, gr, fr);
#pragma unroll
        for (int c = 0; c < 3; ++c) F[c] = (real)0.5 * (fl[c] + fr[c] - al * (R[c] - L[c]));
        store_face(s, Fs, n, f, F, m, t);
    }
    __syncthreads();
    update1(s, n, bc, i0, (real)K.c[0], m, fvl::F_LLF_BAD);
}

#ifdef FV_FAST
// A reactive attempt of the FP32 form (Step.heat: the chemistry's energy and
// fraction increments per cell): the exact change the heating makes to face
// f's |S_L|, |S*|, |S_R| and to |u| + a of its two cells, from the committed
// and the heated states both evaluated in double. With no chemistry in
// either cell the two evaluations are identical and every change is zero.
__device__ void heat_change1(const Step& s, int n, int bc, int f, const real* L, const real* R, double* dS,
                             double& da, double& db) {
    const real* dE = (const real*)s.heat;
    const real* dl = dE + n;
    const int il = ghost_src(f - 1, n, bc), ir = ghost_src(f, n, bc);
    const double ll = s.lam ? (double)((const real*)s.lam)[il] : 0.0;
    const double lr = s.lam ? (double)((const real*)s.lam)[ir] : 0.0;
    const double llh = ll + (double)dl[il], lrh = lr + (double)dl[ir];
    auto gam = [&](double l) { return s.lam ? (4.5 - l) / (3.5 - l) : s.gamma; };
    const double Lc[3] = {L[0], L[1], L[2]}, Rc[3] = {R[0], R[1], R[2]};
    const double Lh[3] = {L[0], L[1], (double)L[2] + (double)dE[il]};
    const double Rh[3] = {R[0], R[1], (double)R[2] + (double)dE[ir]};
    double Sc[3], Sh[3];
    waves1(Lc, Rc, gam(ll), gam(lr), s.wavespeed, Sc);
    waves1(Lh, Rh, gam(llh), gam(lrh), s.wavespeed, Sh);
    for (int c = 0; c < 3; ++c) dS[c] = fabs(Sh[c]) - fabs(Sc[c]);
    const P1T<double> ac = prim1(Lc, gam(ll)), ah = prim1(Lh, gam(llh));
    const P1T<double> bc0 = prim1(Rc, gam(lr)), bh = prim1(Rh, gam(lrh));
    da = (fabs(ah.v) + sound1(ah, gam(llh))) - (fabs(ac.v) + sound1(ac, gam(ll)));
    db = (fabs(bh.v) + sound1(bh, gam(lrh))) - (fabs(bc0.v) + sound1(bc0, gam(lr)));
}
#endif

// CFL maxima of faces [f0, f0+256): max |S| (slot 0) and max |u|+a over the
// padded cells (slot 1); for a reactive attempt of the FP32 form (Step.heat)
// those of the committed state raised by the heating's exact change.
__device__ void cfl1d_faces(const Step& s, int n, int bc, int f0, double* partials, long pblk) {
    __shared__ double red[4][2];
    const int f = f0 + threadIdx.x;
    double m0 = 0, m1 = 0;
    if (f <= n) {
        real L[3], R[3], S[3];
        cell1((const real*)s.U, n, bc, f, L);
        cell1((const real*)s.U, n, bc, f + 1, R);
        real gl = gamma1(s, n, bc, f), gr = gamma1(s, n, bc, f + 1);
        waves1(L, R, gl, gr, s.wavespeed, S);
        double dS[3] = {0, 0, 0}, da = 0, db = 0;
#ifdef FV_FAST
        if (s.heat) heat_change1(s, n, bc, f, L, R, dS, da, db);
#endif
        for (int c = 0; c < 3; ++c) m0 = fmax(m0, (double)fabs(S[c]) + dS[c]);
        P1 a = prim1(L, gl);
        m1 = (double)(fabs(a.v) + sound1(a, gl)) + da;
        if (f == n) {
            P1 b = prim1(R, gr);
            m1 = fmax(m1, (double)(fabs(b.v) + sound1(b, gr)) + db);
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
        double* p = partials + pblk * fvl::NSLOT;
        for (int k = 0; k < fvl::NSLOT; ++k) p[k] = 0.0;
        for (int k = 0; k < 4; ++k) {
            p[0] = fmax(p[0], red[k][0]);
            p[1] = fmax(p[1], red[k][1]);
        }
    }
}

template <int LIM>
__global__ void __launch_bounds__(256) k_step1d(Step s, int n, int bc, real* Fs) {
    __shared__ Lds1 m;
    step1d_tile<LIM>(s, n, bc, blockIdx.x, Fs, m);
}

__global__ void __launch_bounds__(256) k_llf1d(Step s, int n, int bc, real* Fs) {
    __shared__ Lds1 m;
    llf1d_tile(s, n, bc, blockIdx.x, Fs, m);
}

__global__ void __launch_bounds__(256) k_cfl1(Step s, int n, int bc, double* partials) {
    cfl1d_faces(s, n, bc, blockIdx.x * 256, partials, blockIdx.x);
}

__global__ void k_species1(Step s, int n, int bc, const real* Fs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const fvl::Coef K = fvl::coef_of(s);
    const real* U = (const real*)s.U;
    const real rn = rmax(((const real*)s.Uout)[i], (real)1e-12);
    const int sbc = bc == BC_PERIODIC ? BC_PERIODIC : BC_OPEN;  // scalar_bc has no mirror
    for (int j = 0; j < s.nq; ++j) {
        const real* q = (const real*)s.q[j];
        real qm = q[ghost_src(i - 1, n, sbc)], qc = q[i], qp = q[ghost_src(i + 1, n, sbc)];
        real f0 = Fs[i], f1 = Fs[i + 1];
        real sl = f0 * (f0 >= 0 ? qm : qc), sh = f1 * (f1 >= 0 ? qc : qp);
        real rq = U[i] * qc - (real)K.c[0] * (sh - sl);
        ((real*)s.qout[j])[i] = rmin(rmax(rq / rn, (real)0), (real)1);
    }
}

// Boundary integral dt*vol/dx * (F_last - F_first) of the final fluxes.
__device__ __forceinline__ void bound1(double w, const real* Fs, int n, double* out) {
    for (int c = 0; c < 3; ++c) out[c] = w * (double)(Fs[c * (n + 1) + n] - Fs[c * (n + 1)]);
}

__global__ void k_bound1(Step s, int n, const real* Fs) {
    if (threadIdx.x != 0) return;
    const fvl::Coef K = fvl::coef_of(s);
    double* p = s.partials;
    for (int k = 0; k < fvl::NSLOT; ++k) p[k] = 0.0;
    bound1(K.w[0], Fs, n, p + fvl::S_OUT);
}

// Work-list forms: instance i owns blocks first[i] .. first[i+1]-1, one
// tile of T1 cells (or 256 CFL faces) per block; one limiter per launch.
template <int LIM>
__global__ void __launch_bounds__(256) k_batch1d(const fvl::BatchItem* items, const int* first, int n) {
    __shared__ Lds1 m;
    const int i = fvl::owner(first, n, blockIdx.x);
    const fvl::BatchItem& it = fvl::item_at(items, i);
    const Step& s = it.step;
    if (s.coef && s.coef[9] == 0.0) return;
    step1d_tile<LIM>(s, it.grid.nx, it.grid.bc[0], blockIdx.x - first[i], (real*)it.scratch, m);
}

__global__ void __launch_bounds__(256) k_batch_llf1d(const fvl::BatchItem* items, const int* first, int n) {
    __shared__ Lds1 m;
    const int i = fvl::owner(first, n, blockIdx.x);
    const fvl::BatchItem& it = fvl::item_at(items, i);
    if (it.step.coef && it.step.coef[9] == 0.0) return;
    llf1d_tile(it.step, it.grid.nx, it.grid.bc[0], blockIdx.x - first[i], (real*)it.scratch, m);
}

// CFL maxima of each 1D instance's output state (after the step) or, with
// `input` set, of its input state (the first step of a new instance).
__global__ void __launch_bounds__(256) k_batch_cfl1d(const fvl::BatchItem* items, const int* first, int n,
                                                     int input) {
    const int i = fvl::owner(first, n, blockIdx.x);
    const fvl::BatchItem& it = fvl::item_at(items, i);
    if (!input && it.step.coef && it.step.coef[9] == 0.0) return;
    Step s = it.step;
    if (!input) s.U = s.Uout;
    const int local = blockIdx.x - first[i];
    cfl1d_faces(s, it.grid.nx, it.grid.bc[0], local * 256, s.partials, it.cfl_first + local);
}

size_t step1d_scratch(const Grid& g) { return sizeof(real) * (size_t)(3 * (g.nx + 1) + 64); }
int step1d_blocks(const Grid&) { return 1; }
int step1d_tiles(const Grid& g) { return (g.nx + T1 - 1) / T1; }
int cfl1d_blocks(const Grid& g) { return (g.nx + 1 + 255) / 256; }

template <int LIM>
static void launch1(const Grid& g, const Step& s, real* Fs, hipStream_t st) {
    k_step1d<LIM><<<step1d_tiles(g), 256, 0, st>>>(s, g.nx, g.bc[0], Fs);
}

hipError_t step1d(const Grid& g, const Step& s, void* scratch, hipStream_t st) {
    const int n = g.nx, bc = g.bc[0];
    real* Fs = (real*)scratch;
    switch (s.limiter) {
        case VANLEER: launch1<VANLEER>(g, s, Fs, st); break;
        case ALBADA: launch1<ALBADA>(g, s, Fs, st); break;
        case MINBEE: launch1<MINBEE>(g, s, Fs, st); break;
        case SUPERBEE: launch1<SUPERBEE>(g, s, Fs, st); break;
        default: launch1<UNOS>(g, s, Fs, st); break;
    }
    k_llf1d<<<step1d_tiles(g), 256, 0, st>>>(s, n, bc, Fs);
    if (s.nq > 0) k_species1<<<(n + 255) / 256, 256, 0, st>>>(s, n, bc, Fs);
    k_bound1<<<1, 64, 0, st>>>(s, n, Fs);
    return hipGetLastError();
}

hipError_t cfl1d(const Grid& g, const Step& s, hipStream_t st) {
    k_cfl1<<<cfl1d_blocks(g), 256, 0, st>>>(s, g.nx, g.bc[0], s.partials);
    return hipGetLastError();
}

hipError_t batch1d(const fvl::BatchItem* items, const int* first, int n, int blocks, int limiter, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    switch (limiter) {
        case VANLEER: k_batch1d<VANLEER><<<blocks, 256, 0, st>>>(items, first, n); break;
        case ALBADA: k_batch1d<ALBADA><<<blocks, 256, 0, st>>>(items, first, n); break;
        case MINBEE: k_batch1d<MINBEE><<<blocks, 256, 0, st>>>(items, first, n); break;
        case SUPERBEE: k_batch1d<SUPERBEE><<<blocks, 256, 0, st>>>(items, first, n); break;
        default: k_batch1d<UNOS><<<blocks, 256, 0, st>>>(items, first, n); break;
    }
    return hipGetLastError();
}

hipError_t batch_llf1d(const fvl::BatchItem* items, const int* first, int n, int blocks, hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    k_batch_llf1d<<<blocks, 256, 0, st>>>(items, first, n);
    return hipGetLastError();
}

hipError_t batch_cfl1d(const fvl::BatchItem* items, const int* first, int n, int blocks, int input,
                       hipStream_t st) {
    if (blocks <= 0) return hipSuccess;
    k_batch_cfl1d<<<blocks, 256, 0, st>>>(items, first, n, input);
    return hipGetLastError();
}

}  // namespace FV_NS
