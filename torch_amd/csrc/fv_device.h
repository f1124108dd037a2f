// Device-side building blocks of the Cartesian finite-volume step.
//
// Compiled twice from the same source: with FV_REAL=double (the plain method,
// -ffp-contract=off so every product and sum rounds like the reference's
// NumPy arithmetic) and with FV_REAL=float (the FP32 form, contraction and
// fast division/square root). Operation order follows the reference
// expressions cited beside each function so the FP64 build stays within a
// few ulps of python/cfdSolver (torchFlux.py, torchSolver3D.py,
// torchSolver2D.py, torchSolver.py).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef FV_REAL
#define FV_REAL double
#endif

typedef FV_REAL real;

namespace fv {

enum Limiter { VANLEER = 0, ALBADA = 1, MINBEE = 2, SUPERBEE = 3, UNOS = 4 };
enum Boundary { BC_OPEN = 0, BC_PERIODIC = 1, BC_REFLECTIVE = 2, BC_INFLOW = 3 };

// Floors and clamps: the non-NaN operand wins (a floor maps NaN to the
// floor, like the comparisons the reference's floors reduce to here). Every
// call site has a non-NaN second operand or operands that cannot tie at +-0,
// so the single-instruction IEEE maxNum/minNum is exact.
__device__ __forceinline__ real rmax(real a, real b) { return fmax(a, b); }
__device__ __forceinline__ real rmin(real a, real b) { return fmin(a, b); }
// NumPy's maximum/minimum propagate NaN; so do these (a NaN operand wins).
template <typename T>
__device__ __forceinline__ T nmax(T a, T b) { return (a != a || b != b) ? (a + b) : (a > b ? a : b); }
template <typename T>
__device__ __forceinline__ T nmin(T a, T b) { return (a != a || b != b) ? (a + b) : (a < b ? a : b); }

// |x| < 1e-14 -> +1e-14 (torchFlux.py:190, torchSolver3D.py:90).
__device__ __forceinline__ real guard14(real x) { return fabs(x) < (real)1e-14 ? (real)1e-14 : x; }

// Ghost-cell source index along one axis of length n for a (possibly
// far out-of-range) index i, and whether i is a ghost. Periodic wraps,
// Open repeats the edge cell, Reflective mirrors (np.pad wrap/edge/
// symmetric, torchSolver3D.py:47-70). Indices beyond two ghost layers are
// first clamped so that masked lanes still read inside the array.
__device__ __forceinline__ int ghost_src(int i, int n, int kind) {
    if (i < -2) i = -2;
    if (i > n + 1) i = n + 1;
    if (i >= 0 && i < n) return i;
    if (kind == BC_PERIODIC) return i < 0 ? i + n : i - n;
    if (kind == BC_REFLECTIVE) return i < 0 ? -i - 1 : 2 * n - 1 - i;
    return i < 0 ? 0 : n - 1;  // open / transmissive (also Inflow's upper side)
}

__device__ __forceinline__ bool ghost_flip(int i, int n, int kind) {
    return kind == BC_REFLECTIVE && (i < 0 || i >= n);
}

// Limiter phi(r) of the 2D/3D MUSCL schemes (torchSolver3D.py:91-110), in
// the arithmetic type T of the caller.
template <int LIM, typename T>
__device__ __forceinline__ T phi3(T r) {
    if (LIM == VANLEER) {
        T pos = nmax(r, (T)0);
        return ((T)2 * pos) / ((T)1 + pos);
    } else if (LIM == ALBADA) {
        T pos = nmax(r, (T)0);
        return (pos * pos + pos) / (pos * pos + (T)1);
    } else if (LIM == MINBEE) {
        return nmax((T)0, nmin(r, (T)1));
    } else if (LIM == SUPERBEE) {
        return nmax((T)0, nmax(nmin((T)2 * r, (T)1), nmin(r, (T)2)));
    } else {
        return nmax((T)0, nmin(nmin(nmin((T)2, (T)2 * r), ((T)2 + r) / (T)3), ((T)1 + (T)2 * r) / (T)3));
    }
}

// 2D limiters differ: VanLeer/Albada use where(r > 0, ..., 0) and Unos has no
// (1+2r)/3 bound (torchSolver2D.py:144-168).
template <int LIM, typename T>
__device__ __forceinline__ T phi2(T r) {
    if (LIM == VANLEER) {
        T pos = nmax(r, (T)0);
        return r > (T)0 ? ((T)2 * pos) / ((T)1 + pos) : (T)0;
    } else if (LIM == ALBADA) {
        T pos = nmax(r, (T)0);
        return r > (T)0 ? (pos * pos + pos) / (pos * pos + (T)1) : (T)0;
    } else if (LIM == MINBEE) {
        return nmax((T)0, nmin(r, (T)1));
    } else if (LIM == SUPERBEE) {
        return nmax((T)0, nmax(nmin((T)2 * r, (T)1), nmin(r, (T)2)));
    } else {
        return nmax((T)0, nmin(nmin((T)2, (T)2 * r), ((T)2 + r) / (T)3));
    }
}

// Largest |minus| for which the reference's float64 phi(minus / 1e-14) is
// finite, so that phi * plus is a zero when plus == 0. Albada squares r
// (pos * pos overflows from |r| ~ 1.34e154, i.e. |minus| ~ 1.34e140, and
// inf / inf is NaN); VanLeer doubles it (overflow from |minus| ~ 8.9e293);
// MINBEE, Superbee and Unos clamp it and stay finite for every non-NaN r.
// Below these bounds (with a margin) the zero is exact; at and above them the
// literal expression runs and yields the reference's NaN.
template <int LIM>
__device__ __forceinline__ constexpr double zero_plus_bound() {
    return LIM == ALBADA ? 1e140 : 1e280;
}

// Reciprocal, division and square root. The FP32 form uses the hardware
// reciprocal and square root (one instruction each, about 1 ulp). The FP64
// plain method refines the hardware reciprocal and reciprocal square root
// with Newton-Raphson steps in fused multiply-adds instead of the IEEE
// division and square-root sequences (about twice the instructions, with
// scaling and special-case fix-ups): 1/x by r' = r + r (1 - x r), twice; a/b
// as q = a r plus one residual correction q' = q + r (a - b q); sqrt(x) by
// the coupled iteration on g ~ sqrt(x), h ~ 1/(2 sqrt(x)) from y = rsq(x)
// (g = x y, h = y/2; r = 1/2 - g h; g' = g + g r, h' = h + h r) and a final
// residual step g'' = g' + h' (x - g'^2). Each lands within a few ulp of the
// correctly rounded result, far inside the 1e-12 bar. They are used where
// the operands are positive and normal by construction (floored densities
// and pressures, guarded denominators) or where any non-finite result
// already decides the step (a non-positive predicted density makes the
// predictor non-finite either way). crcp / csqrt stay correctly rounded in
// the FP64 build: the CFL speeds, whose dt must round like the reference's.
#ifdef FV_FAST
__device__ __forceinline__ real rcp(real x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ real dv(real a, real b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ real rsqrt_(real x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ real crcp(real x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ real csqrt(real x) { return __builtin_amdgcn_sqrtf(x); }
#else
__device__ __forceinline__ real rcp(real x) {
    real r = __builtin_amdgcn_rcp(x);
    real e = fma(-x, r, (real)1);
    r = fma(e, r, r);
    e = fma(-x, r, (real)1);
    return fma(e, r, r);
}
__device__ __forceinline__ real dv(real a, real b) {
    const real r = rcp(b);
    const real q = a * r;
    return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ real rsqrt_(real x) {
    const real y = __builtin_amdgcn_rsq(x);
    real g = x * y, h = (real)0.5 * y;
    const real r = fma(-g, h, (real)0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    return fma(fma(-g, g, x), h, g);
}
__device__ __forceinline__ real crcp(real x) { return (real)1 / x; }
__device__ __forceinline__ real csqrt(real x) { return sqrt(x); }
#endif

// Limited slope phi(minus/plus) * plus (torchSolver3D.py:86-111,
// torchSolver2D.py:144-168) without the division by `plus` where the
// limiter allows it. With r = minus/plus, phi(r)*plus is piecewise linear in
// (minus, plus) for MINBEE, Superbee and Unos, and a single quotient for
// VanLeer (2 m p / (m + p)) and Albada ((m^2 + m p) p / (m^2 + p^2)); the
// piece is chosen by comparing |minus| with multiples of |plus| (same sign
// <=> r > 0). At the piece boundaries the limiter is continuous, so a tie
// decided one ulp differently moves the slope by at most an ulp. When
// |plus| < 1e-14 the reference divides by +1e-14 instead (so r may take
// minus's sign alone); that branch is evaluated literally.
template <int LIM, int DIM>
__device__ __forceinline__ real limited(real minus, real plus) {
    // plus == +-0 (uniform neighbours): phi(minus / 1e-14) is finite and >= 0
    // below the limiter's bound, so the product is a zero of plus's sign, up
    // to the sign of a zero slope (which adding it to the cell value erases).
    if (plus == (real)0 && fabs((double)minus) < zero_plus_bound<LIM>()) return plus;
    // Only when some lane of the wave needs it (the test is wave-uniform, so
    // the rare literal path is a real branch, not speculated code). The
    // quotient and phi are evaluated in double in both builds: the FP32 form
    // then overflows to the reference's NaN (and takes the first-order
    // fallback) for exactly the operands where the float64 reference does.
    if (__any(fabs(plus) < (real)1e-14)) {
        if (fabs(plus) < (real)1e-14) {
            const double r = (double)minus / 1e-14;
            return (real)((DIM == 3 ? phi3<LIM>(r) : phi2<LIM>(r)) * (double)plus);
        }
    }
    if (!(minus * plus > (real)0)) {
        // r <= 0 (or NaN): every limiter gives phi = 0 except through NaN.
        return (minus == minus && plus == plus) ? (real)0 * plus : minus + plus;
    }
    const real am = fabs(minus), ap = fabs(plus);
    if (LIM == MINBEE) {
        return am >= ap ? plus : minus;
    } else if (LIM == SUPERBEE) {
        if ((real)2 * am <= ap) return (real)2 * minus;
        if (am <= ap) return plus;
        if (am <= (real)2 * ap) return minus;
        return (real)2 * plus;
    } else if (LIM == UNOS) {
        const real third = (real)(1.0 / 3.0);
        real m = rmin((real)2 * ap, (real)2 * am);
        m = rmin(m, ((real)2 * ap + am) * third);
        if (DIM == 3) m = rmin(m, (ap + (real)2 * am) * third);
        return plus > (real)0 ? m : -m;
    } else if (LIM == VANLEER) {
        return dv((real)2 * minus * plus, plus + minus);
    } else {
        return dv((minus * minus + minus * plus) * plus, minus * minus + plus * plus);
    }
}

// Physical flux along momentum component m of a 5-component state, no
// floors (torchSolver3D.py:37-44, 77-83).
__device__ __forceinline__ void flux3(const real* u, real g, int m, real* f) {
    const real ir = rcp(u[0]);
    real vel = u[m] * ir;
    real p = (g - (real)1) * (u[4] - ((real)0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3])) * ir);
#pragma unroll
    for (int c = 0; c < 5; ++c) f[c] = u[c] * vel;
    f[m] = f[m] + p;
    f[4] = vel * (u[4] + p);
}

// internal energy > 1e-12 decided exactly: a cheap test with a relative
// margin far wider than its rounding settles almost every state; states
// within the margin of the threshold take the literal IEEE division.
__device__ __forceinline__ bool energy_ok(real e_total, real half_s2, real rho) {
#ifdef FV_FAST
    return e_total - half_s2 * rcp(rho) > (real)1e-12;
#else
    const real approx = half_s2 * (real)__builtin_amdgcn_rcp((double)rho);
    const real slack = fabs(approx) * (real)1e-6 + (real)1e-300;
    const real e = e_total - approx;
    if (e - slack > (real)1e-12) return true;
    if (e + slack <= (real)1e-12) return false;
    return e_total - half_s2 / rho > (real)1e-12;
#endif
}

// physical_state on one 5-component state (torchSolver.py:1038-1046).
// Evaluated without short-circuit branches (bitwise &), so a wave tests
// all its states with straight-line code.
__device__ __forceinline__ bool physical5(const real* s) {
    const bool finite = isfinite(s[0]) & isfinite(s[1]) & isfinite(s[2]) & isfinite(s[3]) & isfinite(s[4]);
    return finite & (s[0] > (real)1e-12) &
           energy_ok(s[4], (real)0.5 * (s[1] * s[1] + s[2] * s[2] + s[3] * s[3]), s[0]);
}

__device__ __forceinline__ bool physical4(const real* s) {
    const bool finite = isfinite(s[0]) & isfinite(s[1]) & isfinite(s[2]) & isfinite(s[3]);
    return finite & (s[0] > (real)1e-12) & energy_ok(s[3], (real)0.5 * (s[1] * s[1] + s[2] * s[2]), s[0]);
}

// Normal-basis HLLC flux (torchFlux.py:125-248). NV = 5 (3D) or 4 (2D);
// rows [rho, m_n, m_t..., E]. Region choice reproduces the reference's
// ordered masks L, *L, *R, R (last assignment wins); a face that matches
// none (NaN speeds) yields NaN.
template <int NV, bool CHECKED = false>
__device__ __forceinline__ void hllc(const real* ul, const real* ur, real gl, real gr, real* out) {
    const int E = NV - 1;
    real gbar = (real)0.5 * (gl + gr);
    real rl = rmax(ul[0], (real)1e-12), rr = rmax(ur[0], (real)1e-12);
    const real irl = rcp(rl), irr = rcp(rr);
    real unl = ul[1] * irl, unr = ur[1] * irr;
    real utl[NV - 3], utr[NV - 3];
    real sl2 = 0, sr2 = 0;
#pragma unroll
    for (int t = 0; t < NV - 3; ++t) {
        utl[t] = ul[2 + t] * irl;
        utr[t] = ur[2 + t] * irr;
        sl2 = t == 0 ? utl[t] * utl[t] : sl2 + utl[t] * utl[t];
        sr2 = t == 0 ? utr[t] * utr[t] : sr2 + utr[t] * utr[t];
    }
    real kel = ((real)0.5 * rl) * (unl * unl + sl2);
    real ker = ((real)0.5 * rr) * (unr * unr + sr2);
    real pl = rmax((gl - (real)1) * (ul[E] - kel), (real)1e-12);
    real pr = rmax((gr - (real)1) * (ur[E] - ker), (real)1e-12);
    real al = rsqrt_((gl * pl) * irl), ar = rsqrt_((gr * pr) * irr);

    real sql = rsqrt_(rl), sqr = rsqrt_(rr);
    real inv = rcp(sql + sqr);
    real u_roe = (sql * unl + sqr * unr) * inv;
    real kv = 0;
#pragma unroll
    for (int t = 0; t < NV - 3; ++t) {
        real v = (sql * utl[t] + sqr * utr[t]) * inv;
        kv = t == 0 ? v * v : kv + v * v;
    }
    real hl = (ul[E] + pl) * irl, hr = (ur[E] + pr) * irr;
    real h_roe = (sql * hl + sqr * hr) * inv;
    real ke_roe = (real)0.5 * (u_roe * u_roe + kv);
    real a_roe = rsqrt_(rmax((gbar - (real)1) * (h_roe - ke_roe), (real)1e-12));
    real s_l, s_r;
    if (CHECKED && sizeof(real) == sizeof(float)) {
        // CHECKED: both states passed physical5 (finite, density and internal
        // energy above 1e-12) or the step is flagged and these fluxes are
        // replaced by the first-order pass. The speeds are then never NaN
        // (at worst +-inf in FP32), so the FP32 form takes the plain
        // hardware min/max for the NaN-propagating ones.
        s_l = fmin(unl - al, u_roe - a_roe);
        s_r = fmax(unr + ar, u_roe + a_roe);
    } else {
        s_l = nmin(unl - al, u_roe - a_roe);
        s_r = nmax(unr + ar, u_roe + a_roe);
    }
    real den = guard14(rl * (s_l - unl) - rr * (s_r - unr));
    real s_star = dv(pr - pl + (rl * unl) * (s_l - unl) - (rr * unr) * (s_r - unr), den);

    // Final region after the ordered masks.
    int region;  // 0 L, 1 *L, 2 *R, 3 R, -1 none
    if ((real)0 >= s_r) region = 3;
    else if (s_star <= (real)0 && (real)0 <= s_r) region = 2;
    else if (s_l <= (real)0 && (real)0 <= s_star) region = 1;
    else if ((real)0 <= s_l) region = 0;
    else region = -1;
    if (region < 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = (real)NAN;
        return;
    }
    bool left = region <= 1;
    const real* u = left ? ul : ur;
    real rho = left ? rl : rr, un = left ? unl : unr, p = left ? pl : pr;
    real s = left ? s_l : s_r;
    real f[NV];
    f[0] = rho * un;
    f[1] = rho * (un * un) + p;
#pragma unroll
    for (int t = 0; t < NV - 3; ++t) f[2 + t] = (rho * un) * (left ? utl[t] : utr[t]);
    f[E] = un * (u[E] + p);
    bool star = (region == 1 || region == 2) && !(fabs(s - s_star) < (real)1e-12);
    if (!star) {
#pragma unroll
        for (int c = 0; c < NV; ++c) out[c] = f[c];
        return;
    }
    real fac = dv(s - un, guard14(s - s_star));
    real rs = rmax(rho * fac, (real)1e-12);
    real us[NV];
    us[0] = rs;
    us[1] = rs * s_star;
#pragma unroll
    for (int t = 0; t < NV - 3; ++t) us[2 + t] = rs * (left ? utl[t] : utr[t]);
    us[E] = rs * (u[E] * (left ? irl : irr) + (s_star - un) * (s_star + dv(p, rho * guard14(s - un))));
#pragma unroll
    for (int c = 0; c < NV; ++c) out[c] = f[c] + s * (us[c] - u[c]);
}

// HLLC with equal left and right states (and gammas): the physical flux
// with the solver's floors, i.e. its F_L. The full solver returns F_L
// itself for a supersonic face and F_L + S_L (U*_L - U_L) with U*_L equal to
// U_L up to rounding otherwise, so this differs from it by rounding only.
template <int NV>
__device__ __forceinline__ void hllc_equal(const real* u, real g, real* out) {
    const int E = NV - 1;
    const real rho = rmax(u[0], (real)1e-12), ir = rcp(rho);
    const real un = u[1] * ir;
    real ut[NV - 3], s2 = 0;
#pragma unroll
    for (int t = 0; t < NV - 3; ++t) {
        ut[t] = u[2 + t] * ir;
        s2 = t == 0 ? ut[t] * ut[t] : s2 + ut[t] * ut[t];
    }
    const real ke = ((real)0.5 * rho) * (un * un + s2);
    const real p = rmax((g - (real)1) * (u[E] - ke), (real)1e-12);
    out[0] = rho * un;
    out[1] = rho * (un * un) + p;
#pragma unroll
    for (int t = 0; t < NV - 3; ++t) out[2 + t] = (rho * un) * ut[t];
    out[E] = un * (u[E] + p);
}

// HLLC between Cartesian-ordered 5-component states across a face normal to
// axis `dir` (0 x, 1 y, 2 z): rotate to [rho, m_n, m_t(asc), E] and back
// (torchFlux.py:36-39, 75-76). Faces between equal states (uniform regions)
// take the solver's own flux of that state without the wave-speed work.
template <bool CHECKED = false>
__device__ __forceinline__ void hllc_cart5(const real* l, const real* r, real gl, real gr, int dir, real* f) {
    real a[5], b[5], o[5];
    int t0 = dir == 0 ? 2 : 1, t1 = dir == 2 ? 2 : 3;
    a[0] = l[0]; a[1] = l[dir + 1]; a[2] = l[t0]; a[3] = l[t1]; a[4] = l[4];
    b[0] = r[0]; b[1] = r[dir + 1]; b[2] = r[t0]; b[3] = r[t1]; b[4] = r[4];
    const bool equal = (a[0] == b[0]) & (a[1] == b[1]) & (a[2] == b[2]) & (a[3] == b[3]) & (a[4] == b[4]) &
                       (gl == gr);
    if (equal) hllc_equal<5>(a, gl, o);
    else hllc<5, CHECKED>(a, b, gl, gr, o);
    f[0] = o[0]; f[dir + 1] = o[1]; f[t0] = o[2]; f[t1] = o[3]; f[4] = o[4];
}

__device__ __forceinline__ void hllc_cart4(const real* l, const real* r, real gl, real gr, int dir, real* f) {
    real a[4], b[4], o[4];
    int t = dir == 0 ? 2 : 1;
    a[0] = l[0]; a[1] = l[dir + 1]; a[2] = l[t]; a[3] = l[3];
    b[0] = r[0]; b[1] = r[dir + 1]; b[2] = r[t]; b[3] = r[3];
    const bool equal = (a[0] == b[0]) & (a[1] == b[1]) & (a[2] == b[2]) & (a[3] == b[3]) & (gl == gr);
    if (equal) hllc_equal<4>(a, gl, o);
    else hllc<4>(a, b, gl, gr, o);
    f[0] = o[0]; f[dir + 1] = o[1]; f[t] = o[2]; f[3] = o[3];
}

// Range guard of transported fractions (torchSolver3D.py:254-255): the
// reference's 1e-10 absorbs double rounding of rho*q/rho; FP32 storage
// rounds that quotient to ~6e-8, so the FP32 form widens the guard to 1e-6.
constexpr real kSpecTol = sizeof(real) == sizeof(double) ? (real)1e-10 : (real)1e-6;

__device__ __forceinline__ real gamma_of(real lam) { return ((real)4.5 - lam) / ((real)3.5 - lam); }
// Per-cell gamma from a lambda array (gamma_calc, torchSolver.py:71-78) or
// taken as stored when the array already holds gamma.
__device__ __forceinline__ real gamma_from(real v, bool direct) { return direct ? v : gamma_of(v); }

// Floored primitives and sound speed of a 2D cell (torchSolver2D.py:38-58,
// 85-91, 420-447): the 2D CFL, LLF and inlet-state arithmetic.
struct Prim2 {
    real rho, ux, vy, p, a;
};
// Contraction is off: the 2D CFL speeds come from here, and they must round
// the same in the fused update and in a standalone pass over a stored state.
__device__ __forceinline__ Prim2 prim2(const real* u, real g) {
#pragma clang fp contract(off)
    Prim2 q;
    q.rho = rmax(u[0], (real)1e-12);
    q.ux = u[1] / q.rho;
    q.vy = u[2] / q.rho;
    real ke = ((real)0.5 * q.rho) * (q.ux * q.ux + q.vy * q.vy);
    q.p = rmax((g - (real)1) * (u[3] - ke), (real)1e-12);
    q.a = sqrt((g * rmax(q.p, (real)1e-12)) / rmax(q.rho, (real)1e-12));
    return q;
}

// Whole-wave shifts by one lane on the DPP path (wave_shr:1 = 0x138,
// wave_shl:1 = 0x130): a VALU move instead of an LDS permute round trip.
// Lane i receives lane i-1 (up) or lane i+1 (down); the edge lane (0 or 63)
// receives 0. Callers use neither edge lane's result.
__device__ __forceinline__ int dpp_up1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_down1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float shfl_up1(float v) { return __int_as_float(dpp_up1(__float_as_int(v))); }
__device__ __forceinline__ float shfl_down1(float v) { return __int_as_float(dpp_down1(__float_as_int(v))); }
__device__ __forceinline__ double shfl_up1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_up1((int)b), hi = dpp_up1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ double shfl_down1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_down1((int)b), hi = dpp_down1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Wave/block max of a nonnegative double (order independent, so exact).
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace fv
