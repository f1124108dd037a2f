// Metric/face finite volumes: second-order primitive MUSCL with Rusanov
// faces on logically uniform Cartesian, cylindrical, annular and spherical
// grids with per-face boundaries, advanced by SSPRK2
// (python/cfdSolver/torchFiniteVolume.py:16-276, torchGeometry.py:93-213,
// torchBoundaries.py:177-246, torchFlux.py:80-122).
//
// One thread per cell. For each direction a thread evaluates the two faces
// of its cell: each face's left/right states come from the padded line
// (interior cells plus one boundary plane per end), with limited slopes of
// the primitives (zeroed where half a slope would take density or pressure
// to <= 1e-12; wrapped on periodic axes) and mirrored states on walls. A
// face flux is a pure function of the line, so the two cells that share a
// face compute it bit for bit alike, and the area-weighted divergence is
// the reference's diff / volume. The thread of each line's first cell also
// evaluates the line's last face, so the outward boundary flux is summed as
// the reference sums it (last minus first per line). Volumes, areas and the
// geometric sources come from per-axis factor tables (outer products in the
// reference's order).
#include "fv_device.h"
#include "fv_launch.h"

namespace FV_NS {
using namespace fv;
using fvl::MFace;
using fvl::FACE_OUTFLOW;
using fvl::FACE_WALL;
using fvl::FACE_PERIODIC;
using fvl::FACE_INFLOW;
using fvl::FACE_PRESSURE_OUTLET;
using fvl::FACE_STAGNATION;
using fvl::GEO_CARTESIAN;
using fvl::GEO_CYLINDRICAL;
using fvl::GEO_SPHERICAL;
using fvl::S_MSRC;

namespace {

constexpr int MT = 256;  // threads per block

// NaN-propagating maximum (np.maximum).
__device__ __forceinline__ real nmax(real a, real b) { return (a > b || a != a) ? a : b; }

struct Pad {      // a padded line entry: primitives, gamma, fractions
    real p[5];
    real g;
    real q[2];
};

struct Line {     // one grid line along direction d through a cell
    int d, n;     // direction and cells along it
    long base;    // cell index of the line's first cell
    long stride;  // index step along d
};

struct Ctx {
    const real* S;
    const real* q[2];
    long N;
    int dim, nf, nq, reactive;
    real gamma;
};

__device__ __forceinline__ void to_prim(const real* u, int nf, real g, real* p) {
    const real rho = u[0];
    p[0] = rho;
    real s2 = 0;
    for (int k = 1; k < nf - 1; ++k) {
        p[k] = u[k] / rho;
        s2 = k == 1 ? u[k] * u[k] : s2 + u[k] * u[k];
    }
    p[nf - 1] = (g - (real)1) * (u[nf - 1] - ((real)0.5 * s2) / rho);
}

__device__ __forceinline__ void to_cons(const real* p, int nf, real g, real* u) {
    const real rho = p[0];
    u[0] = rho;
    real s2 = 0;
    for (int k = 1; k < nf - 1; ++k) {
        u[k] = p[k] * rho;
        s2 = k == 1 ? p[k] * p[k] : s2 + p[k] * p[k];
    }
    u[nf - 1] = p[nf - 1] / (g - (real)1) + ((real)0.5 * rho) * s2;
}

__device__ __forceinline__ Pad load_cell(const Ctx& c, long i) {
    Pad e;
    real u[5];
    for (int k = 0; k < c.nf; ++k) u[k] = c.S[k * c.N + i];
    e.g = c.reactive ? gamma_of(c.q[0][i]) : c.gamma;
    to_prim(u, c.nf, e.g, e.p);
    for (int j = 0; j < 2; ++j) e.q[j] = j < c.nq ? c.q[j][i] : (real)0;
    return e;
}

// Boundary-plane state from the edge cell (torchBoundaries.py:177-246);
// side -1 low, +1 high. Errors the reference raises set flag bits.
__device__ void ghost(Pad& e, const MFace& f, int d, int side, int nf, int nq, int reactive, int& bad) {
    const int nrm = d + 1;
    if (f.type == FACE_WALL) {
        e.p[nrm] = -e.p[nrm];
    } else if (f.type == FACE_INFLOW) {
        if (reactive) e.g = gamma_of((real)f.lam);
        if (nq > 0) e.q[0] = (real)f.lam;
        if (nq > 1) e.q[1] = (real)f.beta;
        for (int k = 0; k < nf; ++k) e.p[k] = (real)f.prim[k];
    } else if (f.type == FACE_STAGNATION) {
        const real gin = e.g;
        const real g = reactive ? gamma_of((real)f.lam) : gin;
        const real a_in = sqrt((gin * e.p[nf - 1]) / e.p[0]);
        const real riemann = (real)(-side) * e.p[nrm] - ((real)2 * a_in) / (gin - (real)1);
        const real a0sq = (g * (real)f.gas_constant) * (real)f.total_temperature;
        const real k = (g + (real)1) / (g - (real)1);
        const real disc = riemann * riemann - k * (((real)0.5 * (g - (real)1)) * (riemann * riemann) - a0sq);
        if (disc < (real)0) bad |= fvl::F_BC_RESERVOIR;
        real a = (-riemann + sqrt(disc)) / k;
        real un = riemann + ((real)2 * a) / (g - (real)1);
        if (un >= a) {
            a = sqrt(((real)2 * a0sq) / (g + (real)1));
            un = a;
        }
        if (un < (real)0 || a <= (real)0) bad |= fvl::F_BC_SUPPLY;
        for (int m = 1; m < nf - 1; ++m) e.p[m] = 0;
        e.p[nrm] = (real)(-side) * un;
        e.p[nf - 1] = (real)f.total_pressure * pow((a * a) / a0sq, g / (g - (real)1));
        e.p[0] = (g * e.p[nf - 1]) / (a * a);
        e.g = g;
        if (nq > 0) e.q[0] = (real)f.lam;
        if (nq > 1) e.q[1] = (real)f.beta;
    } else if (f.type == FACE_PRESSURE_OUTLET) {
        const real g = e.g;
        const real a = sqrt((g * e.p[nf - 1]) / e.p[0]);
        const real un = (real)side * e.p[nrm];
        const bool sub = un < a;
        const real P = (real)f.pressure;
        const real rho = e.p[0] * pow(P / e.p[nf - 1], (real)1 / g);
        const real a_new = sqrt((g * P) / rho);
        const real vel = un + ((real)2 * (a - a_new)) / (g - (real)1);
        const bool back = sub && (vel < (real)0 || un < (real)0);
        if (sub) {
            e.p[0] = rho;
            e.p[nf - 1] = P;
            e.p[nrm] = (real)side * vel;
        }
        if (back) {
            if (!f.has_backflow) {
                bad |= fvl::F_BC_BACKFLOW;
            } else {
                for (int k = 0; k < nf; ++k) e.p[k] = (real)f.prim[k];
                if (reactive) e.g = gamma_of((real)f.lam);
                if (nq > 0) e.q[0] = (real)f.lam;
                if (nq > 1) e.q[1] = (real)f.beta;
            }
        }
    }
}

// Padded entry jj in [0, n + 1] of a line (torchFiniteVolume.py:33-63).
__device__ Pad padded(const Ctx& c, const fvl::Metric& m, const Line& L, int jj, int& bad) {
    if (jj >= 1 && jj <= L.n) return load_cell(c, L.base + (jj - 1) * L.stride);
    const bool low = jj == 0;
    const MFace& f = m.face[2 * L.d + (low ? 0 : 1)];
    if (f.type == FACE_PERIODIC) return load_cell(c, L.base + (low ? L.n - 1 : 0) * L.stride);
    Pad e = load_cell(c, L.base + (low ? 0 : L.n - 1) * L.stride);
    ghost(e, f, L.d, low ? -1 : 1, c.nf, c.nq, c.reactive, bad);
    return e;
}

// Limited slope at interior padded index jj (1..n) with the positivity mask.
template <int LIM>
__device__ void slope_interior(const Ctx& c, const fvl::Metric& m, const Line& L, int jj, real* s, int& bad) {
    const Pad a = padded(c, m, L, jj - 1, bad), b = padded(c, m, L, jj, bad), z = padded(c, m, L, jj + 1, bad);
    for (int k = 0; k < c.nf; ++k) s[k] = limited<LIM, 3>(b.p[k] - a.p[k], z.p[k] - b.p[k]);
    const int e = c.nf - 1;
    if (b.p[0] - (real)0.5 * fabs(s[0]) <= (real)1e-12 || b.p[e] - (real)0.5 * fabs(s[e]) <= (real)1e-12)
        for (int k = 0; k < c.nf; ++k) s[k] = s[k] * (real)0;
}

// Slope at padded index jj (0..n+1): zero on boundary planes except on
// periodic axes, which copy the wrapped interior slope.
template <int LIM>
__device__ void slope_at(const Ctx& c, const fvl::Metric& m, const Line& L, int jj, int first_order, real* s,
                         int& bad) {
    for (int k = 0; k < 5; ++k) s[k] = 0;
    if (first_order) return;
    if (jj >= 1 && jj <= L.n) return slope_interior<LIM>(c, m, L, jj, s, bad);
    if (m.face[2 * L.d].type == FACE_PERIODIC) slope_interior<LIM>(c, m, L, jj == 0 ? L.n : 1, s, bad);
}

// Rusanov flux of primitive face states along d (torchFlux.py:80-100).
__device__ void rusanov(const real* l, const real* r, real gl, real gr, int d, int nf, real* F) {
    real ul[5], ur[5], fl[5], fr[5];
    to_cons(l, nf, gl, ul);
    to_cons(r, nf, gr, ur);
    const int n = d + 1, e = nf - 1;
    for (int k = 0; k < nf; ++k) {
        fl[k] = ul[k] * l[n];
        fr[k] = ur[k] * r[n];
    }
    fl[n] = fl[n] + l[e];
    fr[n] = fr[n] + r[e];
    fl[e] = fl[e] + l[n] * l[e];
    fr[e] = fr[e] + r[n] * r[e];
    const real s = nmax(fabs(l[n]) + sqrt((gl * l[e]) / l[0]), fabs(r[n]) + sqrt((gr * r[e]) / r[0]));
    for (int k = 0; k < nf; ++k) F[k] = (real)0.5 * ((fl[k] + fr[k]) - s * (ur[k] - ul[k]));
}

// Flux through face f (0..n) of a line, and the fractions either side.
template <int LIM>
__device__ void face_flux(const Ctx& c, const fvl::Metric& m, const Line& L, int f, int first_order, real* F,
                          real* ql, real* qr, int& bad) {
    Pad a = padded(c, m, L, f, bad), b = padded(c, m, L, f + 1, bad);
    real sa[5], sb[5];
    slope_at<LIM>(c, m, L, f, first_order, sa, bad);
    slope_at<LIM>(c, m, L, f + 1, first_order, sb, bad);
    real lft[5], rgt[5];
    for (int k = 0; k < c.nf; ++k) {
        lft[k] = a.p[k] + (real)0.5 * sa[k];
        rgt[k] = b.p[k] - (real)0.5 * sb[k];
    }
    real gl = a.g, gr = b.g;
    const int n = L.d + 1;
    if (f == 0 && m.face[2 * L.d].type == FACE_WALL) {
        for (int k = 0; k < c.nf; ++k) lft[k] = rgt[k];
        gl = gr;
        lft[n] = -lft[n];
    }
    if (f == L.n && m.face[2 * L.d + 1].type == FACE_WALL) {
        for (int k = 0; k < c.nf; ++k) rgt[k] = lft[k];
        gr = gl;
        rgt[n] = -rgt[n];
    }
    rusanov(lft, rgt, gl, gr, L.d, c.nf, F);
    for (int j = 0; j < 2; ++j) {
        ql[j] = a.q[j];
        qr[j] = b.q[j];
    }
}

__device__ __forceinline__ double product(const double* const* f, int dim, const int* idx) {
    double v = 1.0;
    for (int a = 0; a < dim; ++a) v = v * f[a][idx[a]];
    return v;
}

// Area of direction d's face `face` at the cell's other coordinates.
__device__ __forceinline__ double area_at(const fvl::Metric& m, int dim, int d, const int* idx, int face) {
    int j[3] = {idx[0], idx[1], idx[2]};
    j[d] = face;
    return product(m.af[d], dim, j);
}

// Geometric momentum sources (torchGeometry.py:184-213) from the cell's
// primitives and its area divergences.
__device__ void geometric_source(const fvl::Metric& m, int dim, const real* p, double rad, double cot, real* out) {
    const int nf = dim + 2;
    for (int k = 0; k < nf; ++k) out[k] = 0;
    if (m.geometry == GEO_CARTESIAN) return;
    const real rho = p[0], pr = p[nf - 1];
    const real radial = (real)rad;
    out[1] = pr * radial;
    if (m.geometry == GEO_SPHERICAL) {
        const real inv_r = radial / (real)2;
        real s2 = 0;
        for (int k = 2; k < nf - 1; ++k) s2 = k == 2 ? p[k] * p[k] : s2 + p[k] * p[k];
        out[1] = out[1] + (rho * s2) * inv_r;
        if (dim >= 2) {
            const real cotr = (real)cot;
            out[2] = pr * cotr - ((rho * p[1]) * p[2]) * inv_r;
            if (dim == 3) {
                out[2] = out[2] + (rho * (p[3] * p[3])) * cotr;
                out[3] = ((-rho) * p[3]) * (p[1] * inv_r + p[2] * cotr);
            }
        }
    } else if (dim >= 2 && !(m.geometry == GEO_CYLINDRICAL && dim == 2)) {
        out[1] = out[1] + (rho * (p[2] * p[2])) * radial;
        out[2] = (((-rho) * p[1]) * p[2]) * radial;
    }
}

// physical_state of one state (torchSolver.py:1038-1046).
__device__ __forceinline__ bool physical_n(const real* u, int nf) {
    bool finite = true;
    real s2 = 0;
    for (int k = 0; k < nf; ++k) finite = finite & isfinite(u[k]);
    for (int k = 1; k < nf - 1; ++k) s2 = k == 1 ? u[k] * u[k] : s2 + u[k] * u[k];
    return finite & (u[0] > (real)1e-12) & (u[nf - 1] - ((real)0.5 * s2) / u[0] > (real)1e-12);
}

// Deterministic block sums of `n` slots (tree in LDS) into partials.
__device__ void block_sums(const double* v, int first_slot, int n, double* partials) {
    __shared__ double red[MT];
    for (int k = 0; k < n; ++k) {
        red[threadIdx.x] = v[k];
        __syncthreads();
        for (int s = MT / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) partials[(long)blockIdx.x * fvl::NSLOT + first_slot + k] = red[0];
        __syncthreads();
    }
}

__device__ void block_max(double v, double* partials) {
    __shared__ double red[MT];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = MT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double a = red[threadIdx.x], b = red[threadIdx.x + s];
            red[threadIdx.x] = (a > b || a != a) ? a : b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[(long)blockIdx.x * fvl::NSLOT + 0] = red[0];
        for (int k = 1; k < fvl::NSLOT; ++k) partials[(long)blockIdx.x * fvl::NSLOT + k] = 0.0;
    }
}

__device__ __forceinline__ void cell_coords(const fvl::Grid& g, long i, int* idx) {
    idx[0] = (int)(i % g.nx);
    idx[1] = (int)((i / g.nx) % g.ny);
    idx[2] = (int)(i / ((long)g.nx * g.ny));
}

__device__ __forceinline__ Line line_of(const fvl::Grid& g, int d, long i, const int* idx) {
    Line L;
    L.d = d;
    L.n = d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
    L.stride = d == 0 ? 1 : (d == 1 ? (long)g.nx : (long)g.nx * g.ny);
    L.base = i - (long)idx[d] * L.stride;
    return L;
}

template <int LIM>
__global__ void __launch_bounds__(MT) k_metric_stage(fvl::Grid g, fvl::Metric m, fvl::MStage s) {
    const long N = (long)g.nx * g.ny * g.nz;
    const long i = blockIdx.x * (long)MT + threadIdx.x;
    const int dim = g.dim, nf = dim + 2;
    double sums[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // outward x5, integrated source x5
    int bad = 0;
    if (i < N) {
        Ctx c;
        c.S = (const real*)s.S;
        c.q[0] = (const real*)s.qS[0];
        c.q[1] = (const real*)s.qS[1];
        c.N = N;
        c.dim = dim;
        c.nf = nf;
        c.nq = s.nq;
        c.reactive = s.reactive;
        c.gamma = (real)s.gamma;
        int idx[3];
        cell_coords(g, i, idx);
        const double V = product(m.vf, dim, idx);
        const Pad me = load_cell(c, i);
        double rad = 0, cot = 0;
        if (m.geometry != GEO_CARTESIAN) {
            rad = (area_at(m, dim, 0, idx, idx[0] + 1) - area_at(m, dim, 0, idx, idx[0])) / V;
            if (dim >= 2) cot = (area_at(m, dim, 1, idx, idx[1] + 1) - area_at(m, dim, 1, idx, idx[1])) / V;
        }
        real src[5], k[5], kq[2] = {0, 0};
        geometric_source(m, dim, me.p, rad, cot, src);
        for (int f = 0; f < nf; ++f) k[f] = src[f];
        for (int f = 0; f < nf; ++f) sums[5 + f] = (double)src[f] * V;
        const real Vr = (real)V;
        for (int d = 0; d < dim; ++d) {
            const Line L = line_of(g, d, i, idx);
            const int at = idx[d];
            real Flo[5], Fhi[5], ql[2], qr[2], hl[2], hr[2];
            face_flux<LIM>(c, m, L, at, s.first_order, Flo, ql, qr, bad);
            face_flux<LIM>(c, m, L, at + 1, s.first_order, Fhi, hl, hr, bad);
            const real Alo = (real)area_at(m, dim, d, idx, at), Ahi = (real)area_at(m, dim, d, idx, at + 1);
            real Wlo[5], Whi[5];
            for (int f = 0; f < nf; ++f) {
                Wlo[f] = Flo[f] * Alo;
                Whi[f] = Fhi[f] * Ahi;
                k[f] = k[f] - (Whi[f] - Wlo[f]) / Vr;
            }
            for (int j = 0; j < s.nq; ++j) {
                const real mlo = (Flo[0] * (Flo[0] >= (real)0 ? ql[j] : qr[j])) * Alo;
                const real mhi = (Fhi[0] * (Fhi[0] >= (real)0 ? hl[j] : hr[j])) * Ahi;
                kq[j] = kq[j] - (mhi - mlo) / Vr;
            }
            if (at == 0) {
                // last minus first face of this line (torchFiniteVolume.py:209-217)
                real Fl[5], a2[2], b2[2];
                face_flux<LIM>(c, m, L, L.n, s.first_order, Fl, a2, b2, bad);
                const real Alast = (real)area_at(m, dim, d, idx, L.n);
                for (int f = 0; f < nf; ++f) sums[f] += (double)(Fl[f] * Alast - Wlo[f]);
            }
        }
        real out[5], qo[2] = {0, 0};
        const real* S = (const real*)s.S;
        if (s.mode == 0) {
            for (int f = 0; f < nf; ++f) out[f] = k[f];
            for (int j = 0; j < s.nq; ++j) qo[j] = kq[j];
        } else {
            const real* U0 = (const real*)s.U0;
            const real dt = (real)s.dt;
            for (int f = 0; f < nf; ++f)
                out[f] = s.mode == 1 ? S[f * N + i] + dt * k[f]
                                     : (real)0.5 * ((U0[f * N + i] + S[f * N + i]) + dt * k[f]);
            if (!physical_n(out, nf)) bad |= fvl::F_M_UNPHYS;
            for (int j = 0; j < s.nq; ++j) {
                const real held = U0[i] * ((const real*)s.q0[j])[i];
                real q;
                if (s.mode == 1) {
                    q = (held + dt * kq[j]) / out[0];
                } else {
                    q = ((real)0.5 * ((held + S[i] * ((const real*)s.qS[j])[i]) + dt * kq[j])) / out[0];
                }
                if (q < (real)-1e-12 || q > (real)1 + (real)1e-12) bad |= fvl::F_M_SPEC;
                qo[j] = rmin(rmax(q, (real)0), (real)1);
            }
        }
        for (int f = 0; f < nf; ++f) ((real*)s.out)[f * N + i] = out[f];
        for (int j = 0; j < s.nq; ++j) ((real*)s.qout[j])[i] = qo[j];
    }
    if (bad) atomicOr(s.flags, bad);
    block_sums(sums, fvl::S_OUT, 5, s.partials);
    block_sums(sums + 5, S_MSRC, 5, s.partials);
}

// CFL rate: sum over directions of max(A_lo s_lo, A_hi s_hi) / V with face
// speeds max(|u_d| + a) of the two sides; boundary faces use the boundary
// plane's state (torchFiniteVolume.py:66-98).
__global__ void __launch_bounds__(MT) k_metric_cfl(fvl::Grid g, fvl::Metric m, const real* U, const real* lam,
                                                   double gamma, int reactive, double* partials, int* flags) {
    const long N = (long)g.nx * g.ny * g.nz;
    const long i = blockIdx.x * (long)MT + threadIdx.x;
    const int dim = g.dim, nf = dim + 2;
    double rate = 0;
    int bad = 0;
    if (i < N) {
        Ctx c;
        c.S = U;
        c.q[0] = lam;
        c.q[1] = nullptr;
        c.N = N;
        c.dim = dim;
        c.nf = nf;
        c.nq = 0;
        c.reactive = reactive;
        c.gamma = (real)gamma;
        int idx[3];
        cell_coords(g, i, idx);
        const double V = product(m.vf, dim, idx);
        real r = 0;
        for (int d = 0; d < dim; ++d) {
            const Line L = line_of(g, d, i, idx);
            const int at = idx[d];
            auto speed = [&](int jj) {
                const Pad e = padded(c, m, L, jj, bad);
                return fabs(e.p[d + 1]) + sqrt((e.g * e.p[nf - 1]) / e.p[0]);
            };
            const real sc = speed(at + 1);
            const real slo = nmax(speed(at), sc), shi = nmax(sc, speed(at + 2));
            const real a_lo = (real)area_at(m, dim, d, idx, at) * slo;
            const real a_hi = (real)area_at(m, dim, d, idx, at + 1) * shi;
            r = r + nmax(a_lo, a_hi) / (real)V;
        }
        rate = (double)r;
    }
    if (bad) atomicOr(flags, bad);
    block_max(rate, partials);
}

// Reaction source integral on a curvilinear grid: sum over cells of
// (((old - start) + state) - hydro) * V for the energy row
// (torchSession.py:559-567; the other rows are exactly zero).
__global__ void __launch_bounds__(MT) k_metric_source(fvl::Grid g, fvl::Metric m, const real* start,
                                                      const real* old, const real* state, const real* hydro,
                                                      double* partials) {
    const long N = (long)g.nx * g.ny * g.nz;
    const long i = blockIdx.x * (long)MT + threadIdx.x;
    const int dim = g.dim, e = dim + 1;
    double v[1] = {0};
    if (i < N) {
        int idx[3];
        cell_coords(g, i, idx);
        const double V = product(m.vf, dim, idx);
        const long o = (long)e * N + i;
        v[0] = (double)((((old[o] - start[o]) + state[o]) - hydro[o]) * (real)V);
    }
    block_sums(v, S_MSRC + e, 1, partials);
}

// Stand-alone Rusanov over face arrays (rusanov_flux_grid, torchFlux.py:103-122).
__global__ void k_rusanov_faces(int nv, int dir, long mfaces, const real* l, const real* r, const real* gl,
                                const real* gr, real* out) {
    const long f = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (f >= mfaces) return;
    real a[5], b[5], F[5];
    for (int k = 0; k < nv; ++k) {
        a[k] = l[k * mfaces + f];
        b[k] = r[k * mfaces + f];
    }
    rusanov(a, b, gl[f], gr[f], dir, nv, F);
    for (int k = 0; k < nv; ++k) out[k * mfaces + f] = F[k];
}

}  // namespace

int metric_blocks(const fvl::Grid& g) {
    const long N = (long)g.nx * g.ny * g.nz;
    return (int)((N + MT - 1) / MT);
}

hipError_t metric_stage(const fvl::Grid& g, const fvl::Metric& m, const fvl::MStage& s, hipStream_t st) {
    const int blocks = metric_blocks(g);
    switch (s.limiter) {
        case VANLEER: k_metric_stage<VANLEER><<<blocks, MT, 0, st>>>(g, m, s); break;
        case ALBADA: k_metric_stage<ALBADA><<<blocks, MT, 0, st>>>(g, m, s); break;
        case MINBEE: k_metric_stage<MINBEE><<<blocks, MT, 0, st>>>(g, m, s); break;
        case SUPERBEE: k_metric_stage<SUPERBEE><<<blocks, MT, 0, st>>>(g, m, s); break;
        default: k_metric_stage<UNOS><<<blocks, MT, 0, st>>>(g, m, s); break;
    }
    return hipGetLastError();
}

hipError_t metric_cfl(const fvl::Grid& g, const fvl::Metric& m, const void* U, const void* lam, double gamma,
                      int reactive, double* partials, int* flags, hipStream_t st) {
    k_metric_cfl<<<metric_blocks(g), MT, 0, st>>>(g, m, (const real*)U, (const real*)lam, gamma, reactive,
                                                  partials, flags);
    return hipGetLastError();
}

hipError_t metric_source(const fvl::Grid& g, const fvl::Metric& m, const void* start, const void* old,
                         const void* state, const void* hydro, double* partials, hipStream_t st) {
    k_metric_source<<<metric_blocks(g), MT, 0, st>>>(g, m, (const real*)start, (const real*)old, (const real*)state,
                                                     (const real*)hydro, partials);
    return hipGetLastError();
}

hipError_t rusanov_faces(int nv, int dir, long mfaces, const void* l, const void* r, const void* gl, const void* gr,
                         void* out, hipStream_t st) {
    if (mfaces <= 0) return hipSuccess;
    const int blocks = (int)((mfaces + 255) / 256);
    k_rusanov_faces<<<blocks, 256, 0, st>>>(nv, dir, mfaces, (const real*)l, (const real*)r, (const real*)gl,
                                            (const real*)gr, (real*)out);
    return hipGetLastError();
}

}  // namespace FV_NS
