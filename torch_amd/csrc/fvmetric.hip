// Metric/face finite volumes: second-order primitive MUSCL with Rusanov
// faces on logically uniform Cartesian, cylindrical, annular and spherical
// grids with per-face boundaries, advanced by SSPRK2
// (python/cfdSolver/torchFiniteVolume.py:16-276, torchGeometry.py:93-213,
// torchBoundaries.py:177-246, torchFlux.py:80-122).
//
// A pass converts its state to primitives once per cell. A stage then takes
// one thread per face along each direction: the face's left/right states
// come from the padded line (interior cells plus one boundary plane per
// end), with limited slopes of the primitives (zeroed where half a slope
// would take density or pressure to <= 1e-12; wrapped on periodic axes) and
// mirrored states on walls; the Rusanov flux times the face area is stored
// once, and one thread per cell forms the reference's diff / volume from
// its stored faces, the geometric source and the SSPRK2 stage. The outward
// boundary flux sums the weighted fluxes of each line's last face minus its
// first (the threads of the end cells contribute them). Kernels are
// instantiated per dimension so every component index is a compile-time
// constant. Volumes, areas and the geometric sources come from per-axis
// factor tables (outer products in the reference's order).
//
// The adaptive leaves (further down) reuse the face states, boundary planes,
// Rusanov flux and stage output on a leaf mesh with an explicit face list.
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

template <int NF>
struct Pad {      // a padded line entry: primitives, gamma, fractions
    real p[NF];
    real g;
    real q[2];
};

struct Line {     // one grid line along a direction through a cell
    int n;        // cells along it
    long base;    // cell index of the line's first cell
    long stride;  // index step along the line
};

struct Ctx {
    const real* S;
    const real* q[2];
    long N;
    int nq, reactive;
    real gamma;
    const real* P = nullptr;  // primitives then gamma per cell (k_metric_prim), for load_prim
};

template <int NF>
__device__ __forceinline__ void to_prim(const real* u, real g, real* p) {
    const real rho = u[0];
    p[0] = rho;
    real s2 = 0;
#pragma unroll
    for (int k = 1; k < NF - 1; ++k) {
        p[k] = u[k] / rho;
        s2 = k == 1 ? u[k] * u[k] : s2 + u[k] * u[k];
    }
    p[NF - 1] = (g - (real)1) * (u[NF - 1] - ((real)0.5 * s2) / rho);
}

template <int NF>
__device__ __forceinline__ void to_cons(const real* p, real g, real* u) {
    const real rho = p[0];
    u[0] = rho;
    real s2 = 0;
#pragma unroll
    for (int k = 1; k < NF - 1; ++k) {
        u[k] = p[k] * rho;
        s2 = k == 1 ? p[k] * p[k] : s2 + p[k] * p[k];
    }
    u[NF - 1] = p[NF - 1] / (g - (real)1) + ((real)0.5 * rho) * s2;
}

template <int NF>
__device__ __forceinline__ Pad<NF> load_cell(const Ctx& c, long i) {
    Pad<NF> e;
    real u[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) u[k] = c.S[k * c.N + i];
    e.g = c.reactive ? gamma_of(c.q[0][i]) : c.gamma;
    to_prim<NF>(u, e.g, e.p);
    e.q[0] = c.nq > 0 ? c.q[0][i] : (real)0;
    e.q[1] = c.nq > 1 ? c.q[1][i] : (real)0;
    return e;
}

// The same entry from the precomputed primitives (c.P).
template <int NF>
__device__ __forceinline__ Pad<NF> load_prim(const Ctx& c, long i) {
    Pad<NF> e;
#pragma unroll
    for (int k = 0; k < NF; ++k) e.p[k] = c.P[k * c.N + i];
    e.g = c.P[NF * c.N + i];
    e.q[0] = c.nq > 0 ? c.q[0][i] : (real)0;
    e.q[1] = c.nq > 1 ? c.q[1][i] : (real)0;
    return e;
}

// Primitives and gamma of every cell, once per state, for the stencil
// kernels (each cell is an entry of up to fifteen padded stencils).
template <int NF>
__global__ void __launch_bounds__(256) k_metric_prim(Ctx c, real* out) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= c.N) return;
    const Pad<NF> e = load_cell<NF>(c, i);
#pragma unroll
    for (int k = 0; k < NF; ++k) out[k * c.N + i] = e.p[k];
    out[NF * c.N + i] = e.g;
}

// Boundary-plane state from the edge cell (torchBoundaries.py:177-246);
// side -1 low, +1 high. Errors the reference raises set flag bits.
template <int NF, int D>
__device__ __forceinline__ void ghost(Pad<NF>& e, const MFace& f, int side, int nq, int reactive, int& bad) {
    constexpr int nrm = D + 1, E = NF - 1;
    if (f.type == FACE_WALL) {
        e.p[nrm] = -e.p[nrm];
    } else if (f.type == FACE_INFLOW) {
        if (reactive) e.g = gamma_of((real)f.lam);
        if (nq > 0) e.q[0] = (real)f.lam;
        if (nq > 1) e.q[1] = (real)f.beta;
#pragma unroll
        for (int k = 0; k < NF; ++k) e.p[k] = (real)f.prim[k];
    } else if (f.type == FACE_STAGNATION) {
        const real gin = e.g;
        const real g = reactive ? gamma_of((real)f.lam) : gin;
        const real a_in = sqrt((gin * e.p[E]) / e.p[0]);
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
#pragma unroll
        for (int m = 1; m < NF - 1; ++m) e.p[m] = 0;
        e.p[nrm] = (real)(-side) * un;
        e.p[E] = (real)f.total_pressure * pow((a * a) / a0sq, g / (g - (real)1));
        e.p[0] = (g * e.p[E]) / (a * a);
        e.g = g;
        if (nq > 0) e.q[0] = (real)f.lam;
        if (nq > 1) e.q[1] = (real)f.beta;
    } else if (f.type == FACE_PRESSURE_OUTLET) {
        const real g = e.g;
        const real a = sqrt((g * e.p[E]) / e.p[0]);
        const real un = (real)side * e.p[nrm];
        const bool sub = un < a;
        const real P = (real)f.pressure;
        const real rho = e.p[0] * pow(P / e.p[E], (real)1 / g);
        const real a_new = sqrt((g * P) / rho);
        const real vel = un + ((real)2 * (a - a_new)) / (g - (real)1);
        const bool back = sub && (vel < (real)0 || un < (real)0);
        if (sub) {
            e.p[0] = rho;
            e.p[E] = P;
            e.p[nrm] = (real)side * vel;
        }
        if (back) {
            if (!f.has_backflow) {
                bad |= fvl::F_BC_BACKFLOW;
            } else {
#pragma unroll
                for (int k = 0; k < NF; ++k) e.p[k] = (real)f.prim[k];
                if (reactive) e.g = gamma_of((real)f.lam);
                if (nq > 0) e.q[0] = (real)f.lam;
                if (nq > 1) e.q[1] = (real)f.beta;
            }
        }
    }
}

// Padded entry jj in [0, n + 1] of a line (torchFiniteVolume.py:33-63).
template <int NF, int D>
__device__ __forceinline__ Pad<NF> padded(const Ctx& c, const fvl::Metric& m, const Line& L, int jj, int& bad) {
    if (jj >= 1 && jj <= L.n) return load_prim<NF>(c, L.base + (jj - 1) * L.stride);
    const bool low = jj == 0;
    const MFace& f = m.face[2 * D + (low ? 0 : 1)];
    if (f.type == FACE_PERIODIC) return load_prim<NF>(c, L.base + (low ? L.n - 1 : 0) * L.stride);
    Pad<NF> e = load_prim<NF>(c, L.base + (low ? 0 : L.n - 1) * L.stride);
    ghost<NF, D>(e, f, low ? -1 : 1, c.nq, c.reactive, bad);
    return e;
}

// Limited slope of entry b between a and z, zeroed where half of it would
// take density or pressure to <= 1e-12 (torchFiniteVolume.py:143-156).
template <int NF, int LIM>
__device__ __forceinline__ void slope3(const Pad<NF>& a, const Pad<NF>& b, const Pad<NF>& z, real* s) {
#pragma unroll
    for (int k = 0; k < NF; ++k) s[k] = limited<LIM, 3>(b.p[k] - a.p[k], z.p[k] - b.p[k]);
    if (b.p[0] - (real)0.5 * fabs(s[0]) <= (real)1e-12 || b.p[NF - 1] - (real)0.5 * fabs(s[NF - 1]) <= (real)1e-12) {
#pragma unroll
        for (int k = 0; k < NF; ++k) s[k] = s[k] * (real)0;
    }
}

// Rusanov flux of primitive face states along D (torchFlux.py:80-100).
template <int NF, int D>
__device__ __forceinline__ void rusanov(const real* l, const real* r, real gl, real gr, real* F) {
    constexpr int n = D + 1, E = NF - 1;
    real ul[NF], ur[NF], fl[NF], fr[NF];
    to_cons<NF>(l, gl, ul);
    to_cons<NF>(r, gr, ur);
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        fl[k] = ul[k] * l[n];
        fr[k] = ur[k] * r[n];
    }
    fl[n] = fl[n] + l[E];
    fr[n] = fr[n] + r[E];
    fl[E] = fl[E] + l[n] * l[E];
    fr[E] = fr[E] + r[n] * r[E];
    const real s = nmax(fabs(l[n]) + sqrt((gl * l[E]) / l[0]), fabs(r[n]) + sqrt((gr * r[E]) / r[0]));
#pragma unroll
    for (int k = 0; k < NF; ++k) F[k] = (real)0.5 * ((fl[k] + fr[k]) - s * (ur[k] - ul[k]));
}

// Face flux from the entries either side and their slopes, with the wall
// mirror at a boundary face (torchFiniteVolume.py:163-188).
template <int NF, int D>
__device__ __forceinline__ void face(const Pad<NF>& a, const real* sa, const Pad<NF>& b, const real* sb,
                                     bool wall_lo, bool wall_hi, real* F) {
    real lft[NF], rgt[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        lft[k] = a.p[k] + (real)0.5 * sa[k];
        rgt[k] = b.p[k] - (real)0.5 * sb[k];
    }
    real gl = a.g, gr = b.g;
    if (wall_lo) {
#pragma unroll
        for (int k = 0; k < NF; ++k) lft[k] = rgt[k];
        gl = gr;
        lft[D + 1] = -lft[D + 1];
    }
    if (wall_hi) {
#pragma unroll
        for (int k = 0; k < NF; ++k) rgt[k] = lft[k];
        gr = gl;
        rgt[D + 1] = -rgt[D + 1];
    }
    rusanov<NF, D>(lft, rgt, gl, gr, F);
}

__device__ __forceinline__ double product(const double* const* f, int dim, const int* idx) {
    double v = 1.0;
    for (int a = 0; a < dim; ++a) v = v * f[a][idx[a]];
    return v;
}

template <int DIM>
__device__ __forceinline__ double volume_at(const fvl::Metric& m, const int* idx) {
    double v = 1.0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) v = v * m.vf[a][idx[a]];
    return v;
}

// Area of direction D's face `at` at the cell's other coordinates.
template <int DIM, int D>
__device__ __forceinline__ double area_at(const fvl::Metric& m, const int* idx, int at) {
    double v = 1.0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) v = v * m.af[D][a][a == D ? at : idx[a]];
    return v;
}

// Geometric momentum sources (torchGeometry.py:184-213) from the cell's
// primitives and its area divergences.
template <int DIM>
__device__ __forceinline__ void geometric_source(int geometry, const real* p, double rad, double cot, real* out) {
    constexpr int NF = DIM + 2;
#pragma unroll
    for (int k = 0; k < NF; ++k) out[k] = 0;
    if (geometry == GEO_CARTESIAN) return;
    const real rho = p[0], pr = p[NF - 1];
    const real radial = (real)rad;
    out[1] = pr * radial;
    if (geometry == GEO_SPHERICAL) {
        const real inv_r = radial / (real)2;
        real s2 = 0;
#pragma unroll
        for (int k = 2; k < NF - 1; ++k) s2 = k == 2 ? p[k] * p[k] : s2 + p[k] * p[k];
        out[1] = out[1] + (rho * s2) * inv_r;
        if (DIM >= 2) {
            const real cotr = (real)cot;
            out[2] = pr * cotr - ((rho * p[1]) * p[2]) * inv_r;
            if (DIM == 3) {
                out[2] = out[2] + (rho * (p[3] * p[3])) * cotr;
                out[3] = ((-rho) * p[3]) * (p[1] * inv_r + p[2] * cotr);
            }
        }
    } else if (DIM >= 2 && !(geometry == GEO_CYLINDRICAL && DIM == 2)) {
        out[1] = out[1] + (rho * (p[2] * p[2])) * radial;
        out[2] = (((-rho) * p[1]) * p[2]) * radial;
    }
}

// physical_state of one state (torchSolver.py:1038-1046).
template <int NF>
__device__ __forceinline__ bool physical_n(const real* u) {
    bool finite = true;
    real s2 = 0;
#pragma unroll
    for (int k = 0; k < NF; ++k) finite = finite & isfinite(u[k]);
#pragma unroll
    for (int k = 1; k < NF - 1; ++k) s2 = k == 1 ? u[k] * u[k] : s2 + u[k] * u[k];
    return finite & (u[0] > (real)1e-12) & (u[NF - 1] - ((real)0.5 * s2) / u[0] > (real)1e-12);
}

// Deterministic block sum of one slot (tree in LDS) into partials.
__device__ __forceinline__ void block_sum(double v, int slot, double* partials) {
    __shared__ double red[MT];
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int s = MT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[(long)blockIdx.x * fvl::NSLOT + slot] = red[0];
    __syncthreads();
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

template <int D>
__device__ __forceinline__ Line line_of(const fvl::Grid& g, long i, const int* idx) {
    Line L;
    L.n = D == 0 ? g.nx : (D == 1 ? g.ny : g.nz);
    L.stride = D == 0 ? 1 : (D == 1 ? (long)g.nx : (long)g.nx * g.ny);
    L.base = i - (long)idx[D] * L.stride;
    return L;
}

// Stage output of cell i from its change k and species change kq (SSPRK2,
// torchFiniteVolume.py:232-276 and torchAdaptive.py:169-206): mode 0 stores
// the change, mode 1 S + dt k, mode 2 ((U0 + S) + dt k) / 2, fractions from
// the held partial densities; flags nonphysical states and fractions
// outside [0, 1] and stores the clipped fractions.
template <int NF>
__device__ __forceinline__ void stage_out(const fvl::MStage& s, long N, long i, const real* k, const real* kq,
                                          int& bad) {
    real out[NF], qo[2] = {0, 0};
    const real* S = (const real*)s.S;
    if (s.mode == 0) {
#pragma unroll
        for (int f = 0; f < NF; ++f) out[f] = k[f];
#pragma unroll
        for (int j = 0; j < 2; ++j) qo[j] = kq[j];
    } else {
        const real* U0 = (const real*)s.U0;
        const real dt = (real)(s.dtp ? *s.dtp : s.dt);
#pragma unroll
        for (int f = 0; f < NF; ++f)
            out[f] = s.mode == 1 ? S[f * N + i] + dt * k[f]
                                 : (real)0.5 * ((U0[f * N + i] + S[f * N + i]) + dt * k[f]);
        if (!physical_n<NF>(out)) bad |= fvl::F_M_UNPHYS;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= s.nq) break;
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
#pragma unroll
    for (int f = 0; f < NF; ++f) ((real*)s.out)[f * N + i] = out[f];
#pragma unroll
    for (int j = 0; j < 2; ++j)
        if (j < s.nq) ((real*)s.qout[j])[i] = qo[j];
}

// Faces along direction D: their counts per axis and the face index of a
// cell's low (at) or high (at + 1) face.
template <int D>
__device__ __forceinline__ long face_count_of(const fvl::Grid& g) {
    return (long)(g.nx + (D == 0)) * (g.ny + (D == 1)) * (g.nz + (D == 2));
}

template <int D>
__device__ __forceinline__ long face_index(const fvl::Grid& g, const int* idx, int at) {
    const int fx = D == 0 ? at : idx[0], fy = D == 1 ? at : idx[1], fz = D == 2 ? at : idx[2];
    return ((long)fz * (g.ny + (D == 1)) + fy) * (g.nx + (D == 0)) + fx;
}

// Limited slope of padded entry jj of a line (torchFiniteVolume.py:143-160):
// interior entries from their neighbours; the boundary entries take the
// slope of the entry they wrap to on a periodic axis, none otherwise.
template <int NF, int D, int LIM>
__device__ __forceinline__ void entry_slope(const Ctx& c, const fvl::Metric& m, const Line& L, int jj,
                                            const Pad<NF>& e, const Pad<NF>* prev, const Pad<NF>* next, real* s,
                                            int& bad) {
    const int n = L.n;
#pragma unroll
    for (int q = 0; q < NF; ++q) s[q] = 0;
    if (jj >= 1 && jj <= n) {
        const Pad<NF> a = prev ? *prev : padded<NF, D>(c, m, L, jj - 1, bad);
        const Pad<NF> z = next ? *next : padded<NF, D>(c, m, L, jj + 1, bad);
        slope3<NF, LIM>(a, e, z, s);
    } else if (m.face[2 * D].type == FACE_PERIODIC) {
        const int w = jj == 0 ? n : 1;
        slope3<NF, LIM>(padded<NF, D>(c, m, L, w - 1, bad), padded<NF, D>(c, m, L, w, bad),
                        padded<NF, D>(c, m, L, w + 1, bad), s);
    }
}

// One face along D (torchFiniteVolume.py:163-229): the padded entries either
// side (fa, fa + 1) with their limited slopes, the wall mirror on a wall
// face, the Rusanov flux times the face area, and the fractions it moves.
// Both cells of a face read the same stored values.
template <int DIM, int D, int LIM>
__global__ void __launch_bounds__(MT) k_metric_faces(fvl::Grid g, fvl::Metric m, fvl::MStage s) {
    constexpr int NF = DIM + 2;
    const long nfD = face_count_of<D>(g);
    const long f = blockIdx.x * (long)MT + threadIdx.x;
    int bad = 0;
    if (f < nfD) {
        Ctx c;
        c.S = (const real*)s.S;
        c.q[0] = (const real*)s.qS[0];
        c.q[1] = (const real*)s.qS[1];
        c.N = (long)g.nx * g.ny * g.nz;
        c.nq = s.nq;
        c.reactive = s.reactive;
        c.gamma = (real)s.gamma;
        c.P = (const real*)m.prim;
        const int wx = g.nx + (D == 0), wy = g.ny + (D == 1);
        int idx[3] = {(int)(f % wx), (int)((f / wx) % wy), (int)(f / ((long)wx * wy))};
        const int fa = idx[D];
        idx[D] = 0;
        const Line L = line_of<D>(g, (long)idx[0] + (long)g.nx * (idx[1] + (long)g.ny * idx[2]), idx);
        const int n = L.n;
        const Pad<NF> A = padded<NF, D>(c, m, L, fa, bad);
        const Pad<NF> B = padded<NF, D>(c, m, L, fa + 1, bad);
        real sa[NF], sb[NF];
#pragma unroll
        for (int q = 0; q < NF; ++q) sa[q] = sb[q] = 0;
        if (!s.first_order) {
            entry_slope<NF, D, LIM>(c, m, L, fa, A, nullptr, &B, sa, bad);
            entry_slope<NF, D, LIM>(c, m, L, fa + 1, B, &A, nullptr, sb, bad);
        }
        const bool wall_lo = fa == 0 && m.face[2 * D].type == FACE_WALL;
        const bool wall_hi = fa == n && m.face[2 * D + 1].type == FACE_WALL;
        real F[NF];
        face<NF, D>(A, sa, B, sb, wall_lo, wall_hi, F);
        const real area = (real)area_at<DIM, D>(m, idx, fa);
        real* W = (real*)m.wf[D];
#pragma unroll
        for (int v = 0; v < NF; ++v) W[v * nfD + f] = F[v] * area;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= s.nq) break;
            W[(NF + j) * nfD + f] = (F[0] * (F[0] >= (real)0 ? A.q[j] : B.q[j])) * area;
        }
    }
    if (bad) atomicOr(s.flags, bad);
}

// Change of one cell along D from its two stored faces: -(W_hi - W_lo) / V
// into k, the species change into kq, the boundary faces into out_sum.
template <int DIM, int D>
__device__ __forceinline__ void gather(const fvl::Metric& m, const fvl::Grid& g, const int* idx, int nq, real Vr,
                                       real* k, real* kq, double* out_sum) {
    constexpr int NF = DIM + 2;
    const long nfD = face_count_of<D>(g);
    const int at = idx[D], n = D == 0 ? g.nx : (D == 1 ? g.ny : g.nz);
    const long lo = face_index<D>(g, idx, at), hi = face_index<D>(g, idx, at + 1);
    const real* W = (const real*)m.wf[D];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const real wlo = W[f * nfD + lo], whi = W[f * nfD + hi];
        k[f] = k[f] - (whi - wlo) / Vr;
        if (at == n - 1) out_sum[f] += (double)whi;
        if (at == 0) out_sum[f] -= (double)wlo;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (j >= nq) break;
        kq[j] = kq[j] - (W[(NF + j) * nfD + hi] - W[(NF + j) * nfD + lo]) / Vr;
    }
}

// Cell update (torchFiniteVolume.py:125-229, 232-276): geometric source,
// the stored faces direction by direction, then the SSPRK2 stage.
template <int DIM>
__global__ void __launch_bounds__(MT) k_metric_update(fvl::Grid g, fvl::Metric m, fvl::MStage s) {
    constexpr int NF = DIM + 2;
    const long N = (long)g.nx * g.ny * g.nz;
    const long i = blockIdx.x * (long)MT + threadIdx.x;
    double sums[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // outward x5, integrated source x5
    int bad = 0;
    if (i < N) {
        Ctx c;
        c.N = N;
        c.nq = s.nq;
        c.q[0] = (const real*)s.qS[0];
        c.q[1] = (const real*)s.qS[1];
        c.P = (const real*)m.prim;
        int idx[3];
        cell_coords(g, i, idx);
        const double V = volume_at<DIM>(m, idx);
        const Pad<NF> me = load_prim<NF>(c, i);
        double rad = 0, cot = 0;
        if (m.geometry != GEO_CARTESIAN) {
            rad = (area_at<DIM, 0>(m, idx, idx[0] + 1) - area_at<DIM, 0>(m, idx, idx[0])) / V;
            if (DIM >= 2) cot = (area_at<DIM, (DIM >= 2 ? 1 : 0)>(m, idx, idx[1] + 1) -
                                 area_at<DIM, (DIM >= 2 ? 1 : 0)>(m, idx, idx[1])) / V;
        }
        real k[NF], kq[2] = {0, 0};
        geometric_source<DIM>(m.geometry, me.p, rad, cot, k);
#pragma unroll
        for (int f = 0; f < NF; ++f) sums[5 + f] = (double)k[f] * V;
        const real Vr = (real)V;
        gather<DIM, 0>(m, g, idx, s.nq, Vr, k, kq, sums);
        if (DIM >= 2) gather<DIM, (DIM >= 2 ? 1 : 0)>(m, g, idx, s.nq, Vr, k, kq, sums);
        if (DIM >= 3) gather<DIM, (DIM >= 3 ? 2 : 0)>(m, g, idx, s.nq, Vr, k, kq, sums);
        stage_out<NF>(s, N, i, k, kq, bad);
    }
    if (bad) atomicOr(s.flags, bad);
#pragma unroll
    for (int k = 0; k < 5; ++k) block_sum(sums[k], fvl::S_OUT + k, s.partials);
#pragma unroll
    for (int k = 0; k < 5; ++k) block_sum(sums[5 + k], S_MSRC + k, s.partials);
}

// Larger |u_D| + a of a face's two sides, boundary faces with the boundary
// plane's state (torchFiniteVolume.py:66-98).
template <int DIM, int D>
__device__ __forceinline__ real face_rate(const Ctx& c, const fvl::Metric& m, const fvl::Grid& g, long i,
                                          const int* idx, int& bad) {
    constexpr int NF = DIM + 2;
    const Line L = line_of<D>(g, i, idx);
    const int at = idx[D];
    auto speed = [&](int jj) {
        const Pad<NF> e = padded<NF, D>(c, m, L, jj, bad);
        return fabs(e.p[D + 1]) + sqrt((e.g * e.p[NF - 1]) / e.p[0]);
    };
    const real sc = speed(at + 1);
    const real slo = nmax(speed(at), sc), shi = nmax(sc, speed(at + 2));
    const real a_lo = (real)area_at<DIM, D>(m, idx, at) * slo;
    const real a_hi = (real)area_at<DIM, D>(m, idx, at + 1) * shi;
    return nmax(a_lo, a_hi);
}

#ifdef FV_FAST
// The exact change of face_rate, in double, when each entry's sound speed
// rises by its increment `inc` (per cell; a periodic ghost takes its source
// cell's, an inflow ghost none, any other ghost its edge cell's, exact for
// walls and extrapolation and a first-order estimate for stagnation inlets
// and pressure outlets). Zero when no entry's cell reacted.
template <int DIM, int D>
__device__ __forceinline__ double face_rate_change(const Ctx& c, const fvl::Metric& m, const fvl::Grid& g, long i,
                                                   const int* idx, int& bad, const real* inc) {
    constexpr int NF = DIM + 2;
    const Line L = line_of<D>(g, i, idx);
    const int at = idx[D];
    auto raise = [&](int jj) -> double {
        if (jj >= 1 && jj <= L.n) return (double)inc[L.base + (jj - 1) * L.stride];
        const bool low = jj == 0;
        const MFace& f = m.face[2 * D + (low ? 0 : 1)];
        if (f.type == FACE_INFLOW) return 0.0;
        if (f.type == FACE_PERIODIC) return (double)inc[L.base + (low ? L.n - 1 : 0) * L.stride];
        return (double)inc[L.base + (low ? 0 : L.n - 1) * L.stride];
    };
    auto speed = [&](int jj) {
        const Pad<NF> e = padded<NF, D>(c, m, L, jj, bad);
        return (double)(fabs(e.p[D + 1]) + sqrt((e.g * e.p[NF - 1]) / e.p[0]));
    };
    auto mx = [](double a, double b) { return (a > b || a != a) ? a : b; };
    const double s0 = speed(at), s1 = speed(at + 1), s2 = speed(at + 2);
    const double h0 = s0 + raise(at), h1 = s1 + raise(at + 1), h2 = s2 + raise(at + 2);
    const double alo = area_at<DIM, D>(m, idx, at), ahi = area_at<DIM, D>(m, idx, at + 1);
    return mx(alo * mx(h0, h1), ahi * mx(h1, h2)) - mx(alo * mx(s0, s1), ahi * mx(s1, s2));
}
#endif

// CFL rate: sum over directions of max(A_lo s_lo, A_hi s_hi) / V. With
// `inc` (an FP32 reactive attempt over the committed state): that rate
// raised by the heating's exact change (face_rate_change).
template <int DIM>
__global__ void __launch_bounds__(MT) k_metric_cfl(fvl::Grid g, fvl::Metric m, const real* U, const real* lam,
                                                   double gamma, int reactive, double* partials, int* flags,
                                                   const real* inc) {
    const long N = (long)g.nx * g.ny * g.nz;
    const long i = blockIdx.x * (long)MT + threadIdx.x;
    double rate = 0;
    int bad = 0;
    if (i < N) {
        Ctx c;
        c.S = U;
        c.q[0] = lam;
        c.q[1] = nullptr;
        c.N = N;
        c.nq = 0;
        c.reactive = reactive;
        c.gamma = (real)gamma;
        c.P = (const real*)m.prim;
        int idx[3];
        cell_coords(g, i, idx);
        const real V = (real)volume_at<DIM>(m, idx);
        real r = 0;
        r = r + face_rate<DIM, 0>(c, m, g, i, idx, bad) / V;
        if (DIM >= 2) r = r + face_rate<DIM, (DIM >= 2 ? 1 : 0)>(c, m, g, i, idx, bad) / V;
        if (DIM >= 3) r = r + face_rate<DIM, (DIM >= 3 ? 2 : 0)>(c, m, g, i, idx, bad) / V;
        rate = (double)r;
#ifdef FV_FAST
        if (inc) {
            double d = face_rate_change<DIM, 0>(c, m, g, i, idx, bad, inc);
            if (DIM >= 2) d += face_rate_change<DIM, (DIM >= 2 ? 1 : 0)>(c, m, g, i, idx, bad, inc);
            if (DIM >= 3) d += face_rate_change<DIM, (DIM >= 3 ? 2 : 0)>(c, m, g, i, idx, bad, inc);
            rate = rate + d / volume_at<DIM>(m, idx);
        }
#endif
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
    double v = 0;
    if (i < N) {
        int idx[3];
        cell_coords(g, i, idx);
        const double V = product(m.vf, dim, idx);
        const long o = (long)e * N + i;
        v = (double)((((old[o] - start[o]) + state[o]) - hydro[o]) * (real)V);
    }
    block_sum(v, S_MSRC + e, partials);
}

// Stand-alone Rusanov over face arrays (rusanov_flux_grid, torchFlux.py:103-122).
template <int NF, int D>
__global__ void k_rusanov_faces(long mfaces, const real* l, const real* r, const real* gl, const real* gr,
                                real* out) {
    const long f = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (f >= mfaces) return;
    real a[NF], b[NF], F[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
        a[k] = l[k * mfaces + f];
        b[k] = r[k * mfaces + f];
    }
    rusanov<NF, D>(a, b, gl[f], gr[f], F);
#pragma unroll
    for (int k = 0; k < NF; ++k) out[k * mfaces + f] = F[k];
}

// -- adaptive leaves ------------------------------------------------------------
// Dyadic leaves with an explicit face list (torchAdaptive.py:14-166). Face
// values and fluxes take one thread per face; gradients, updates and CFL
// rates one thread per leaf, which gathers its faces through the incidence
// list in the order the reference scatters them with np.add.at (faces where
// the leaf is left, then where it is right), so each sum is formed in the
// reference's order and no atomics are needed.

__device__ __forceinline__ real nmin(real a, real b) { return (a < b || a != a) ? a : b; }

// Boundary-plane state of a boundary face along runtime direction D.
template <int NF, int DIM>
__device__ __forceinline__ void ghost_along(Pad<NF>& e, const MFace& f, int side, int D, int nq, int reactive,
                                            int& bad) {
    if (D == 0) ghost<NF, 0>(e, f, side, nq, reactive, bad);
    else if (DIM >= 2 && D == 1) ghost<NF, (DIM >= 2 ? 1 : 0)>(e, f, side, nq, reactive, bad);
    else if (DIM >= 3) ghost<NF, (DIM >= 3 ? 2 : 0)>(e, f, side, nq, reactive, bad);
}

template <int NF, int DIM>
__device__ __forceinline__ void rusanov_along(const real* l, const real* r, real gl, real gr, int D, real* F) {
    if (D == 0) rusanov<NF, 0>(l, r, gl, gr, F);
    else if (DIM >= 2 && D == 1) rusanov<NF, (DIM >= 2 ? 1 : 0)>(l, r, gl, gr, F);
    else rusanov<NF, (DIM >= 3 ? 2 : 0)>(l, r, gl, gr, F);
}

// Raw face values (face_values without reconstruction, torchAdaptive.py:14-57):
// the primitives, gamma and fractions of the leaves either side, the
// boundary plane's state on the outer side of a boundary face.
template <int DIM>
__global__ void __launch_bounds__(MT) k_amr_faces(fvl::Amr a, Ctx c, int* flags) {
    constexpr int NF = DIM + 2;
    const long F = a.nface;
    const long f = blockIdx.x * (long)MT + threadIdx.x;
    int bad = 0;
    if (f < F) {
        const int side = a.side[f], D = a.dir[f];
        Pad<NF> L, R;
        if (side == -1) {
            R = load_cell<NF>(c, a.right[f]);
            L = R;
            ghost_along<NF, DIM>(L, a.face[2 * D], -1, D, c.nq, c.reactive, bad);
        } else if (side == 1) {
            L = load_cell<NF>(c, a.left[f]);
            R = L;
            ghost_along<NF, DIM>(R, a.face[2 * D + 1], 1, D, c.nq, c.reactive, bad);
        } else {
            L = load_cell<NF>(c, a.left[f]);
            R = load_cell<NF>(c, a.right[f]);
        }
        real* FL = (real*)a.FL;
        real* FR = (real*)a.FR;
#pragma unroll
        for (int v = 0; v < NF; ++v) {
            FL[v * F + f] = L.p[v];
            FR[v * F + f] = R.p[v];
        }
        real* G = (real*)a.G;
        G[f] = L.g;
        G[F + f] = R.g;
        real* Q = (real*)a.Q;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= c.nq) break;
            Q[(2 * j) * F + f] = L.q[j];
            Q[(2 * j + 1) * F + f] = R.q[j];
        }
    }
    if (bad) atomicOr(flags, bad);
}

// Bounded least-squares gradients of the primitives of one leaf
// (torchAdaptiveGeometry.py:334-372): the weighted displacement-difference
// vector through the inverse normal matrix, scaled per variable by the
// largest theta in [0, 1] that keeps every face value within the range of
// the leaf and its face neighbours.
template <int DIM>
__global__ void __launch_bounds__(MT) k_amr_gradient(fvl::Amr a, Ctx c) {
    constexpr int NF = DIM + 2;
    const long F = a.nface;
    const long n = blockIdx.x * (long)MT + threadIdx.x;
    if (n >= a.nleaf) return;
    const Pad<NF> me = load_cell<NF>(c, n);
    const real* FL = (const real*)a.FL;
    const real* FR = (const real*)a.FR;
    real vec[DIM][NF], lo[NF], hi[NF];
#pragma unroll
    for (int v = 0; v < NF; ++v) {
        lo[v] = hi[v] = me.p[v];
#pragma unroll
        for (int d = 0; d < DIM; ++d) vec[d][v] = 0;
    }
    const long e0 = a.start[n], e1 = a.start[n + 1];
    for (long e = e0; e < e1; ++e) {
        const long f = a.cface[e];
        const bool is_left = a.crole[e] < 0;
        const real w = (real)a.weight[f];
        real diff[NF];
#pragma unroll
        for (int v = 0; v < NF; ++v) {
            const real l = FL[v * F + f], r = FR[v * F + f];
            diff[v] = r - l;
            const real other = is_left ? r : l;
            lo[v] = nmin(lo[v], other);
            hi[v] = nmax(hi[v], other);
        }
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
            const real wd = w * (real)a.disp[f * DIM + d];
#pragma unroll
            for (int v = 0; v < NF; ++v) vec[d][v] = vec[d][v] + wd * diff[v];
        }
    }
    real grad[DIM][NF];
    const double* inv = a.inverse + n * DIM * DIM;
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int v = 0; v < NF; ++v) {
            real acc = (real)inv[i * DIM] * vec[0][v];
#pragma unroll
            for (int j = 1; j < DIM; ++j) acc = acc + (real)inv[i * DIM + j] * vec[j][v];
            grad[i][v] = acc;
        }
    real theta[NF];
#pragma unroll
    for (int v = 0; v < NF; ++v) theta[v] = 1;
    for (long e = e0; e < e1; ++e) {
        const long f = a.cface[e];
        const double* off = (a.crole[e] < 0 ? a.offl : a.offr) + f * DIM;
#pragma unroll
        for (int v = 0; v < NF; ++v) {
            real change = (real)off[0] * grad[0][v];
#pragma unroll
            for (int d = 1; d < DIM; ++d) change = change + (real)off[d] * grad[d][v];
            const real allowed = change >= (real)0 ? hi[v] - me.p[v] : lo[v] - me.p[v];
            real ratio = fabs(change) > (real)1e-14 ? allowed / change : (real)1;
            ratio = ratio < (real)0 ? (real)0 : (ratio > (real)1 ? (real)1 : ratio);
            theta[v] = nmin(theta[v], ratio);
        }
    }
    real* out = (real*)a.grad + n * DIM * NF;
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int v = 0; v < NF; ++v) out[d * NF + v] = grad[d][v] * theta[v];
}

// Face fluxes (torchAdaptive.py:59-77, 106-146, 149-166): reconstructed
// face values (when second order), the wall mirror after reconstruction,
// the Rusanov flux weighted by the face area, the fractions it moves, the
// area-weighted wave speed for the CFL rate, and the boundary faces'
// outward sums.
template <int DIM>
__global__ void __launch_bounds__(MT) k_amr_flux(fvl::Amr a, int nq, int reconstruct, double* partials) {
    constexpr int NF = DIM + 2, E = NF - 1;
    const long F = a.nface;
    const long f = blockIdx.x * (long)MT + threadIdx.x;
    double out[NF];
#pragma unroll
    for (int v = 0; v < NF; ++v) out[v] = 0;
    if (f < F) {
        const int D = a.dir[f], side = a.side[f], li = a.left[f], ri = a.right[f];
        const real* FL = (const real*)a.FL;
        const real* FR = (const real*)a.FR;
        const real* G = (const real*)a.G;
        real l[NF], r[NF];
#pragma unroll
        for (int v = 0; v < NF; ++v) {
            l[v] = FL[v * F + f];
            r[v] = FR[v * F + f];
        }
        real gl = G[f], gr = G[F + f];
        if (reconstruct) {
            const real* grad = (const real*)a.grad;
            if (li >= 0) {
                const real* g = grad + (long)li * DIM * NF;
                const double* off = a.offl + f * DIM;
#pragma unroll
                for (int v = 0; v < NF; ++v) {
                    real change = (real)off[0] * g[v];
#pragma unroll
                    for (int d = 1; d < DIM; ++d) change = change + (real)off[d] * g[d * NF + v];
                    l[v] = l[v] + change;
                }
            }
            if (ri >= 0) {
                const real* g = grad + (long)ri * DIM * NF;
                const double* off = a.offr + f * DIM;
#pragma unroll
                for (int v = 0; v < NF; ++v) {
                    real change = (real)off[0] * g[v];
#pragma unroll
                    for (int d = 1; d < DIM; ++d) change = change + (real)off[d] * g[d * NF + v];
                    r[v] = r[v] + change;
                }
            }
            if (side == -1 && a.face[2 * D].type == FACE_WALL) {
#pragma unroll
                for (int v = 0; v < NF; ++v) l[v] = r[v];
                l[D + 1] = -l[D + 1];
                gl = gr;
            } else if (side == 1 && a.face[2 * D + 1].type == FACE_WALL) {
#pragma unroll
                for (int v = 0; v < NF; ++v) r[v] = l[v];
                r[D + 1] = -r[D + 1];
                gr = gl;
            }
        }
        real flux[NF];
        rusanov_along<NF, DIM>(l, r, gl, gr, D, flux);
        const real A = (real)a.area[f];
        real* W = (real*)a.W;
#pragma unroll
        for (int v = 0; v < NF; ++v) {
            const real w = flux[v] * A;
            W[v * F + f] = w;
            if (side != 0) out[v] = side > 0 ? (double)w : -(double)w;
        }
        const real* Q = (const real*)a.Q;
        real* M = (real*)a.M;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            if (j >= nq) break;
            M[j * F + f] = (flux[0] * A) * (flux[0] >= (real)0 ? Q[(2 * j) * F + f] : Q[(2 * j + 1) * F + f]);
        }
        const real s = nmax(fabs(l[D + 1]) + sqrt((gl * l[E]) / l[0]), fabs(r[D + 1]) + sqrt((gr * r[E]) / r[0]));
        ((real*)a.S)[f] = s * A;
    }
#pragma unroll
    for (int v = 0; v < NF; ++v) block_sum(out[v], fvl::S_OUT + v, partials);
}

// Leaf update (torchAdaptive.py:80-146, 169-206): geometric source times
// volume plus the signed weighted fluxes of the leaf's faces, over the
// volume; then the stage output.
template <int DIM>
__global__ void __launch_bounds__(MT) k_amr_update(fvl::Amr a, fvl::MStage s) {
    constexpr int NF = DIM + 2;
    const long N = a.nleaf, F = a.nface;
    const long n = blockIdx.x * (long)MT + threadIdx.x;
    double sums[NF];
#pragma unroll
    for (int v = 0; v < NF; ++v) sums[v] = 0;
    int bad = 0;
    if (n < N) {
        Ctx c;
        c.S = (const real*)s.S;
        c.q[0] = (const real*)s.qS[0];
        c.q[1] = (const real*)s.qS[1];
        c.N = N;
        c.nq = s.nq;
        c.reactive = s.reactive;
        c.gamma = (real)s.gamma;
        const Pad<NF> me = load_cell<NF>(c, n);
        const double V = a.volume[n];
        const real Vr = (real)V;
        real src[NF], chg[NF], kq[2] = {0, 0};
        geometric_source<DIM>(a.geometry, me.p, a.div[n], DIM >= 2 ? a.div[N + n] : 0.0, src);
#pragma unroll
        for (int v = 0; v < NF; ++v) {
            chg[v] = src[v] * Vr;
            sums[v] = (double)src[v] * V;
        }
        const real* W = (const real*)a.W;
        const real* M = (const real*)a.M;
        for (long e = a.start[n]; e < a.start[n + 1]; ++e) {
            const long f = a.cface[e];
            const bool is_left = a.crole[e] < 0;
#pragma unroll
            for (int v = 0; v < NF; ++v) {
                const real w = W[v * F + f];
                chg[v] = chg[v] + (is_left ? -w : w);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (j >= s.nq) break;
                const real m = M[j * F + f];
                kq[j] = kq[j] + (is_left ? -m : m);
            }
        }
        real k[NF];
#pragma unroll
        for (int v = 0; v < NF; ++v) k[v] = chg[v] / Vr;
#pragma unroll
        for (int j = 0; j < 2; ++j) kq[j] = kq[j] / Vr;
        stage_out<NF>(s, N, n, k, kq, bad);
    }
    if (bad) atomicOr(s.flags, bad);
#pragma unroll
    for (int v = 0; v < NF; ++v) block_sum(sums[v], S_MSRC + v, s.partials);
}

// CFL rate of one leaf: the area-weighted face speeds summed direction by
// direction (faces where the leaf is left, then right), 0.5 rate / volume.
template <int DIM>
__global__ void __launch_bounds__(MT) k_amr_cfl(fvl::Amr a, double* partials) {
    const long n = blockIdx.x * (long)MT + threadIdx.x;
    double v = 0;
    if (n < a.nleaf) {
        const real* S = (const real*)a.S;
        real rate = 0;
        const long e0 = a.start[n], e1 = a.start[n + 1];
#pragma unroll
        for (int d = 0; d < DIM; ++d)
            for (long e = e0; e < e1; ++e) {
                const long f = a.cface[e];
                if (a.dir[f] == d) rate = rate + S[f];
            }
        v = (double)(((real)0.5 * rate) / (real)a.volume[n]);
    }
    block_max(v, partials);
}

// Reaction source integral of the energy row over the leaves:
// sum of (((old - start) + state) - hydro) * volume (torchAdaptive.py:464-467).
__global__ void __launch_bounds__(MT) k_amr_source(fvl::Amr a, int e, const real* start, const real* old,
                                                   const real* state, const real* hydro, double* partials) {
    const long N = a.nleaf;
    const long n = blockIdx.x * (long)MT + threadIdx.x;
    double v = 0;
    if (n < N) {
        const long o = (long)e * N + n;
        v = (double)((((old[o] - start[o]) + state[o]) - hydro[o]) * (real)a.volume[n]);
    }
    block_sum(v, S_MSRC + e, partials);
}

template <int DIM>
hipError_t amr_eval_dim(const fvl::Amr& a, const Ctx& c, int reconstruct, int* flags, double* partials,
                        hipStream_t st) {
    const int fb = (int)((a.nface + MT - 1) / MT), lb = (int)((a.nleaf + MT - 1) / MT);
    k_amr_faces<DIM><<<fb, MT, 0, st>>>(a, c, flags);
    if (reconstruct) k_amr_gradient<DIM><<<lb, MT, 0, st>>>(a, c);
    k_amr_flux<DIM><<<fb, MT, 0, st>>>(a, c.nq, reconstruct, partials);
    return hipGetLastError();
}

template <int DIM, int D, int LIM>
void launch_faces(const fvl::Grid& g, const fvl::Metric& m, const fvl::MStage& s, hipStream_t st) {
    const long nf = (long)(g.nx + (D == 0)) * (g.ny + (D == 1)) * (g.nz + (D == 2));
    k_metric_faces<DIM, D, LIM><<<(int)((nf + MT - 1) / MT), MT, 0, st>>>(g, m, s);
}

template <int DIM, int LIM>
void launch_all_faces(const fvl::Grid& g, const fvl::Metric& m, const fvl::MStage& s, hipStream_t st) {
    launch_faces<DIM, 0, LIM>(g, m, s, st);
    if (DIM >= 2) launch_faces<DIM, (DIM >= 2 ? 1 : 0), LIM>(g, m, s, st);
    if (DIM >= 3) launch_faces<DIM, (DIM >= 3 ? 2 : 0), LIM>(g, m, s, st);
}

template <int DIM>
hipError_t launch_stage(const fvl::Grid& g, const fvl::Metric& m, const fvl::MStage& s, int blocks,
                        hipStream_t st) {
    switch (s.limiter) {
        case VANLEER: launch_all_faces<DIM, VANLEER>(g, m, s, st); break;
        case ALBADA: launch_all_faces<DIM, ALBADA>(g, m, s, st); break;
        case MINBEE: launch_all_faces<DIM, MINBEE>(g, m, s, st); break;
        case SUPERBEE: launch_all_faces<DIM, SUPERBEE>(g, m, s, st); break;
        default: launch_all_faces<DIM, UNOS>(g, m, s, st); break;
    }
    k_metric_update<DIM><<<blocks, MT, 0, st>>>(g, m, s);
    return hipGetLastError();
}

}  // namespace

int metric_blocks(const fvl::Grid& g) {
    const long N = (long)g.nx * g.ny * g.nz;
    return (int)((N + MT - 1) / MT);
}

hipError_t metric_prims(const fvl::Grid& g, const fvl::Metric& m, const void* S, const void* q0, const void* q1,
                        int nq, double gamma, int reactive, hipStream_t st) {
    Ctx c;
    c.S = (const real*)S;
    c.q[0] = (const real*)q0;
    c.q[1] = (const real*)q1;
    c.N = (long)g.nx * g.ny * g.nz;
    c.nq = nq;
    c.reactive = reactive;
    c.gamma = (real)gamma;
    const int blocks = (int)((c.N + 255) / 256);
    real* out = (real*)m.prim;
    if (g.dim == 1) k_metric_prim<3><<<blocks, 256, 0, st>>>(c, out);
    else if (g.dim == 2) k_metric_prim<4><<<blocks, 256, 0, st>>>(c, out);
    else k_metric_prim<5><<<blocks, 256, 0, st>>>(c, out);
    return hipGetLastError();
}

hipError_t metric_stage(const fvl::Grid& g, const fvl::Metric& m, const fvl::MStage& s, hipStream_t st) {
    const int blocks = metric_blocks(g);
    hipError_t e = metric_prims(g, m, s.S, s.qS[0], s.qS[1], s.nq, s.gamma, s.reactive, st);
    if (e != hipSuccess) return e;
    if (g.dim == 1) return launch_stage<1>(g, m, s, blocks, st);
    if (g.dim == 2) return launch_stage<2>(g, m, s, blocks, st);
    return launch_stage<3>(g, m, s, blocks, st);
}

hipError_t metric_cfl(const fvl::Grid& g, const fvl::Metric& m, const void* U, const void* lam, double gamma,
                      int reactive, double* partials, int* flags, const void* increments, hipStream_t st) {
#ifndef FV_FAST
    if (increments) return hipErrorInvalidValue;
#endif
    const int blocks = metric_blocks(g);
    const real* u = (const real*)U;
    const real* l = (const real*)lam;
    const real* inc = (const real*)increments;
    hipError_t e = metric_prims(g, m, U, lam, nullptr, 0, gamma, reactive, st);
    if (e != hipSuccess) return e;
    if (g.dim == 1) k_metric_cfl<1><<<blocks, MT, 0, st>>>(g, m, u, l, gamma, reactive, partials, flags, inc);
    else if (g.dim == 2) k_metric_cfl<2><<<blocks, MT, 0, st>>>(g, m, u, l, gamma, reactive, partials, flags, inc);
    else k_metric_cfl<3><<<blocks, MT, 0, st>>>(g, m, u, l, gamma, reactive, partials, flags, inc);
    return hipGetLastError();
}

hipError_t metric_source(const fvl::Grid& g, const fvl::Metric& m, const void* start, const void* old,
                         const void* state, const void* hydro, double* partials, hipStream_t st) {
    k_metric_source<<<metric_blocks(g), MT, 0, st>>>(g, m, (const real*)start, (const real*)old, (const real*)state,
                                                     (const real*)hydro, partials);
    return hipGetLastError();
}

int amr_blocks(long n) { return (int)((n + MT - 1) / MT); }

hipError_t amr_eval(int dim, const fvl::Amr& a, const void* S, const void* q0, const void* q1, int nq, double gamma,
                    int reactive, int reconstruct, int* flags, double* partials, hipStream_t st) {
    if (a.nface <= 0 || a.nleaf <= 0) return hipSuccess;
    Ctx c;
    c.S = (const real*)S;
    c.q[0] = (const real*)q0;
    c.q[1] = (const real*)q1;
    c.N = a.nleaf;
    c.nq = nq;
    c.reactive = reactive;
    c.gamma = (real)gamma;
    if (dim == 1) return amr_eval_dim<1>(a, c, reconstruct, flags, partials, st);
    if (dim == 2) return amr_eval_dim<2>(a, c, reconstruct, flags, partials, st);
    return amr_eval_dim<3>(a, c, reconstruct, flags, partials, st);
}

hipError_t amr_update(int dim, const fvl::Amr& a, const fvl::MStage& s, hipStream_t st) {
    const int lb = amr_blocks(a.nleaf);
    if (dim == 1) k_amr_update<1><<<lb, MT, 0, st>>>(a, s);
    else if (dim == 2) k_amr_update<2><<<lb, MT, 0, st>>>(a, s);
    else k_amr_update<3><<<lb, MT, 0, st>>>(a, s);
    return hipGetLastError();
}

hipError_t amr_cfl(int dim, const fvl::Amr& a, double* partials, hipStream_t st) {
    const int lb = amr_blocks(a.nleaf);
    if (dim == 1) k_amr_cfl<1><<<lb, MT, 0, st>>>(a, partials);
    else if (dim == 2) k_amr_cfl<2><<<lb, MT, 0, st>>>(a, partials);
    else k_amr_cfl<3><<<lb, MT, 0, st>>>(a, partials);
    return hipGetLastError();
}

hipError_t amr_source(int dim, const fvl::Amr& a, const void* start, const void* old, const void* state,
                      const void* hydro, double* partials, hipStream_t st) {
    k_amr_source<<<amr_blocks(a.nleaf), MT, 0, st>>>(a, dim + 1, (const real*)start, (const real*)old,
                                                     (const real*)state, (const real*)hydro, partials);
    return hipGetLastError();
}

hipError_t rusanov_faces(int nv, int dir, long mfaces, const void* l, const void* r, const void* gl, const void* gr,
                         void* out, hipStream_t st) {
    if (mfaces <= 0) return hipSuccess;
    const int blocks = (int)((mfaces + 255) / 256);
    const real *a = (const real*)l, *b = (const real*)r, *ga = (const real*)gl, *gb = (const real*)gr;
    real* o = (real*)out;
#define RUSANOV_CASE(NF, D) \
    if (nv == NF && dir == D) k_rusanov_faces<NF, D><<<blocks, 256, 0, st>>>(mfaces, a, b, ga, gb, o);
    RUSANOV_CASE(3, 0)
    RUSANOV_CASE(4, 0)
    RUSANOV_CASE(4, 1)
    RUSANOV_CASE(5, 0)
    RUSANOV_CASE(5, 1)
    RUSANOV_CASE(5, 2)
#undef RUSANOV_CASE
    return hipGetLastError();
}

}  // namespace FV_NS
