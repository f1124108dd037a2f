// Launch interface between the C-ABI layer (capi.cpp) and the kernel
// translation units. Every kernel TU is compiled twice, into namespace
// fv64 (double storage and arithmetic) and fv32 (float storage, float bulk
// arithmetic, double reductions and kinetics). Pointers are untyped here:
// they address `double` arrays in fv64 and `float` arrays in fv32.
#pragma once

#include <hip/hip_runtime.h>

namespace fvl {

// Reduction slots written per block and folded by finalize().
enum Slot {
    S_MAX0 = 0,      // max(|u|+a) along x (3D/2D) or the 1D wave bound
    S_MAX1 = 1,
    S_MAX2 = 2,
    S_OUT = 3,       // 5 outward boundary-flux sums (weighted by dt*vol/h)
    S_SRC = 8,       // source integral of the energy row (reactive)
    S_TOTAL = 9,     // 5 conserved totals (optional)
    NSLOT = 16
};

// Device flag words (int), set with atomicOr.
enum Flag {
    F_PRED_BAD = 1,      // a predicted face state failed physical_state
    F_UPD_BAD = 2,       // second-order update unphysical
    F_LLF_BAD = 4,       // first-order update unphysical -> RuntimeError
    F_SPEC_BAD = 8,      // species left [-tol, 1+tol] (second-order fluxes; tol kSpecTol)
    F_SPEC_BAD_LLF = 16, // same with first-order fluxes
    F_REACT_BAD = 32,    // chemistry input outside its domain -> ValueError
    F_STATE_BAD = 64,    // state after chemistry unphysical
    // metric/face step (fvmetric.hip)
    F_M_UNPHYS = 128,    // an SSPRK2 stage left physical_state
    F_M_SPEC = 256,      // an SSPRK2 stage's fractions left [-1e-12, 1+1e-12]
    F_BC_RESERVOIR = 512,   // stagnation inlet: no subsonic reservoir solution
    F_BC_SUPPLY = 1024,     // stagnation inlet: reservoir cannot supply inflow
    F_BC_BACKFLOW = 2048    // pressure outlet backflow without a backflow state
};

struct Grid {
    int dim, nx, ny, nz;
    int bc[3];         // x, y, z boundary kinds (fv::Boundary)
    double h[3];       // spacing x, y, z
    double inlet[4];   // 2D Inflow conservative state (rho, mx, my, E)
    int wavespeed;     // 1D wave-speed estimate: 0 pressure based, 1 enthalpy based
};

// Face boundary of the metric/face scheme (torchBoundaries.py:8-15, 177-246).
enum FaceType { FACE_OUTFLOW = 0, FACE_WALL, FACE_PERIODIC, FACE_INFLOW, FACE_PRESSURE_OUTLET, FACE_STAGNATION };
struct MFace {
    int type, has_backflow;
    double prim[5];    // inflow primitive, or the backflow primitive (dim+2 used)
    double pressure;   // pressure outlet
    double total_pressure, total_temperature, gas_constant;  // stagnation inlet
    double lam, beta;  // boundary fractions
};
// Metric of a logically uniform grid: cell volumes and face areas are outer
// products of per-axis factors (torchGeometry.py:113-182), evaluated on the
// fly as ((1 * f_x) * f_y) * f_z.
enum Geometry { GEO_CARTESIAN = 0, GEO_CYLINDRICAL, GEO_ANNULAR, GEO_SPHERICAL };
struct Metric {
    const double* vf[3];      // volume factor per axis (n_a entries)
    const double* af[3][3];   // area factors of direction d over axis a (n_a, or n_d + 1 when a == d)
    MFace face[6];            // x_min, x_max, y_min, y_max, z_min, z_max
    int geometry;
    void* prim;               // (dim+3) x ncell scratch: primitives, then gamma, of the state in use
    void* wf[3];              // per direction: (dim+4) x nface weighted fluxes, then moved fractions
};
// One SSPRK2 stage of the metric scheme (torchFiniteVolume.py:232-276).
// mode 0 writes the semi-discrete change k (and species change) itself;
// mode 1 writes S + dt k with fractions (U0_rho q0 + dt kq) / rho_out;
// mode 2 writes 0.5 ((U0 + S) + dt k) with (0.5 ((U0_rho q0 + S_rho qS) + dt kq)) / rho_out.
struct MStage {
    const void* S;            // rhs input state
    const void* qS[2];        // its fractions (lambda, beta) or null
    const void* U0;           // step input state
    const void* q0[2];
    void* out;
    void* qout[2];
    double dt, gamma;
    const double* dtp;        // dt in device memory, used instead of dt when set
    int mode, first_order, limiter, nq, reactive;
    int* flags;
    double* partials;         // NSLOT per block: S_OUT.. outward, S_MSRC.. integrated source
};
enum { S_MSRC = 8 };          // 5 integrated geometric-source sums (metric stages)
// Adaptive (dyadic AMR) leaves as flat arrays (torchAdaptiveGeometry.LeafMesh):
// faces with their left/right leaves (-1 outside the domain), direction,
// boundary side, area and centre offsets; leaves with volume, area
// divergence (dim x nleaf) and least-squares inverse matrix (nleaf x dim x
// dim); per-leaf face incidence as CSR in the order the reference
// accumulates (faces where the leaf is left, then where it is right; role
// -1 / +1). Scratch arrays hold one evaluation's intermediate values.
struct Amr {
    long nleaf, nface;
    const int* left;
    const int* right;
    const int* dir;
    const int* side;
    const double* area;
    const double* offl;    // nface x dim
    const double* offr;    // nface x dim
    const double* disp;    // nface x dim
    const double* weight;  // nface
    const double* volume;  // nleaf
    const double* div;     // dim x nleaf
    const double* inverse; // nleaf x dim x dim
    const long* start;     // nleaf + 1
    const int* cface;
    const int* crole;
    MFace face[6];
    int geometry;
    void* FL;    // nf x nface face values (left side)
    void* FR;    // nf x nface (right side)
    void* G;     // 2 x nface: gamma left, right
    void* Q;     // 4 x nface: fraction 0 left/right, fraction 1 left/right
    void* grad;  // nleaf x dim x nf
    void* W;     // nf x nface area-weighted flux
    void* M;     // 2 x nface transported fractions
    void* S;     // nface area-weighted face speed (CFL)
};


struct Step {
    const void* U;     // input state, (dim+2) fields of ncell values
    // Energy row of the input state at Ue + (dim+1) * ncell, or null when it
    // is U's own (a reactive step's first half-step rewrites only that row).
    const void* Ue;
    void* Uout;        // output state
    const void* lam;   // per-cell lambda for gamma(lambda), or null
    int lam_is_gamma;  // lam holds gamma itself (functional calls with a gamma array)
    double gamma;      // scalar gamma when lam is null
    const void* q[2];  // transported fractions (lambda, beta) or null
    void* qout[2];
    int nq;
    double dt;
    double c[3];       // dt / h
    double hc[3];      // 0.5 * dt / h
    double w[3];       // dt * volume / h (boundary integral weights)
    void* F[3];        // optional face-flux dumps (reference layout) or null
    int* flags;
    double* partials;  // NSLOT doubles per block
    int zchunk;
    int limiter;
    int want_cfl;      // compute wave maxima of the output (inert runs)
    // 1D options
    int controller, variant, wavespeed;
    double cfl;
    // When set, the step coefficients come from this device array instead
    // of c/hc/w: [dt/h x3, 0.5*dt/h x3, dt*vol/h x3, hydro gate, ...]
    // (enum CoefSlot).
    const double* coef;
    // 1D CFL pass of an FP32-form reactive attempt: the chemistry's energy
    // and fraction increments per cell (two rows of n values), or null.
    const void* heat;
};

// Step coefficients: from the host (Step.c/hc/w) or, for graph-replayed
// batches, from the device array Step.coef written by the batch post kernel.
struct Coef {
    double c[3], hc[3], w[3];
    bool active;
};
__device__ __forceinline__ Coef coef_of(const Step& s) {
    Coef k;
    if (s.coef) {
        for (int d = 0; d < 3; ++d) {
            k.c[d] = s.coef[d];
            k.hc[d] = s.coef[3 + d];
            k.w[d] = s.coef[6 + d];
        }
        k.active = s.coef[9] != 0.0;
    } else {
        for (int d = 0; d < 3; ++d) {
            k.c[d] = s.c[d];
            k.hc[d] = s.hc[d];
            k.w[d] = s.w[d];
        }
        k.active = true;
    }
    return k;
}

struct React {
    const void* U;     // input state
    void* Uout;
    const void* lam;
    void* lamout;
    const void* beta;  // null for Arrhenius without induction array
    void* betaout;
    const void* Ustart;  // for the source integral (final half-step), or null
    const void* Uold;
    int model;           // 0 arrhenius, 1 znd
    double dt;
    double q_heat, activation_energy, pre_exponential;
    int* flags;
    double* partials;
    int want_cfl;
    // Write only the energy row of Uout (its other rows are U's, unchanged
    // by the chemistry): Uout may then be U itself, an in-place update.
    int energy_only;
    // A first half-step attempt, whose maxima only decide whether dt is
    // accepted (the FP32 form then reports the committed cell's speeds plus
    // the chemistry's exact speed increment; see react_cells_dim).
    int attempt;
    // Attempts of the FP32 form outside 3D Cartesian sessions, for the CFL
    // pass over the committed state: each cell's sound-speed increment
    // (INC_SPEED, one row of ncell values: 2D Cartesian and metric grids) or
    // its energy and fraction increments (INC_HEAT, two rows: 1D Cartesian).
    void* increments;
    int increment_kind;
};
enum IncrementKind { INC_SPEED = 1, INC_HEAT = 2 };

// One instance of a batch (device memory). Partial slots of an instance:
// [second-order | first-order | CFL pass | first chemistry half | second].
struct BatchItem {
    Step step;
    Grid grid;
    int tiles[3];      // 3D: tiles in x, y and z chunks; 2D: strips in x, row chunks
    int llf_first;     // first slot of the first-order pass
    int cfl_first;     // first slot of the CFL pass
    void* scratch;     // 1D work arrays
    // Reactive instances (3D): chemistry half-steps around the hydro step
    // and the four rotating buffers of state, lambda and beta.
    int reactive, znd;
    React r1, r2;
    int r1_first, r2_first;
    void* slot[3][4];  // [U, lambda, beta][buffer]
};

// Step coefficients of a batch instance (BatchCtl.coef): dt/h x3,
// 0.5*dt/h x3, dt*vol/h x3, then the hydro-pass gate, the chemistry
// half-step duration dt/2 and the chemistry gate.
enum CoefSlot { K_HYDRO = 9, K_HALF = 10, K_CHEM = 11, NCOEF = 12 };

// Pointers of a reactive instance for the step whose committed buffer is
// `cur` (A), over three rotating buffers: R(dt/2) writes the energy row and
// fractions of B (density and momenta stay A's: the chemistry does not
// change them), hydro reads {A's density and momenta, B's energy} and
// writes C with species, R(dt/2) updates C's energy row and fractions in
// place with the source integral against A and B (torchSession.py:398-515).
// C becomes the next step's A.
constexpr int kReactSlots = 3;
__host__ __device__ inline void set_reactive_slots(BatchItem& it, int cur) {
    const int A = cur, B = (cur + 1) % kReactSlots, C = (cur + 2) % kReactSlots;
    void* const* U = it.slot[0];
    void* const* L = it.slot[1];
    void* const* Bt = it.slot[2];
    it.r1.U = U[A];
    it.r1.Uout = U[B];
    it.r1.energy_only = 1;
    it.r1.lam = L[A];
    it.r1.lamout = L[B];
    it.r1.beta = it.znd ? Bt[A] : nullptr;
    it.r1.betaout = it.znd ? Bt[B] : nullptr;
    it.r1.Ustart = nullptr;
    it.r1.Uold = nullptr;
    it.r1.attempt = 1;
    it.step.U = U[A];
    it.step.Ue = U[B];
    it.step.Uout = U[C];
    it.step.lam = L[B];
    it.step.q[0] = L[B];
    it.step.qout[0] = L[C];
    it.step.q[1] = it.znd ? Bt[B] : nullptr;
    it.step.qout[1] = it.znd ? Bt[C] : nullptr;
    it.r2.U = U[C];
    it.r2.Uout = U[C];
    it.r2.energy_only = 1;
    it.r2.lam = L[C];
    it.r2.lamout = L[C];
    it.r2.beta = it.znd ? Bt[C] : nullptr;
    it.r2.betaout = it.znd ? Bt[C] : nullptr;
    it.r2.Ustart = U[A];
    it.r2.Uold = U[B];
    it.r2.attempt = 0;
}

// Per-instance control record of a batch (device memory). The host fills
// the constants at admission; the batch kernels advance the rest on the
// device, as Simulation.step does on the host (torchSession.py:374-522).
struct BatchCtl {
    double t_end, volume, cfl;
    double h[3];
    double time, dt;     // current time; dt of the step (or attempt) in flight
    double maxima[3];    // wave maxima of the committed state
    double outward[5];   // accumulated boundary integral
    double source[5];    // accumulated source integral (reactive)
    double coef[NCOEF];  // Step.coef of the step in flight (enum Coef)
    int dim, max_steps;
    int nblk, nllf, ncfl, nreact;  // partial slots of the passes
    int steps, fallbacks;
    int retry;           // failed chemistry attempts of the step in flight
    int retries;         // all failed attempts so far
    int status;          // 0, or the TORCH_* code that stopped the instance
    int active;          // 1 stepping, 2 waiting for its first dt, 0 stopped
    int cur;             // buffer holding the committed state
    int reactive;
    int ticket;          // post blocks done with their slot (0 between launches)
    double folded[9];    // slot totals of the step being committed
};

// Blocks of the post kernel per instance: one per folded partial slot.
constexpr int kPostSlots = 9;

// Instance of block b in a work list: first[i] <= b < first[i+1].
__device__ __forceinline__ int owner(const int* first, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Work-list kernels read their item through the constant address space:
// items change only between launches (host uploads, the post kernel), and
// constant loads may be issued again where needed instead of holding every
// field of the item in registers for the whole kernel.
__device__ __forceinline__ const BatchItem& item_at(const BatchItem* items, int i) {
    typedef const __attribute__((address_space(4))) BatchItem CItem;
    return *(const BatchItem*)((const CItem*)items + i);
}


struct Launch {
    int block_x, block_y;  // tile shape chosen by the autotuner
    hipStream_t stream;
};

}  // namespace fvl

#define FVL_DECLARE(NS)                                                                   \
    namespace NS {                                                                        \
    int step3d_blocks(const fvl::Grid& g, const fvl::Launch& l, int zchunk);              \
    hipError_t step3d(const fvl::Grid& g, const fvl::Step& s, const fvl::Launch& l);      \
    hipError_t llf3d(const fvl::Grid& g, const fvl::Step& s, hipStream_t st);             \
    int llf3d_blocks(const fvl::Grid& g);                                                 \
    hipError_t step2d(const fvl::Grid& g, const fvl::Step& s, const fvl::Launch& l);      \
    int step2d_blocks(const fvl::Grid& g, const fvl::Launch& l);                          \
    hipError_t llf2d(const fvl::Grid& g, const fvl::Step& s, hipStream_t st);             \
    int llf2d_blocks(const fvl::Grid& g);                                                 \
    hipError_t cfl2d(const fvl::Grid& g, const void* U, const void* lam, double gamma,    \
                     double* partials, int lam_is_gamma, const void* increments,           \
                     hipStream_t st);                                                     \
    int cfl2d_blocks(const fvl::Grid& g);                                                 \
    hipError_t cfl1d(const fvl::Grid& g, const fvl::Step& s, hipStream_t st);             \
    int cfl1d_blocks(const fvl::Grid& g);                                                 \
    hipError_t step1d(const fvl::Grid& g, const fvl::Step& s, void* scratch, hipStream_t st); \
    size_t step1d_scratch(const fvl::Grid& g);                                            \
    int step1d_blocks(const fvl::Grid& g);                                                \
    hipError_t cfl_cells(const fvl::Grid& g, const void* U, const void* lam, double gamma,\
                         double* partials, int lam_is_gamma, const void* increments,      \
                         hipStream_t st);                                                 \
    int cfl_blocks(const fvl::Grid& g);                                                   \
    hipError_t react(const fvl::Grid& g, const fvl::React& r, hipStream_t st);            \
    int react_blocks(const fvl::Grid& g);                                                 \
    hipError_t hllc_faces(int nv, int dir, long m, const void* l, const void* r,          \
                          const void* gl, const void* gr, void* out, hipStream_t st);     \
    hipError_t totals(const fvl::Grid& g, const void* U, double* partials, hipStream_t st);\
    hipError_t convert_in(const double* src, void* dst, long n, hipStream_t st);          \
    hipError_t convert_out(const void* src, double* dst, long n, hipStream_t st);         \
    /* batches: items/first in device memory; first[n] = total blocks */                 \
    hipError_t batch3d(const fvl::BatchItem* items, const int* first, int n, int blocks,   \
                       int bx, int by, int limiter, int nq, hipStream_t st);               \
    hipError_t batch_llf3d(const fvl::BatchItem* items, const int* first, int n,          \
                           int blocks, hipStream_t st);                                    \
    hipError_t batch_cfl3d(const fvl::BatchItem* items, const int* first, int n,          \
                           int blocks, hipStream_t st);                                    \
    hipError_t batch2d(const fvl::BatchItem* items, const int* first, int n, int blocks,   \
                       int limiter, hipStream_t st);                                       \
    hipError_t batch_llf2d(const fvl::BatchItem* items, const int* first, int n,          \
                           int blocks, hipStream_t st);                                    \
    hipError_t batch_cfl2d(const fvl::BatchItem* items, const int* first, int n,          \
                           int blocks, hipStream_t st);                                    \
    hipError_t batch1d(const fvl::BatchItem* items, const int* first, int n, int blocks,   \
                       int limiter, hipStream_t st);                                       \
    hipError_t batch_llf1d(const fvl::BatchItem* items, const int* first, int n,          \
                           int blocks, hipStream_t st);                                    \
    hipError_t batch_cfl1d(const fvl::BatchItem* items, const int* first, int n,          \
                           int blocks, int input, hipStream_t st);                         \
    hipError_t batch_post(fvl::BatchItem* items, fvl::BatchCtl* ctl, int n, int prime,    \
                          hipStream_t st);                                                 \
    hipError_t batch_react(const fvl::BatchItem* items, const int* first, int n,          \
                           int blocks, int phase, hipStream_t st);                         \
    hipError_t batch_check(fvl::BatchItem* items, fvl::BatchCtl* ctl, int n, hipStream_t st); \
    int strip2d_rows(const fvl::Grid& g);                                                 \
    int step1d_tiles(const fvl::Grid& g);                                                 \
    /* metric/face scheme */                                                              \
    hipError_t metric_stage(const fvl::Grid& g, const fvl::Metric& m, const fvl::MStage& s, \
                            hipStream_t st);                                              \
    hipError_t metric_cfl(const fvl::Grid& g, const fvl::Metric& m, const void* U,         \
                          const void* lam, double gamma, int reactive, double* partials,   \
                          int* flags, const void* increments, hipStream_t st);             \
    hipError_t metric_source(const fvl::Grid& g, const fvl::Metric& m, const void* start,  \
                             const void* old, const void* state, const void* hydro,        \
                             double* partials, hipStream_t st);                            \
    hipError_t rusanov_faces(int nv, int dir, long m, const void* l, const void* r,        \
                             const void* gl, const void* gr, void* out, hipStream_t st);   \
    int metric_blocks(const fvl::Grid& g);                                                \
    /* adaptive leaves (fvmetric.hip) */                                                  \
    hipError_t amr_eval(int dim, const fvl::Amr& a, const void* S, const void* q0,        \
                        const void* q1, int nq, double gamma, int reactive,               \
                        int reconstruct, int* flags, double* partials, hipStream_t st);   \
    hipError_t amr_update(int dim, const fvl::Amr& a, const fvl::MStage& s, hipStream_t st); \
    hipError_t amr_cfl(int dim, const fvl::Amr& a, double* partials, hipStream_t st);      \
    hipError_t amr_source(int dim, const fvl::Amr& a, const void* start, const void* old,  \
                          const void* state, const void* hydro, double* partials,          \
                          hipStream_t st);                                                 \
    int amr_blocks(long n);                                                               \
    }

FVL_DECLARE(fv64)
FVL_DECLARE(fv32)

namespace fvl {
// Fold per-block partials: maxima are max-reduced, sums added in block order
// (deterministic). Writes NSLOT doubles to out.
hipError_t finalize(const double* partials, int nblocks, double* out, hipStream_t st);
// The same, skipped on the device unless (*flags & mask) is set.
hipError_t finalize_if(const double* partials, int nblocks, double* out, const int* flags, int mask,
                       hipStream_t st);
}
