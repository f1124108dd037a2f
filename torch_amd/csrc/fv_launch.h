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
    F_SPEC_BAD = 8,      // species left [-1e-10, 1+1e-10] (second-order fluxes)
    F_SPEC_BAD_LLF = 16, // same with first-order fluxes
    F_REACT_BAD = 32,    // chemistry input outside its domain -> ValueError
    F_STATE_BAD = 64     // state after chemistry unphysical
};

struct Grid {
    int dim, nx, ny, nz;
    int bc[3];         // x, y, z boundary kinds (fv::Boundary)
    double h[3];       // spacing x, y, z
    double inlet[4];   // 2D Inflow conservative state (rho, mx, my, E)
    int wavespeed;     // 1D wave-speed estimate: 0 pressure based, 1 enthalpy based
};

struct Step {
    const void* U;     // input state, (dim+2) fields of ncell values
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
};

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
};

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
                     double* partials, int lam_is_gamma, hipStream_t st);                 \
    int cfl2d_blocks(const fvl::Grid& g);                                                 \
    hipError_t cfl1d(const fvl::Grid& g, const fvl::Step& s, hipStream_t st);             \
    int cfl1d_blocks(const fvl::Grid& g);                                                 \
    hipError_t step1d(const fvl::Grid& g, const fvl::Step& s, void* scratch, hipStream_t st); \
    size_t step1d_scratch(const fvl::Grid& g);                                            \
    int step1d_blocks(const fvl::Grid& g);                                                \
    hipError_t cfl_cells(const fvl::Grid& g, const void* U, const void* lam, double gamma,\
                         double* partials, int lam_is_gamma, hipStream_t st);            \
    int cfl_blocks(const fvl::Grid& g);                                                   \
    hipError_t react(const fvl::Grid& g, const fvl::React& r, hipStream_t st);            \
    int react_blocks(const fvl::Grid& g);                                                 \
    hipError_t hllc_faces(int nv, int dir, long m, const void* l, const void* r,          \
                          const void* gl, const void* gr, void* out, hipStream_t st);     \
    hipError_t totals(const fvl::Grid& g, const void* U, double* partials, hipStream_t st);\
    hipError_t convert_in(const double* src, void* dst, long n, hipStream_t st);          \
    hipError_t convert_out(const void* src, double* dst, long n, hipStream_t st);         \
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
