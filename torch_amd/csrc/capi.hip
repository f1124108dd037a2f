// C ABI (include/torch_hip.h): device contexts, device-resident session
// states and the Strang-split step driver of Simulation.step
// (python/cfdSolver/torchSession.py:374-522), dispatching to the fv64 or
// fv32 kernel builds.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "capi_common.h"

namespace tcapi {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

int alloc(void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes ? bytes : 16);
    if (e != hipSuccess) return fail(TORCH_HIP_ERROR, std::string("hipMalloc: ") + hipGetErrorString(e));
    return TORCH_OK;
}

int check_config(const torch_config& c) {
    if (c.dim < 1 || c.dim > 3) return fail(TORCH_INVALID, "Invalid dimension");
    for (int d = 0; d < c.dim; ++d) {
        if (c.n[d] < 2) return fail(TORCH_INVALID, "Require cells >= 2");
        if (!(c.spacing[d] > 0) && (!(c.lengths[d] > 0) || !isfinite(c.lengths[d])))
            return fail(TORCH_INVALID, "Require length > 0");
        if (c.bc[d] < 0 || c.bc[d] > 3 || (c.bc[d] == TORCH_INFLOW && !(c.dim == 2 && d == 1)))
            return fail(TORCH_INVALID, "Unsupported session boundary");
    }
    if (c.limiter < 0 || c.limiter > 4) return fail(TORCH_INVALID, "Unknown limiter or chemistry model");
    if (c.kinetic_model < 0 || c.kinetic_model > 1) return fail(TORCH_INVALID, "Unknown limiter or chemistry model");
    if (!(c.gamma > 1) || !isfinite(c.gamma)) return fail(TORCH_INVALID, "gamma must be finite and greater than one");
    if (!(c.cfl > 0) || !isfinite(c.cfl)) return fail(TORCH_INVALID, "CFL must be positive");
    if (c.precision != TORCH_FP64 && c.precision != TORCH_FP32) return fail(TORCH_INVALID, "unknown precision");
    if (c.gamma_mode < 0 || c.gamma_mode > 2 || (c.reactive && c.gamma_mode != 1))
        return fail(TORCH_INVALID, "gamma_mode must be 1 for reactive states, 0 or 2 otherwise");
    // A chosen 3D tile must be one the step kernels are built for (the
    // 768-thread 32x24 one exists in the FP32 form only); 0 picks the default.
    if (c.dim == 3 && c.block_x != 0 && c.block_y != 0) {
        const bool ok = (c.block_x == 64 && (c.block_y == 4 || c.block_y == 8)) ||
                        (c.block_x == 32 && (c.block_y == 8 || c.block_y == 16 ||
                                             (c.block_y == 24 && c.precision == TORCH_FP32)));
        if (!ok) return fail(TORCH_INVALID, "unsupported tile shape");
    }
    return TORCH_OK;
}

void make_grid(const torch_config& c, fvl::Grid* out, double* volume) {
    fvl::Grid& g = *out;
    memset(&g, 0, sizeof(g));
    g.dim = c.dim;
    g.wavespeed = c.wavespeed;
    g.nx = c.n[0];
    g.ny = c.dim >= 2 ? c.n[1] : 1;
    g.nz = c.dim >= 3 ? c.n[2] : 1;
    for (int d = 0; d < 3; ++d) {
        g.bc[d] = d < c.dim ? c.bc[d] : TORCH_OPEN;
        g.h[d] = d < c.dim ? (c.spacing[d] > 0 ? c.spacing[d] : c.lengths[d] / (double)c.n[d]) : 1.0;
    }
    // Volume as np.prod(spacing): ((hx*hy)*hz) over the active axes.
    double v = g.h[0];
    if (c.dim >= 2) v = v * g.h[1];
    if (c.dim >= 3) v = v * g.h[2];
    *volume = v;
    if (c.dim == 2) {
        // Inlet conservative state at gamma 1.4 with the reference's floors
        // (torchSolver2D.py:61-82, 243).
        const double* p = c.inlet_primitive;
        double rho = fmax(p[0], 1e-12), pr = fmax(p[3], 1e-12);
        double ke = 0.5 * rho * (p[1] * p[1] + p[2] * p[2]);
        g.inlet[0] = rho;
        g.inlet[1] = rho * p[1];
        g.inlet[2] = rho * p[2];
        g.inlet[3] = pr / (1.4 - 1.0) + ke;
    }
}

// Pinned staging of FP32 transfers: an upload converts the float64 input
// into it (on up to 8 threads for large arrays) and one DMA moves it; a
// download is one DMA into it, then the widening. Grown on demand and kept
// for the process; every transfer completes before the lock is released,
// so the buffer is free again when the next one takes it.
namespace {
std::mutex g_stage_mu;
float* g_stage = nullptr;
long g_stage_n = 0;

// Elementwise conversion of n values on up to 8 threads (one per 2^20).
template <typename From, typename To>
void convert(const From* src, To* dst, long n) {
    const long per = 1L << 20;
    const int threads = (int)std::min<long>(8, (n + per - 1) / per);
    auto part = [&](int t) {
        const long lo = n * t / threads, hi = n * (t + 1) / threads;
        for (long i = lo; i < hi; ++i) dst[i] = (To)src[i];
    };
    if (threads <= 1) {
        part(0);
        return;
    }
    // A thread that cannot be started (a process limit) leaves its part to
    // this one: no exception crosses the C ABI.
    std::vector<std::thread> pool;
    int started = 1;
    try {
        for (; started < threads; ++started) pool.emplace_back(part, started);
    } catch (...) {
    }
    for (int t = started; t < threads; ++t) part(t);
    part(0);
    for (auto& th : pool) th.join();
}

// The staging buffer holds at least n floats (the caller holds g_stage_mu).
int stage_at_least(long n) {
    if (n <= g_stage_n) return TORCH_OK;
    if (g_stage) (void)hipHostFree(g_stage);
    g_stage = nullptr;
    g_stage_n = 0;
    HIP_TRY(hipHostMalloc((void**)&g_stage, (size_t)n * sizeof(float), hipHostMallocDefault));
    g_stage_n = n;
    return TORCH_OK;
}
}  // namespace

int upload_as(const Ops* ops, hipStream_t st, void* dst, const double* src, long n) {
    if (n <= 0) return TORCH_OK;
    if (ops->elem == sizeof(double)) {
        HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamSynchronize(st));
        return TORCH_OK;
    }
    std::lock_guard<std::mutex> lock(g_stage_mu);
    int rc = stage_at_least(n);
    if (rc) return rc;
    convert(src, g_stage, n);
    HIP_TRY(hipMemcpyAsync(dst, g_stage, n * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    return TORCH_OK;
}

int download_as(const Ops* ops, hipStream_t st, double* dst, const void* src, long n) {
    if (n <= 0) return TORCH_OK;
    if (ops->elem == sizeof(double)) {
        HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return TORCH_OK;
    }
    std::lock_guard<std::mutex> lock(g_stage_mu);
    int rc = stage_at_least(n);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(g_stage, src, n * sizeof(float), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    convert(g_stage, dst, n);
    return TORCH_OK;
}

#define OPS_OF(NS, T)                                                                                    \
    Ops{sizeof(T),        NS::step3d_blocks, NS::step3d,      NS::llf3d,        NS::llf3d_blocks,         \
        NS::step2d,       NS::step2d_blocks, NS::llf2d,       NS::llf2d_blocks,                           \
        NS::step1d,       NS::step1d_scratch, NS::step1d_blocks,                                         \
        NS::cfl_cells,    NS::cfl_blocks,    NS::react,       NS::react_blocks, NS::hllc_faces, NS::totals, \
        NS::convert_in,   NS::convert_out,                                                               \
        NS::batch3d,      NS::batch_llf3d,   NS::batch_cfl3d, NS::batch2d,      NS::batch_llf2d,          \
        NS::batch_cfl2d,  NS::batch1d,       NS::batch_llf1d, NS::batch_cfl1d,  NS::batch_post,           \
        NS::batch_react,  NS::batch_check,   NS::strip2d_rows, NS::step1d_tiles,                       \
        NS::metric_stage, NS::metric_cfl,    NS::metric_source, NS::rusanov_faces, NS::metric_blocks,      \
        NS::amr_eval,     NS::amr_update,    NS::amr_cfl,       NS::amr_source,   NS::amr_blocks}

// Not const: a const table would be promoted to device memory, where the
// host launchers it points to do not exist.
Ops kOps64 = OPS_OF(fv64, double);
Ops kOps32 = OPS_OF(fv32, float);

}  // namespace tcapi

using namespace tcapi;

namespace {

// Fold per-block partials in a fixed order (deterministic sums).
__global__ void k_finalize(const double* partials, int nblocks, double* out, const int* flags, int mask) {
    __shared__ double buf[256];
    if (flags && !(*flags & mask)) return;  // a pass that did not run left no partials
    const int slot = blockIdx.x;
    const bool is_max = slot < 3;
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        double p = partials[(long)b * fvl::NSLOT + slot];
        v = is_max ? fmax(v, p) : v + p;
    }
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            double a = buf[threadIdx.x], b = buf[threadIdx.x + s];
            buf[threadIdx.x] = is_max ? fmax(a, b) : a + b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[slot] = buf[0];
}

__global__ void k_clear_ints(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

hipError_t clear_flags(int* flags, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_ints, dim3(1), dim3(64), 0, st, flags, 2);
    return hipGetLastError();
}

}  // namespace

namespace fvl {
hipError_t finalize(const double* partials, int nblocks, double* out, hipStream_t st) {
    k_finalize<<<NSLOT, 256, 0, st>>>(partials, nblocks, out, nullptr, 0);
    return hipGetLastError();
}
hipError_t finalize_if(const double* partials, int nblocks, double* out, const int* flags, int mask,
                       hipStream_t st) {
    k_finalize<<<NSLOT, 256, 0, st>>>(partials, nblocks, out, flags, mask);
    return hipGetLastError();
}
}  // namespace fvl

// Device buffers of one session. Slots of 4 state buffers rotate so the
// committed state is never overwritten by a step that may still fail.
struct torch_state {
    torch_ctx* ctx;
    torch_config cfg;
    const Ops* ops;
    fvl::Grid grid;
    long ncell;
    int nf;
    void* U[4] = {nullptr, nullptr, nullptr, nullptr};
    void* lam[4] = {nullptr, nullptr, nullptr, nullptr};
    void* beta[4] = {nullptr, nullptr, nullptr, nullptr};
    int cur = 0;
    void* F[3] = {nullptr, nullptr, nullptr};
    void* scratch1d = nullptr;
    double* partials = nullptr;  // 4 regions: hydro, first-order, chemistry, CFL
    int pblocks = 0;
    double* scalars = nullptr;   // 4 x NSLOT finalized values
    int* flags = nullptr;
    double* host = nullptr;      // pinned readback: 4 x NSLOT doubles + flags
    bool maxima_valid = false;
    double maxima[3] = {0, 0, 0};
    // Bumped by every change of the committed state. A 1D or 2D chemistry
    // half-step of the FP32 form (torch_state_react) records the slot it
    // started from and the buffer of its per-cell increments; until the
    // next change, torch_state_wave_maxima reports that slot's maxima raised
    // by the heating's exact change (the FP64 method's view of the heated
    // state, whose sub-ulp heating the FP32 storage drops).
    long version = 0;
    int attempt_slot = -1;
    long attempt_version = -1;
    void* attempt_inc = nullptr;
    double volume = 0;
    bool profile = false, prof_pending = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double prof_ms = 0;
    long prof_n = 0;
    // metric/face scheme (torch_state_set_metric)
    bool has_metric = false;
    fvl::Metric metric;
    double* mfactors = nullptr;
    void* mprim = nullptr;       // primitives + gamma scratch of the metric kernels
    void* mfaces = nullptr;      // face fluxes of a metric stage
};

namespace {

long face_count(const fvl::Grid& g, int d) {
    long nx = g.nx + (d == 0), ny = g.ny + (d == 1), nz = g.nz + (d == 2);
    return nx * ny * nz;
}


fvl::Launch launch_of(const torch_state* s) {
    fvl::Launch l;
    l.block_x = s->cfg.block_x;
    l.block_y = s->cfg.block_y;
    l.stream = s->ctx->stream;
    return l;
}

// dt from the CFL maxima, in the reference's order of operations.
double dt_from_maxima(const torch_state* s, const double* mx) {
    const torch_config& c = s->cfg;
    if (c.dim == 3) {
        double sum = mx[0] / s->grid.h[0];
        sum = sum + mx[1] / s->grid.h[1];
        sum = sum + mx[2] / s->grid.h[2];
        return c.cfl / sum;
    }
    if (c.dim == 2) {
        // max((|u|+a)/dx) == max(|u|+a)/dx because rounding is monotone.
        double total = mx[0] / s->grid.h[0] + mx[1] / s->grid.h[1];
        if (total <= 1e-12) return 1e-3;
        return c.cfl / total;
    }
    // 1D: min(cfl*dx/max|S|, cfl*dx/max(|u|+a)), with the 1e-6 default.
    double a = mx[0] > 0 ? (c.cfl * s->grid.h[0]) / mx[0] : 1e-6;
    double b = c.cfl * s->grid.h[0] / mx[1];
    return b < a ? b : a;
}

double py_min(double a, double b) { return b < a ? b : a; }

int sync_read(torch_state* s) {
    HIP_TRY(hipMemcpyAsync(s->host, s->scalars, 4 * fvl::NSLOT * sizeof(double), hipMemcpyDeviceToHost,
                           s->ctx->stream));
    HIP_TRY(hipMemcpyAsync(s->host + 4 * fvl::NSLOT, s->flags, sizeof(int) * 2, hipMemcpyDeviceToHost,
                           s->ctx->stream));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    if (s->prof_pending) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        s->prof_ms += ms;
        s->prof_n += 1;
        s->prof_pending = false;
    }
    return TORCH_OK;
}

int host_flags(const torch_state* s) {
    int f;
    memcpy(&f, s->host + 4 * fvl::NSLOT, sizeof(int));
    return f;
}

fvl::Step base_step(torch_state* s, int in, int out, double dt, const void* lam) {
    fvl::Step st;
    memset(&st, 0, sizeof(st));
    st.U = s->U[in];
    st.Uout = s->U[out];
    st.lam = lam;
    st.gamma = s->cfg.gamma;
    st.dt = dt;
    for (int d = 0; d < 3; ++d) {
        st.c[d] = dt / s->grid.h[d];
        st.hc[d] = 0.5 * dt / s->grid.h[d];
        st.w[d] = dt * s->volume / s->grid.h[d];
    }
    st.flags = s->flags;
    st.zchunk = s->cfg.zchunk;
    st.limiter = s->cfg.limiter;
    st.controller = s->cfg.controller;
    st.variant = s->cfg.variant;
    st.wavespeed = s->cfg.wavespeed;
    st.cfl = s->cfg.cfl;
    return st;
}

double* region(torch_state* s, int r) { return s->partials + (long)r * s->pblocks * fvl::NSLOT; }

// Wave maxima of the committed state (Simulation._dt), into s->maxima.
int compute_maxima(torch_state* s) {
    const void* lam = s->cfg.gamma_mode ? s->lam[s->cur] : nullptr;
    hipStream_t st = s->ctx->stream;
    HIP_TRY(clear_flags(s->flags, st));
    HIP_TRY(s->ops->cfl_cells(s->grid, s->U[s->cur], lam, s->cfg.gamma, region(s, 3), s->cfg.gamma_mode == 2,
                              nullptr, st));
    HIP_TRY(fvl::finalize(region(s, 3), s->ops->cfl_blocks(s->grid), s->scalars + 3 * fvl::NSLOT, st));
    int rc = sync_read(s);
    if (rc) return rc;
    for (int m = 0; m < 3; ++m) s->maxima[m] = s->host[3 * fvl::NSLOT + m];
    s->maxima_valid = true;
    return TORCH_OK;
}

// 1D/2D CFL maxima of slot `u` (lambda of slot `l`) into region 3, raised
// by a reactive attempt's `increments` when given (FP32 form).
hipError_t run_cfl(torch_state* s, int u, int l, const void* increments = nullptr) {
    const void* lam = s->cfg.gamma_mode ? s->lam[l] : nullptr;
    hipStream_t st = s->ctx->stream;
    hipError_t e = s->ops->cfl_cells(s->grid, s->U[u], lam, s->cfg.gamma, region(s, 3), s->cfg.gamma_mode == 2,
                                     increments, st);
    if (e != hipSuccess) return e;
    return fvl::finalize(region(s, 3), s->ops->cfl_blocks(s->grid), s->scalars + 3 * fvl::NSLOT, st);
}

// Chemistry half-step slot `in` -> slot `out` (state, lambda, beta).
int run_react(torch_state* s, int in, int out, double dt, int start, int old, bool want_cfl,
              void* increments = nullptr, int increment_kind = fvl::INC_SPEED) {
    fvl::React r;
    memset(&r, 0, sizeof(r));
    r.U = s->U[in];
    r.Uout = s->U[out];
    r.lam = s->lam[in];
    r.lamout = s->lam[out];
    r.beta = s->cfg.kinetic_model == 1 ? s->beta[in] : nullptr;
    r.betaout = s->cfg.kinetic_model == 1 ? s->beta[out] : nullptr;
    r.Ustart = start >= 0 ? s->U[start] : nullptr;
    r.Uold = old >= 0 ? s->U[old] : nullptr;
    r.model = s->cfg.kinetic_model;
    r.dt = dt;
    r.q_heat = s->cfg.q_heat;
    r.activation_energy = s->cfg.activation_energy;
    r.pre_exponential = s->cfg.pre_exponential;
    r.flags = s->flags;
    r.partials = region(s, 2);
    r.want_cfl = want_cfl && s->cfg.dim == 3;
    r.attempt = want_cfl && start < 0;
    // A 2D or 1D attempt of the FP32 form leaves its per-cell increments in
    // state slot (in + 3) mod 4 (the second half-step's output, unused until
    // then) and checks the committed state's maxima raised by them; a metric
    // attempt or a standalone half-step passes its own buffer.
    if (r.attempt && s->cfg.dim != 3 && s->cfg.precision == TORCH_FP32) {
        r.increments = s->U[(in + 3) % 4];
        r.increment_kind = s->cfg.dim == 1 ? fvl::INC_HEAT : fvl::INC_SPEED;
    }
    if (increments) {
        r.increments = increments;
        r.increment_kind = increment_kind;
    }
    if (!r.U || !r.Uout || !r.lam || !r.lamout || (s->cfg.kinetic_model == 1 && (!r.beta || !r.betaout)) ||
        (start >= 0 && (!r.Ustart || !r.Uold)))
        return fail(TORCH_INVALID, "internal: chemistry buffers missing for this state");
    hipStream_t st = s->ctx->stream;
    HIP_TRY(s->ops->react(s->grid, r, st));
    HIP_TRY(fvl::finalize(region(s, 2), s->ops->react_blocks(s->grid), s->scalars + 2 * fvl::NSLOT, st));
    if (want_cfl && r.increments) HIP_TRY(run_cfl(s, in, in, r.increments));
    else if (want_cfl && s->cfg.dim != 3) HIP_TRY(run_cfl(s, out, out));
    return TORCH_OK;
}

// Hydro slot `in` -> `out` with optional species; leaves finalized hydro
// values in scalars[0] and first-order ones in scalars[1].
int run_hydro(torch_state* s, int in, int out, double dt, int lam_in, int lam_out, bool want_fluxes) {
    const bool reactive = s->cfg.reactive;
    const void* lam = s->cfg.gamma_mode ? s->lam[lam_in] : nullptr;
    fvl::Step st = base_step(s, in, out, dt, lam);
    if (!st.U || !st.Uout || (s->cfg.gamma_mode && !lam) || (s->cfg.dim == 1 && !s->scratch1d))
        return fail(TORCH_INVALID, "internal: hydro buffers missing for this state");
    if (reactive && (!s->lam[lam_in] || !s->lam[lam_out] || !s->beta[lam_in] || !s->beta[lam_out]))
        return fail(TORCH_INVALID, "internal: species buffers missing for this state");
    if (want_fluxes)
        for (int d = 0; d < s->cfg.dim; ++d)
            if (!s->F[d]) return fail(TORCH_INVALID, "internal: flux buffers missing");
    st.lam_is_gamma = s->cfg.gamma_mode == 2;
    if (reactive) {
        st.nq = s->cfg.kinetic_model == 1 ? 2 : 1;
        st.q[0] = s->lam[lam_in];
        st.qout[0] = s->lam[lam_out];
        st.q[1] = s->beta[lam_in];
        st.qout[1] = s->beta[lam_out];
    }
    st.want_cfl = !s->cfg.gamma_mode;
    if (want_fluxes)
        for (int d = 0; d < s->cfg.dim; ++d) st.F[d] = s->F[d];
    hipStream_t stream = s->ctx->stream;
    HIP_TRY(clear_flags(s->flags, stream));
    if (s->cfg.dim == 3) {
        fvl::Launch l = launch_of(s);
        st.partials = region(s, 0);
        if (s->profile) HIP_TRY(hipEventRecord(s->ev0, stream));
        HIP_TRY(s->ops->step3d(s->grid, st, l));
        if (s->profile) {
            HIP_TRY(hipEventRecord(s->ev1, stream));
            s->prof_pending = true;
        }
        HIP_TRY(fvl::finalize(region(s, 0), s->ops->step3d_blocks(s->grid, l, st.zchunk), s->scalars, stream));
        st.partials = region(s, 1);
        HIP_TRY(s->ops->llf3d(s->grid, st, stream));
        HIP_TRY(fvl::finalize_if(region(s, 1), s->ops->llf3d_blocks(s->grid), s->scalars + fvl::NSLOT, s->flags,
                                 fvl::F_PRED_BAD | fvl::F_UPD_BAD, stream));
    } else if (s->cfg.dim == 2) {
        fvl::Launch l = launch_of(s);
        st.partials = region(s, 0);
        HIP_TRY(s->ops->step2d(s->grid, st, l));
        HIP_TRY(fvl::finalize(region(s, 0), s->ops->step2d_blocks(s->grid, l), s->scalars, stream));
        st.partials = region(s, 1);
        HIP_TRY(s->ops->llf2d(s->grid, st, stream));
        HIP_TRY(fvl::finalize_if(region(s, 1), s->ops->llf2d_blocks(s->grid), s->scalars + fvl::NSLOT, s->flags,
                                 fvl::F_UPD_BAD, stream));
    } else {
        st.partials = region(s, 0);
        HIP_TRY(s->ops->step1d(s->grid, st, s->scratch1d, stream));
        HIP_TRY(fvl::finalize(region(s, 0), s->ops->step1d_blocks(s->grid), s->scalars, stream));
    }
    if (s->cfg.dim != 3 && !reactive) HIP_TRY(run_cfl(s, out, lam_in));
    return TORCH_OK;
}

// -- metric/face scheme -------------------------------------------------------

const int kMetricSlots = 4;

const void* frac(torch_state* s, int slot, int j) {
    if (!s->cfg.reactive) return nullptr;
    return j == 0 ? s->lam[slot] : (s->cfg.kinetic_model == 1 ? s->beta[slot] : nullptr);
}

int metric_nq(const torch_state* s) { return s->cfg.reactive ? (s->cfg.kinetic_model == 1 ? 2 : 1) : 0; }

// Errors the reference raises (ValueError) while building boundary planes.
int boundary_error(int flags) {
    if (flags & fvl::F_BC_RESERVOIR) return fail(TORCH_INVALID, "No subsonic reservoir-inlet solution");
    if (flags & fvl::F_BC_SUPPLY)
        return fail(TORCH_INVALID, "Reservoir cannot supply inflow against the interior pressure");
    if (flags & fvl::F_BC_BACKFLOW)
        return fail(TORCH_INVALID, "Pressure outlet backflow requires an explicit backflow primitive state");
    return TORCH_OK;
}

// cfl_dt of slot u (torchFiniteVolume.py:66-98).
// CFL dt of slot `u`; with `increments` (FP32 reactive attempts) the rates
// of `u`, the committed state, raised by the chemistry's sound-speed change.
int metric_dt(torch_state* s, int u, double* dt, const void* increments = nullptr) {
    hipStream_t st = s->ctx->stream;
    HIP_TRY(clear_flags(s->flags, st));
    HIP_TRY(s->ops->metric_cfl(s->grid, s->metric, s->U[u], frac(s, u, 0), s->cfg.gamma, s->cfg.reactive,
                               region(s, 3), s->flags, increments, st));
    HIP_TRY(fvl::finalize(region(s, 3), s->ops->metric_blocks(s->grid), s->scalars + 3 * fvl::NSLOT, st));
    int rc = sync_read(s);
    if (rc) return rc;
    if ((rc = boundary_error(host_flags(s)))) return rc;
    *dt = s->cfg.cfl / s->host[3 * fvl::NSLOT];
    return TORCH_OK;
}

// One SSPRK2 stage (mode 0: raw change) into slot `out`; returns the flags
// and leaves the outward / source sums in scalars[0].
int metric_stage(torch_state* s, int mode, int S, int U0, int out, double dt, int first_order, int* flags) {
    fvl::MStage m;
    memset(&m, 0, sizeof(m));
    m.S = s->U[S];
    m.U0 = s->U[U0];
    m.out = s->U[out];
    for (int j = 0; j < 2; ++j) {
        m.qS[j] = frac(s, S, j);
        m.q0[j] = frac(s, U0, j);
        m.qout[j] = const_cast<void*>(frac(s, out, j));
    }
    m.dt = dt;
    m.gamma = s->cfg.gamma;
    m.mode = mode;
    m.first_order = first_order;
    m.limiter = s->cfg.limiter;
    m.nq = metric_nq(s);
    m.reactive = s->cfg.reactive;
    m.flags = s->flags;
    m.partials = region(s, 0);
    hipStream_t st = s->ctx->stream;
    HIP_TRY(clear_flags(s->flags, st));
    HIP_TRY(s->ops->metric_stage(s->grid, s->metric, m, st));
    HIP_TRY(fvl::finalize(region(s, 0), s->ops->metric_blocks(s->grid), s->scalars, st));
    int rc = sync_read(s);
    if (rc) return rc;
    *flags = host_flags(s);
    return boundary_error(*flags);
}

// advance (torchFiniteVolume.py:232-276): slot `in` -> `out` through the
// stage slot. Returns TORCH_NONPHYSICAL when both orders lose positivity.
int metric_advance(torch_state* s, int in, int stage, int out, double dt, double* boundary, double* source,
                   int* first_used) {
    const int nf = s->nf;
    for (int first = 0; first < 2; ++first) {
        int flags = 0, rc;
        if ((rc = metric_stage(s, 1, in, in, stage, dt, first, &flags))) return rc;
        if (flags & (fvl::F_M_UNPHYS | fvl::F_M_SPEC)) continue;
        double b1[5], s1[5];
        for (int f = 0; f < nf; ++f) {
            b1[f] = s->host[fvl::S_OUT + f];
            s1[f] = s->host[fvl::S_MSRC + f];
        }
        if ((rc = metric_stage(s, 2, stage, in, out, dt, first, &flags))) return rc;
        if (flags & (fvl::F_M_UNPHYS | fvl::F_M_SPEC)) continue;
        for (int f = 0; f < nf; ++f) {
            boundary[f] = (0.5 * dt) * (b1[f] + s->host[fvl::S_OUT + f]);
            source[f] = (0.5 * dt) * (s1[f] + s->host[fvl::S_MSRC + f]);
        }
        *first_used = first;
        return TORCH_OK;
    }
    return fail(TORCH_NONPHYSICAL, "Metric/face update lost positivity; reduce CFL");
}

// dt of a metric step on the device: cfl / max rate clipped to t_end - time
// (torchSession.py:531-533), into *dt; the host checks it after the step.
__global__ void k_metric_dt(const double* max_rate, double cfl, double time, double t_end, double* dt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double d = cfl / max_rate[0];
        const double rest = t_end - time;
        dt[0] = rest < d ? rest : d;
    }
}

// The common inert metric step without host round trips: cfl, dt and both
// second-order stages are queued back to back and read once. Returns
// TORCH_OK with the step committed, kSlowPath when a stage lost positivity
// (nothing committed; the host-driven step with its retries takes over), or
// an error.
const int kSlowPath = -1;

int metric_fast_step(torch_state* s, double time, double t_end, torch_step_info* info) {
    hipStream_t st = s->ctx->stream;
    const int A = s->cur, B = (A + 1) % kMetricSlots, C = (A + 2) % kMetricSlots;
    double* rate = s->scalars + 3 * fvl::NSLOT;
    double* dtp = s->scalars + 3 * fvl::NSLOT + fvl::NSLOT - 1;
    HIP_TRY(clear_flags(s->flags, st));
    HIP_TRY(s->ops->metric_cfl(s->grid, s->metric, s->U[A], nullptr, s->cfg.gamma, 0, region(s, 3), s->flags,
                               nullptr, st));
    HIP_TRY(fvl::finalize(region(s, 3), s->ops->metric_blocks(s->grid), rate, st));
    k_metric_dt<<<1, 64, 0, st>>>(rate, s->cfg.cfl, time, t_end, dtp);
    HIP_TRY(hipGetLastError());
    for (int k = 0; k < 2; ++k) {
        fvl::MStage m;
        memset(&m, 0, sizeof(m));
        m.S = s->U[k == 0 ? A : B];
        m.U0 = s->U[A];
        m.out = s->U[k == 0 ? B : C];
        m.dtp = dtp;
        m.gamma = s->cfg.gamma;
        m.mode = k + 1;
        m.limiter = s->cfg.limiter;
        m.flags = s->flags + k;
        m.partials = region(s, k);
        HIP_TRY(s->ops->metric_stage(s->grid, s->metric, m, st));
        HIP_TRY(fvl::finalize(region(s, k), s->ops->metric_blocks(s->grid), s->scalars + k * fvl::NSLOT, st));
    }
    int rc = sync_read(s);
    if (rc) return rc;
    int flags[2];
    memcpy(flags, s->host + 4 * fvl::NSLOT, sizeof(flags));
    if ((rc = boundary_error(flags[0] | flags[1]))) return rc;
    const double dt = s->host[3 * fvl::NSLOT + fvl::NSLOT - 1];
    if (!isfinite(dt) || dt <= 0 || time + dt == time)
        return fail(TORCH_STALLED, "Integration stalled before requested final time");
    if ((flags[0] | flags[1]) & (fvl::F_M_UNPHYS | fvl::F_M_SPEC)) return kSlowPath;
    memset(info, 0, sizeof(*info));
    info->dt = dt;
    for (int f = 0; f < s->nf; ++f) {
        info->outward[f] = (0.5 * dt) * (s->host[fvl::S_OUT + f] + s->host[fvl::NSLOT + fvl::S_OUT + f]);
        info->source[f] = (0.5 * dt) * (s->host[fvl::S_MSRC + f] + s->host[fvl::NSLOT + fvl::S_MSRC + f]);
    }
    s->cur = C;
    s->maxima_valid = false;
    s->version += 1;
    return TORCH_OK;
}

int need_metric(const torch_state* s) {
    if (!s->has_metric) return fail(TORCH_INVALID, "state has no metric (torch_state_set_metric)");
    return TORCH_OK;
}

const char* hydro_message(int dim) {
    return dim == 3 ? "3D hydrodynamic update produced a nonphysical state"
                    : "Hydrodynamic update produced a nonphysical state";
}

}  // namespace

// ---------------------------------------------------------------------------
// Allocation-free HLLC face batches (torch_hllc_faces_static). Their staging
// is one file-scope device array, `scratch`, of which every GPU holds its own
// copy for the life of the process. In host code the array's name is only a
// stand-in: its device address is looked up once per device at context
// creation (hipGetSymbolAddress with that device current) and kept in
// g_scratch; kernels and copies take that address, never the name.
//
// Layout for m faces (regions rounded up to mm, a multiple of 64 faces):
//   FP64: left nv*mm | right nv*mm | gl mm | gr mm | out nv*mm        doubles
//   FP32: the float64 inputs left | right | gl | gr ((2 nv + 2) mm doubles),
//         then their float copies and the float result
//         ((3 nv + 2) mm floats); the result is widened back into the start
//         of the double region for the download.
// At nv = 5 that is 136 (FP64) or 164 (FP32) bytes per face; 65536 faces,
// eight of the reference's 8192-face worker batches (torchFlux.py:13), take
// 10.25 MiB. Every region is written before it is read, so nothing needs
// clearing between calls. Calls on one GPU are serialised by a per-device
// mutex (each process has its own copy of the array).
// ---------------------------------------------------------------------------
constexpr long kStaticFaces = 65536;
constexpr size_t kStaticBytes = (size_t)kStaticFaces * ((2 * 5 + 2) * sizeof(double) + (3 * 5 + 2) * sizeof(float));
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kStaticBytes / sizeof(double)];
static void* g_scratch[kMaxDevices];
static std::mutex g_scratch_mu[kMaxDevices];
static std::mutex g_scratch_init_mu;

// Called by torch_ctx_create with `device` current.
static int scratch_init(int device) {
    if (device >= kMaxDevices) return TORCH_OK;  // the static path then reports the device as unsupported
    std::lock_guard<std::mutex> lock(g_scratch_init_mu);
    if (g_scratch[device]) return TORCH_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(TORCH_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = p;
    return TORCH_OK;
}

extern "C" {

int torch_last_error(char* buffer, size_t size) {
    if (!buffer || size == 0) return TORCH_INVALID;
    snprintf(buffer, size, "%s", g_error.c_str());
    return TORCH_OK;
}

int torch_ctx_create(int device, torch_ctx** out) {
    if (!out) return fail(TORCH_INVALID, "null output pointer");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(TORCH_INVALID, "no such HIP device");
    HIP_TRY(hipSetDevice(device));
    torch_ctx* c = new torch_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(TORCH_HIP_ERROR, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    if (int rc = scratch_init(device)) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return TORCH_OK;
}

int torch_ctx_destroy(torch_ctx* ctx) {
    if (!ctx) return TORCH_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TORCH_OK;
}

int torch_state_destroy(torch_state* s) {
    if (!s) return TORCH_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (int i = 0; i < 4; ++i) {
        (void)hipFree(s->U[i]);
        (void)hipFree(s->lam[i]);
        (void)hipFree(s->beta[i]);
    }
    for (int d = 0; d < 3; ++d) (void)hipFree(s->F[d]);
    (void)hipFree(s->mfactors);
    (void)hipFree(s->mprim);
    (void)hipFree(s->mfaces);
    (void)hipFree(s->scratch1d);
    (void)hipFree(s->partials);
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    (void)hipFree(s->scalars);
    (void)hipFree(s->flags);
    (void)hipHostFree(s->host);
    delete s;
    return TORCH_OK;
}

int torch_state_create(torch_ctx* ctx, const torch_config* cfg, torch_state** out) {
    if (!ctx || !cfg || !out) return fail(TORCH_INVALID, "null argument");
    const torch_config& c = *cfg;
    int rc0 = check_config(c);
    if (rc0) return rc0;

    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    torch_state* s = new torch_state();
    s->ctx = ctx;
    s->cfg = c;
    s->ops = c.precision == TORCH_FP64 ? &kOps64 : &kOps32;
    // Default tile, the same as a batch's so that a batch of one folds its
    // sums in the same blocks: FP64 32x16 threads (30x14 cells per plane,
    // two waves per SIMD at 256 VGPRs); FP32 32x8 (three 256-thread blocks
    // per CU at <= 168 VGPRs). The FP32 form's 768-thread 32x24 tile (30x22
    // cells, one block per CU at three waves per SIMD; 86% of its lanes own
    // a cell against 70%) is 5% faster on one 256^3 grid but 2-3% slower on
    // the 32^3..160^3 ignition mix, so it is chosen, not the default.
    if (s->cfg.block_x == 0 || s->cfg.block_y == 0) {
        s->cfg.block_x = 32;
        s->cfg.block_y = c.precision == TORCH_FP32 ? 8 : 16;
    }
    if (s->cfg.zchunk <= 0) s->cfg.zchunk = default_zchunk(c.dim == 3 ? c.n[2] : 1, s->cfg.block_y);
    fvl::Grid& g = s->grid;
    make_grid(c, &g, &s->volume);
    s->ncell = (long)g.nx * g.ny * g.nz;
    s->nf = c.dim + 2;
    const size_t e = s->ops->elem;
    // Inert states with a scalar gamma rotate between 2 slots; states with a
    // per-cell lambda/gamma array use 4 (torch_state_hydro writes slot +2).
    const int nbuf = (c.reactive || c.gamma_mode) ? 4 : 2;
    int rc = TORCH_OK;
    for (int i = 0; i < nbuf && !rc; ++i) rc = alloc(&s->U[i], s->nf * s->ncell * e);
    if (c.gamma_mode)
        for (int i = 0; i < 4 && !rc; ++i) {
            rc = alloc(&s->lam[i], s->ncell * e);
            if (!rc) rc = alloc(&s->beta[i], s->ncell * e);
        }
    fvl::Launch l = launch_of(s);
    int pb = std::max(s->ops->cfl_blocks(g), s->ops->react_blocks(g));
    if (c.dim == 3) pb = std::max(pb, std::max(s->ops->step3d_blocks(g, l, s->cfg.zchunk), s->ops->llf3d_blocks(g)));
    if (c.dim == 2) pb = std::max(pb, std::max(s->ops->step2d_blocks(g, l), s->ops->llf2d_blocks(g)));
    pb = std::max(pb, s->ops->metric_blocks(g));
    if (c.dim == 1) {
        pb = std::max(pb, s->ops->step1d_blocks(g));
        if (!rc) rc = alloc(&s->scratch1d, s->ops->step1d_scratch(g));
    }
    pb = std::max(pb, 256);
    s->pblocks = pb;
    if (!rc) rc = alloc((void**)&s->partials, 4 * (size_t)pb * fvl::NSLOT * sizeof(double));
    if (!rc) rc = alloc((void**)&s->scalars, 4 * fvl::NSLOT * sizeof(double));
    if (!rc) rc = alloc((void**)&s->flags, 16 * sizeof(int));
    if (!rc && hipHostMalloc((void**)&s->host, (4 * fvl::NSLOT + 2) * sizeof(double)) != hipSuccess)
        rc = fail(TORCH_HIP_ERROR, "hipHostMalloc failed");
    if (rc) {
        torch_state_destroy(s);
        return rc;
    }
    *out = s;
    return TORCH_OK;
}

static int upload_field(torch_state* s, void* dst, const double* src, long n) {
    return upload_as(s->ops, s->ctx->stream, dst, src, n);
}

static int download_field(torch_state* s, double* dst, const void* src, long n) {
    return download_as(s->ops, s->ctx->stream, dst, src, n);
}

int torch_state_upload(torch_state* s, const double* U, const double* lam, const double* beta) {
    if (!s || !U) return fail(TORCH_INVALID, "null argument");
    if (s->cfg.gamma_mode && !lam) return fail(TORCH_INVALID, "state needs a lambda (or gamma) array");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    int rc = upload_field(s, s->U[s->cur], U, s->nf * s->ncell);
    if (!rc && s->cfg.gamma_mode) rc = upload_field(s, s->lam[s->cur], lam, s->ncell);
    if (!rc && s->cfg.reactive) {
        if (beta) {
            rc = upload_field(s, s->beta[s->cur], beta, s->ncell);
        } else {
            std::vector<double> ones(s->ncell, 1.0);
            rc = upload_field(s, s->beta[s->cur], ones.data(), s->ncell);
        }
    }
    s->maxima_valid = false;
    s->version += 1;
    return rc;
}

int torch_state_download(torch_state* s, double* U, double* lam, double* beta) {
    if (!s) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    int rc = TORCH_OK;
    if (U) rc = download_field(s, U, s->U[s->cur], s->nf * s->ncell);
    if (!rc && lam && s->cfg.reactive) rc = download_field(s, lam, s->lam[s->cur], s->ncell);
    if (!rc && beta && s->cfg.reactive) rc = download_field(s, beta, s->beta[s->cur], s->ncell);
    return rc;
}

int torch_state_dt(torch_state* s, double* dt_limit) {
    if (!s || !dt_limit) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    if (!s->maxima_valid) {
        int rc = compute_maxima(s);
        if (rc) return rc;
    }
    *dt_limit = dt_from_maxima(s, s->maxima);
    return TORCH_OK;
}

int torch_state_wave_maxima(torch_state* s, double* maxima) {
    if (!s || !maxima) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    if (s->attempt_slot >= 0 && s->attempt_version == s->version) {
        // Right after an FP32 half-step: the maxima of the state it started
        // from raised by the heating's exact change. Not cached: the maxima
        // that set the next dt stay those of the stored state.
        HIP_TRY(clear_flags(s->flags, s->ctx->stream));
        HIP_TRY(run_cfl(s, s->attempt_slot, s->attempt_slot, s->attempt_inc));
        int rc = sync_read(s);
        if (rc) return rc;
        for (int m = 0; m < 3; ++m) maxima[m] = s->host[3 * fvl::NSLOT + m];
        return TORCH_OK;
    }
    if (!s->maxima_valid) {
        int rc = compute_maxima(s);
        if (rc) return rc;
    }
    for (int m = 0; m < 3; ++m) maxima[m] = s->maxima[m];
    return TORCH_OK;
}

int torch_state_step(torch_state* s, double time, double t_end, torch_step_info* info) {
    if (!s || !info) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    memset(info, 0, sizeof(*info));
    const torch_config& c = s->cfg;
    int rc;
    if (!s->maxima_valid && (rc = compute_maxima(s))) return rc;
    double dt = py_min(dt_from_maxima(s, s->maxima), t_end - time);
    if (!isfinite(dt) || dt <= 0 || time + dt == time)
        return fail(TORCH_STALLED, "Integration stalled before requested final time");
    const int A = s->cur;
    const int B = (A + 1) % 4, C = (A + 2) % 4, D = (A + 3) % 4;
    int hin = A, hout = c.reactive ? C : (A + 1) % 2;
    if (c.reactive) {
        int it = 0;
        for (; it < 60; ++it) {
            HIP_TRY(clear_flags(s->flags, s->ctx->stream));
            if ((rc = run_react(s, A, B, dt / 2, -1, -1, true))) return rc;
            if ((rc = sync_read(s))) return rc;
            if (host_flags(s) & fvl::F_REACT_BAD)
                return fail(TORCH_INVALID, "Reaction input outside its domain (state, lambda or beta)");
            double mx[3];
            const double* src = s->host + (c.dim == 3 ? 2 : 3) * fvl::NSLOT;
            for (int m = 0; m < 3; ++m) mx[m] = src[m];
            double allowed = dt_from_maxima(s, mx);
            if (dt <= allowed * (1 + 1e-12)) break;
            dt = py_min(dt / 2, allowed);
        }
        if (it == 60) return fail(TORCH_NO_STABLE_STEP, "Could not find stable reactive step");
        info->retries = it;
        hin = B;
    }
    if ((rc = run_hydro(s, hin, hout, dt, B, C, false))) return rc;
    if (c.reactive) {
        if ((rc = run_react(s, C, D, dt / 2, A, B, true))) return rc;
    }
    if ((rc = sync_read(s))) return rc;
    const int flags = host_flags(s);
    const bool fb = (flags & (fvl::F_PRED_BAD | fvl::F_UPD_BAD)) != 0;
    if (flags & fvl::F_LLF_BAD) return fail(TORCH_NONPHYSICAL, hydro_message(c.dim));
    if (c.dim == 3 && (flags & (fb ? fvl::F_SPEC_BAD_LLF : fvl::F_SPEC_BAD)))
        return fail(TORCH_SPECIES_RANGE, "3D scalar transport left the physical interval");
    if (flags & fvl::F_REACT_BAD) return fail(TORCH_INVALID, "Reaction input outside its domain");
    if (flags & fvl::F_STATE_BAD) return fail(TORCH_REACTION_STATE, "Reaction produced a nonphysical state");
    const double* hv = s->host + (fb && c.dim >= 2 ? fvl::NSLOT : 0);
    info->dt = dt;
    info->fallback = fb;
    for (int f = 0; f < s->nf; ++f) info->outward[f] = hv[fvl::S_OUT + f];
    if (c.reactive) {
        info->source[s->nf - 1] = s->volume * s->host[2 * fvl::NSLOT + fvl::S_SRC];
        const double* mxs = s->host + (c.dim == 3 ? 2 : 3) * fvl::NSLOT;
        for (int m = 0; m < 3; ++m) s->maxima[m] = mxs[m];
    } else {
        const double* mxs = c.dim == 3 ? hv : s->host + 3 * fvl::NSLOT;
        for (int m = 0; m < 3; ++m) s->maxima[m] = mxs[m];
    }
    s->maxima_valid = true;
    for (int m = 0; m < 3; ++m) info->wave_max[m] = s->maxima[m];
    s->cur = c.reactive ? D : hout;
    s->version += 1;
    return TORCH_OK;
}

int torch_state_hydro(torch_state* s, double dt, int want_fluxes, torch_step_info* info) {
    if (!s || !info) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    memset(info, 0, sizeof(*info));
    int rc;
    if (want_fluxes)
        for (int d = 0; d < s->cfg.dim; ++d)
            if (!s->F[d] && (rc = alloc(&s->F[d], s->nf * face_count(s->grid, d) * s->ops->elem))) return rc;
    const int A = s->cur;
    const int out = s->cfg.gamma_mode ? (A + 2) % 4 : (A + 1) % 2;
    if ((rc = run_hydro(s, A, out, dt, A, out, want_fluxes != 0))) return rc;
    if ((rc = sync_read(s))) return rc;
    const int flags = host_flags(s);
    const bool fb = (flags & (fvl::F_PRED_BAD | fvl::F_UPD_BAD)) != 0;
    if (flags & fvl::F_LLF_BAD) return fail(TORCH_NONPHYSICAL, hydro_message(s->cfg.dim));
    const double* hv = s->host + (fb && s->cfg.dim >= 2 ? fvl::NSLOT : 0);
    info->dt = dt;
    info->fallback = fb;
    for (int f = 0; f < s->nf; ++f) info->outward[f] = hv[fvl::S_OUT + f];
    s->cur = out;
    s->version += 1;
    // 1D/2D scalar-gamma inert: run_hydro already folded the CFL maxima of
    // the new state (the pass compute_maxima would launch), so the next
    // dt needs no launch and no wait.
    s->maxima_valid = s->cfg.dim != 3 && !s->cfg.reactive && s->cfg.gamma_mode == 0;
    if (s->maxima_valid)
        for (int m = 0; m < 3; ++m) s->maxima[m] = s->host[3 * fvl::NSLOT + m];
    return TORCH_OK;
}

int torch_state_fluxes(torch_state* s, double* fx, double* fy, double* fz) {
    if (!s) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    double* dst[3] = {fx, fy, fz};
    for (int d = 0; d < s->cfg.dim; ++d) {
        if (!dst[d]) continue;
        if (!s->F[d]) return fail(TORCH_INVALID, "no fluxes recorded");
        int rc = download_field(s, dst[d], s->F[d], s->nf * face_count(s->grid, d));
        if (rc) return rc;
    }
    return TORCH_OK;
}

int torch_state_react(torch_state* s, double dt) {
    if (!s) return fail(TORCH_INVALID, "null argument");
    if (!s->cfg.reactive) return fail(TORCH_INVALID, "state is not reactive");
    if (!isfinite(dt) || dt < 0 || !isfinite(s->cfg.q_heat) || s->cfg.q_heat < 0 ||
        !isfinite(s->cfg.activation_energy) || s->cfg.activation_energy < 0 ||
        !isfinite(s->cfg.pre_exponential) || s->cfg.pre_exponential < 0)
        return fail(TORCH_INVALID, "Reaction time and constants must be finite and nonnegative");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    const int A = s->cur, B = (A + 1) % 4, C = (A + 2) % 4;
    HIP_TRY(clear_flags(s->flags, s->ctx->stream));
    // FP32 form, 1D/2D Cartesian: the half-step's per-cell increments go to
    // slot C, which nothing reads or writes before the state next changes.
    const bool record = s->cfg.precision == TORCH_FP32 && s->cfg.dim != 3 && !s->has_metric;
    int rc = run_react(s, A, B, dt, -1, -1, false, record ? s->U[C] : nullptr,
                       s->cfg.dim == 1 ? fvl::INC_HEAT : fvl::INC_SPEED);
    if (!rc) rc = sync_read(s);
    if (rc) return rc;
    if (host_flags(s) & fvl::F_REACT_BAD)
        return fail(TORCH_INVALID, "Reaction requires a physical state and fractions in [0,1]");
    s->cur = B;
    s->maxima_valid = false;
    s->version += 1;
    s->attempt_slot = record ? A : -1;
    s->attempt_version = s->version;
    s->attempt_inc = record ? s->U[C] : nullptr;
    return TORCH_OK;
}

int torch_state_totals(torch_state* s, double* out) {
    if (!s || !out) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    HIP_TRY(s->ops->totals(s->grid, s->U[s->cur], region(s, 2), s->ctx->stream));
    HIP_TRY(fvl::finalize(region(s, 2), 256, s->scalars + 2 * fvl::NSLOT, s->ctx->stream));
    int rc = sync_read(s);
    if (rc) return rc;
    for (int f = 0; f < s->nf; ++f) out[f] = s->host[2 * fvl::NSLOT + fvl::S_OUT + f] * s->volume;
    return TORCH_OK;
}

int torch_state_profile(torch_state* s, int enable, double* kernel_ms, long* launches) {
    if (!s) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    if (enable && !s->ev0) {
        HIP_TRY(hipEventCreate(&s->ev0));
        HIP_TRY(hipEventCreate(&s->ev1));
    }
    if (kernel_ms) *kernel_ms = s->prof_ms;
    if (launches) *launches = s->prof_n;
    s->prof_ms = 0;
    s->prof_n = 0;
    s->profile = enable != 0;
    return TORCH_OK;
}

int torch_state_set_metric(torch_state* s, int geometry, const double* factors, long nfactors,
                           const torch_face* faces) {
    if (!s || !factors || !faces) return fail(TORCH_INVALID, "null argument");
    if (geometry < TORCH_CARTESIAN || geometry > TORCH_SPHERICAL) return fail(TORCH_INVALID, "Unknown geometry");
    const int dim = s->cfg.dim;
    const int n[3] = {s->grid.nx, s->grid.ny, s->grid.nz};
    long need = 0;
    for (int a = 0; a < dim; ++a) need += n[a];
    for (int d = 0; d < dim; ++d)
        for (int a = 0; a < dim; ++a) need += n[a] + (a == d);
    if (nfactors != need) return fail(TORCH_INVALID, "metric factor table has the wrong length");
    for (long i = 0; i < need; ++i)
        if (!isfinite(factors[i])) return fail(TORCH_INVALID, "metric factors must be finite");
    for (int f = 0; f < 2 * dim; ++f) {
        const int t = faces[f].type;
        if (t < TORCH_FACE_OUTFLOW || t > TORCH_FACE_STAGNATION_INLET)
            return fail(TORCH_INVALID, "Unknown face boundary type");
        if ((t == TORCH_FACE_PERIODIC) != (faces[f ^ 1].type == TORCH_FACE_PERIODIC))
            return fail(TORCH_INVALID, "Periodic boundaries must be paired on an axis");
    }
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    // Four state slots: start, chemistry, stage, hydro.
    const size_t e = s->ops->elem;
    for (int i = 0; i < kMetricSlots; ++i)
        if (!s->U[i]) {
            int rc = alloc(&s->U[i], s->nf * s->ncell * e);
            if (rc) return rc;
        }
    if (!s->mfactors) {
        int rc = alloc((void**)&s->mfactors, need * sizeof(double));
        if (rc) return rc;
    }
    if (!s->mprim) {
        int rc = alloc(&s->mprim, (s->nf + 1) * s->ncell * e);
        if (rc) return rc;
    }
    // Face fluxes of one stage, per direction: (dim+2) weighted fluxes and
    // two moved fractions per face.
    long nfaces[3] = {0, 0, 0}, total = 0;
    for (int d = 0; d < dim; ++d) total += (nfaces[d] = face_count(s->grid, d));
    if (!s->mfaces) {
        int rc = alloc(&s->mfaces, (s->nf + 2) * total * e);
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(s->mfactors, factors, need * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    fvl::Metric& m = s->metric;
    memset(&m, 0, sizeof(m));
    long at = 0;
    for (int a = 0; a < dim; ++a) {
        m.vf[a] = s->mfactors + at;
        at += n[a];
    }
    for (int d = 0; d < dim; ++d)
        for (int a = 0; a < dim; ++a) {
            m.af[d][a] = s->mfactors + at;
            at += n[a] + (a == d);
        }
    m.geometry = geometry;
    m.prim = s->mprim;
    for (long d = 0, off = 0; d < dim; off += (s->nf + 2) * nfaces[d] * (long)e, ++d)
        m.wf[d] = (char*)s->mfaces + off;
    for (int f = 0; f < 2 * dim; ++f) {
        fvl::MFace& o = m.face[f];
        const torch_face& i = faces[f];
        o.type = i.type;
        o.has_backflow = i.has_backflow;
        for (int k = 0; k < 5; ++k) o.prim[k] = i.primitive[k];
        o.pressure = i.pressure;
        o.total_pressure = i.total_pressure;
        o.total_temperature = i.total_temperature;
        o.gas_constant = i.gas_constant;
        o.lam = i.lambda;
        o.beta = i.beta;
    }
    s->has_metric = true;
    return TORCH_OK;
}

int torch_state_metric_dt(torch_state* s, double* dt) {
    if (!s || !dt) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    int rc = need_metric(s);
    if (rc) return rc;
    return metric_dt(s, s->cur, dt);
}

int torch_state_metric_rhs(torch_state* s, int first_order, double* change, double* species, double* outward,
                           double* source) {
    if (!s) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    int rc = need_metric(s);
    if (rc) return rc;
    const int A = s->cur, B = (A + 1) % kMetricSlots;
    int flags = 0;
    if ((rc = metric_stage(s, 0, A, A, B, 0.0, first_order != 0, &flags))) return rc;
    for (int f = 0; f < s->nf; ++f) {
        if (outward) outward[f] = s->host[fvl::S_OUT + f];
        if (source) source[f] = s->host[fvl::S_MSRC + f];
    }
    if (change && (rc = download_as(s->ops, s->ctx->stream, change, s->U[B], s->nf * s->ncell))) return rc;
    for (int j = 0; j < metric_nq(s) && species; ++j)
        if ((rc = download_as(s->ops, s->ctx->stream, species + j * s->ncell, frac(s, B, j), s->ncell))) return rc;
    return TORCH_OK;
}

int torch_state_metric_advance(torch_state* s, double dt, torch_step_info* info) {
    if (!s || !info) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    memset(info, 0, sizeof(*info));
    int rc = need_metric(s);
    if (rc) return rc;
    const int A = s->cur, B = (A + 1) % kMetricSlots, C = (A + 2) % kMetricSlots;
    if ((rc = metric_advance(s, A, B, C, dt, info->outward, info->source, &info->fallback))) return rc;
    info->dt = dt;
    s->cur = C;
    s->maxima_valid = false;
    s->version += 1;
    return TORCH_OK;
}

int torch_state_metric_step(torch_state* s, double time, double t_end, torch_step_info* info) {
    if (!s || !info) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    HIP_TRY(hipSetDevice(s->ctx->device));
    memset(info, 0, sizeof(*info));
    int rc = need_metric(s);
    if (rc) return rc;
    const bool reactive = s->cfg.reactive;
    if (!reactive) {
        rc = metric_fast_step(s, time, t_end, info);
        if (rc != kSlowPath) return rc;
        memset(info, 0, sizeof(*info));
    }
    const int A = s->cur;
    const int B = (A + 1) % kMetricSlots, C = (A + 2) % kMetricSlots, D = (A + 3) % kMetricSlots;
    double dt;
    if ((rc = metric_dt(s, A, &dt))) return rc;
    dt = py_min(dt, t_end - time);
    for (int retry = 0; retry < 20; ++retry) {
        if (!isfinite(dt) || dt <= 0 || time + dt == time)
            return fail(TORCH_STALLED, "Integration stalled before requested final time");
        int old = A;
        if (reactive) {
            HIP_TRY(clear_flags(s->flags, s->ctx->stream));
            // The FP32 form checks the committed state's rates raised by the
            // heating's exact change (increments in slot C, free until the
            // stage below), as its Cartesian sessions do.
            void* incr = s->cfg.precision == TORCH_FP32 ? s->U[C] : nullptr;
            if ((rc = run_react(s, A, B, dt / 2, -1, -1, false, incr))) return rc;
            if ((rc = sync_read(s))) return rc;
            if (host_flags(s) & fvl::F_REACT_BAD)
                return fail(TORCH_INVALID, "Reaction input outside its domain (state, lambda or beta)");
            double allowed;
            if ((rc = incr ? metric_dt(s, A, &allowed, incr) : metric_dt(s, B, &allowed))) return rc;
            if (dt > allowed * (1 + 1e-12)) {
                dt = py_min(dt / 2, allowed);
                info->retries += 1;
                continue;
            }
            old = B;
        }
        // inert: stage B, hydro C; reactive: stage C, hydro D, final C
        const int stage = reactive ? C : B, hydro = reactive ? D : C;
        double boundary[5] = {0, 0, 0, 0, 0}, source[5] = {0, 0, 0, 0, 0};
        int first = 0;
        rc = metric_advance(s, old, stage, hydro, dt, boundary, source, &first);
        if (rc == TORCH_NONPHYSICAL) {
            dt /= 2;
            info->retries += 1;
            continue;
        }
        if (rc) return rc;
        int final_slot = hydro;
        if (reactive) {
            HIP_TRY(clear_flags(s->flags, s->ctx->stream));
            if ((rc = run_react(s, hydro, C, dt / 2, -1, -1, false))) return rc;
            HIP_TRY(s->ops->metric_source(s->grid, s->metric, s->U[A], s->U[B], s->U[C], s->U[hydro], region(s, 1),
                                          s->ctx->stream));
            HIP_TRY(fvl::finalize(region(s, 1), s->ops->metric_blocks(s->grid), s->scalars + fvl::NSLOT,
                                  s->ctx->stream));
            if ((rc = sync_read(s))) return rc;
            const int flags = host_flags(s);
            if (flags & fvl::F_REACT_BAD) return fail(TORCH_INVALID, "Reaction input outside its domain");
            if (flags & fvl::F_STATE_BAD) return fail(TORCH_REACTION_STATE, "Reaction produced a nonphysical state");
            const int e = s->nf - 1;
            source[e] = source[e] + s->host[fvl::NSLOT + fvl::S_MSRC + e];
            final_slot = C;
        }
        info->dt = dt;
        info->fallback = first;
        for (int f = 0; f < s->nf; ++f) {
            info->outward[f] = boundary[f];
            info->source[f] = source[f];
        }
        s->cur = final_slot;
        s->maxima_valid = false;
        s->version += 1;
        return TORCH_OK;
    }
    return fail(TORCH_NONPHYSICAL, "Could not find a positive metric/face step");
}


// A face-array kernel (HLLC or Rusanov) on host arrays: upload, launch, download.
static int faces_call(torch_ctx* ctx, int precision, int nv, int direction, long m, const double* left,
                      const double* right, const double* gl, const double* gr, double* out,
                      hipError_t (*const Ops::*kernel)(int, int, long, const void*, const void*, const void*,
                                                       const void*, void*, hipStream_t)) {
    if (m == 0) return TORCH_OK;
    const Ops* ops = precision == TORCH_FP32 ? &kOps32 : &kOps64;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t e = ops->elem;
    void *dl = nullptr, *dr = nullptr, *dgl = nullptr, *dgr = nullptr, *dout = nullptr;
    int rc = alloc(&dl, nv * m * e);
    if (!rc) rc = alloc(&dr, nv * m * e);
    if (!rc) rc = alloc(&dgl, m * e);
    if (!rc) rc = alloc(&dgr, m * e);
    if (!rc) rc = alloc(&dout, nv * m * e);
    if (!rc) rc = upload_as(ops, ctx->stream, dl, left, nv * m);
    if (!rc) rc = upload_as(ops, ctx->stream, dr, right, nv * m);
    if (!rc) rc = upload_as(ops, ctx->stream, dgl, gl, m);
    if (!rc) rc = upload_as(ops, ctx->stream, dgr, gr, m);
    if (!rc) {
        hipError_t he = (ops->*kernel)(nv, direction, m, dl, dr, dgl, dgr, dout, ctx->stream);
        if (he != hipSuccess) rc = fail(TORCH_HIP_ERROR, hipGetErrorString(he));
    }
    if (!rc) rc = download_as(ops, ctx->stream, out, dout, nv * m);
    (void)hipFree(dl);
    (void)hipFree(dr);
    (void)hipFree(dgl);
    (void)hipFree(dgr);
    (void)hipFree(dout);
    return rc;
}

int torch_hllc_faces_static(torch_ctx* ctx, int precision, int nv, int direction, long m, const double* left,
                            const double* right, const double* gl, const double* gr, double* out) {
    if (!ctx || !left || !right || !gl || !gr || !out) return fail(TORCH_INVALID, "null argument");
    if ((nv != 4 && nv != 5) || direction < 0 || direction >= nv - 2 || m < 0)
        return fail(TORCH_INVALID, "HLLC requires matching four/five-component grids");
    if (precision != TORCH_FP64 && precision != TORCH_FP32) return fail(TORCH_INVALID, "unknown precision");
    if (m > kStaticFaces)
        return fail(TORCH_INVALID, "torch_hllc_faces_static takes at most 65536 faces per call (its staging is a "
                                   "fixed device array); split the batch or use torch_hllc_faces");
    if (m == 0) return TORCH_OK;
    if (ctx->device >= kMaxDevices || !g_scratch[ctx->device])
        return fail(TORCH_HIP_ERROR, "no static face scratch on this device");
    const Ops* ops = precision == TORCH_FP32 ? &kOps32 : &kOps64;
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::lock_guard<std::mutex> slock(g_scratch_mu[ctx->device]);
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const long mm = (m + 63) / 64 * 64;
    double* dl = (double*)g_scratch[ctx->device];
    double* dr = dl + nv * mm;
    double* dgl = dr + nv * mm;
    double* dgr = dgl + mm;
    HIP_TRY(hipMemcpyAsync(dl, left, nv * m * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dr, right, nv * m * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dgl, gl, m * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dgr, gr, m * sizeof(double), hipMemcpyHostToDevice, st));
    const double* result;
    if (precision == TORCH_FP64) {
        double* dout = dgr + mm;
        HIP_TRY(ops->hllc_faces(nv, direction, m, dl, dr, dgl, dgr, dout, st));
        result = dout;
    } else {
        float* fl = (float*)(dgr + mm);
        float* fr = fl + nv * mm;
        float* fgl = fr + nv * mm;
        float* fgr = fgl + mm;
        float* fout = fgr + mm;
        HIP_TRY(ops->convert_in(dl, fl, nv * m, st));
        HIP_TRY(ops->convert_in(dr, fr, nv * m, st));
        HIP_TRY(ops->convert_in(dgl, fgl, m, st));
        HIP_TRY(ops->convert_in(dgr, fgr, m, st));
        HIP_TRY(ops->hllc_faces(nv, direction, m, fl, fr, fgl, fgr, fout, st));
        HIP_TRY(ops->convert_out(fout, dl, nv * m, st));  // the inputs are consumed by now (stream order)
        result = dl;
    }
    HIP_TRY(hipMemcpyAsync(out, result, nv * m * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return TORCH_OK;
}

int torch_hllc_faces(torch_ctx* ctx, int precision, int nv, int direction, long m, const double* left,
                     const double* right, const double* gl, const double* gr, double* out) {
    if (!ctx || !left || !right || !gl || !gr || !out) return fail(TORCH_INVALID, "null argument");
    if ((nv != 4 && nv != 5) || direction < 0 || direction >= nv - 2 || m < 0)
        return fail(TORCH_INVALID, "HLLC requires matching four/five-component grids");
    return faces_call(ctx, precision, nv, direction, m, left, right, gl, gr, out, &Ops::hllc_faces);
}

int torch_rusanov_faces(torch_ctx* ctx, int precision, int nv, int direction, long m, const double* left,
                        const double* right, const double* gl, const double* gr, double* out) {
    if (!ctx || !left || !right || !gl || !gr || !out) return fail(TORCH_INVALID, "null argument");
    if (nv < 3 || nv > 5 || direction < 0 || direction >= nv - 2 || m < 0)
        return fail(TORCH_INVALID, "Rusanov requires (dim+2)-component face arrays and a valid direction");
    return faces_call(ctx, precision, nv, direction, m, left, right, gl, gr, out, &Ops::rusanov_faces);
}

}  // extern "C"
