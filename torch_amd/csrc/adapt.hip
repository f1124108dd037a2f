// C ABI of the adaptive leaves (include/torch_hip.h): one device-resident
// leaf mesh with its state slots, and the step of AdaptiveSimulation
// (python/cfdSolver/torchAdaptive.py:106-206, 436-481) up to its regrid.
// The mesh arrives as flat arrays from the host's leaf tree; every array is
// validated here so that every index a kernel follows is in range.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "capi_common.h"

using namespace tcapi;

struct torch_amr {
    torch_ctx* ctx = nullptr;
    torch_config cfg;
    const Ops* ops = nullptr;
    int nf = 0, nq = 0, order = 2, geometry = 0;
    fvl::MFace faces[6];
    long nleaf = 0, nface = 0, nentry = 0;
    long cap_leaf = 0, cap_face = 0, cap_entry = 0;
    bool has_mesh = false, has_state = false;
    int cur = 0;
    // The scratch arrays hold the checked evaluation (face values, gradients,
    // fluxes, outward sums) of slot cache_slot with reconstruction
    // cache_rec; -1 when they hold none. A stage whose input is that slot
    // reuses it (the step's CFL evaluation is its first stage's).
    int cache_slot = -1, cache_rec = -1;
    void* U[4] = {nullptr, nullptr, nullptr, nullptr};
    void* q[2][4] = {{nullptr, nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr, nullptr}};
    // mesh
    int *left = nullptr, *right = nullptr, *dir = nullptr, *side = nullptr, *cface = nullptr, *crole = nullptr;
    long* start = nullptr;
    double *area = nullptr, *weight = nullptr, *offl = nullptr, *offr = nullptr, *disp = nullptr;
    double *volume = nullptr, *div = nullptr, *inverse = nullptr;
    // one evaluation's face and leaf values
    void *FL = nullptr, *FR = nullptr, *G = nullptr, *Q = nullptr, *grad = nullptr, *W = nullptr, *M = nullptr,
         *S = nullptr;
    double* partials = nullptr;  // 4 regions: faces, leaves, CFL, chemistry
    long pblocks = 0;
    double* scalars = nullptr;   // kRows x NSLOT finalized values
    int* flags = nullptr;
    double* host = nullptr;      // pinned readback: kRows x NSLOT doubles + flags
};

namespace {

const int kSlots = 4;
// Finalized rows: 0 outward sums of an evaluation, 1 leaf sums of a stage,
// 2 the CFL rate (and the queued step's dt in its last slot), 3 unused,
// 4 and 5 the second stage's sums in a queued step.
const int kRows = 6;

__global__ void k_amr_clear(int* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

hipError_t clear(torch_amr* a) {
    hipLaunchKernelGGL(k_amr_clear, dim3(1), dim3(64), 0, a->ctx->stream, a->flags, 2);
    return hipGetLastError();
}

double* region(torch_amr* a, int r) { return a->partials + (long)r * a->pblocks * fvl::NSLOT; }

int sync_read(torch_amr* a) {
    hipStream_t st = a->ctx->stream;
    HIP_TRY(hipMemcpyAsync(a->host, a->scalars, kRows * fvl::NSLOT * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(a->host + kRows * fvl::NSLOT, a->flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return TORCH_OK;
}

int host_flags(const torch_amr* a) {
    int f;
    memcpy(&f, a->host + kRows * fvl::NSLOT, sizeof(int));
    return f;
}

int boundary_error(int flags) {
    if (flags & fvl::F_BC_RESERVOIR) return fail(TORCH_INVALID, "No subsonic reservoir-inlet solution");
    if (flags & fvl::F_BC_SUPPLY)
        return fail(TORCH_INVALID, "Reservoir cannot supply inflow against the interior pressure");
    if (flags & fvl::F_BC_BACKFLOW)
        return fail(TORCH_INVALID, "Pressure outlet backflow requires an explicit backflow primitive state");
    return TORCH_OK;
}

double py_min(double a, double b) { return b < a ? b : a; }

void release(void** p) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
}

template <class T>
void release_t(T** p) {
    void* v = (void*)*p;
    release(&v);
    *p = nullptr;
}

void release_buffers(torch_amr* a) {
    for (int i = 0; i < kSlots; ++i) {
        release(&a->U[i]);
        release(&a->q[0][i]);
        release(&a->q[1][i]);
    }
    for (int** p : {&a->left, &a->right, &a->dir, &a->side, &a->cface, &a->crole}) release_t(p);
    release_t(&a->start);
    for (double** p : {&a->area, &a->weight, &a->offl, &a->offr, &a->disp, &a->volume, &a->div, &a->inverse,
                       &a->partials})
        release_t(p);
    for (void** p : {&a->FL, &a->FR, &a->G, &a->Q, &a->grad, &a->W, &a->M, &a->S}) release(p);
    a->cap_leaf = a->cap_face = a->cap_entry = 0;
}

// Buffers for at least L leaves, F faces and E incidence entries.
int reserve(torch_amr* a, long L, long F, long E) {
    if (L <= a->cap_leaf && F <= a->cap_face && E <= a->cap_entry) return TORCH_OK;
    release_buffers(a);
    const long cl = std::max(L + L / 4, 256L), cf = std::max(F + F / 4, 256L), ce = std::max(E + E / 4, 256L);
    const int d = a->cfg.dim, nf = a->nf;
    const size_t e = a->ops->elem;
    int rc = TORCH_OK;
    for (int i = 0; i < kSlots && !rc; ++i) {
        rc = alloc(&a->U[i], nf * cl * e);
        for (int j = 0; j < a->nq && !rc; ++j) rc = alloc(&a->q[j][i], cl * e);
    }
    if (!rc) rc = alloc((void**)&a->left, cf * sizeof(int));
    if (!rc) rc = alloc((void**)&a->right, cf * sizeof(int));
    if (!rc) rc = alloc((void**)&a->dir, cf * sizeof(int));
    if (!rc) rc = alloc((void**)&a->side, cf * sizeof(int));
    if (!rc) rc = alloc((void**)&a->area, cf * sizeof(double));
    if (!rc) rc = alloc((void**)&a->weight, cf * sizeof(double));
    if (!rc) rc = alloc((void**)&a->offl, cf * d * sizeof(double));
    if (!rc) rc = alloc((void**)&a->offr, cf * d * sizeof(double));
    if (!rc) rc = alloc((void**)&a->disp, cf * d * sizeof(double));
    if (!rc) rc = alloc((void**)&a->volume, cl * sizeof(double));
    if (!rc) rc = alloc((void**)&a->div, cl * d * sizeof(double));
    if (!rc) rc = alloc((void**)&a->inverse, cl * d * d * sizeof(double));
    if (!rc) rc = alloc((void**)&a->start, (cl + 1) * sizeof(long));
    if (!rc) rc = alloc((void**)&a->cface, ce * sizeof(int));
    if (!rc) rc = alloc((void**)&a->crole, ce * sizeof(int));
    if (!rc) rc = alloc(&a->FL, nf * cf * e);
    if (!rc) rc = alloc(&a->FR, nf * cf * e);
    if (!rc) rc = alloc(&a->G, 2 * cf * e);
    if (!rc) rc = alloc(&a->Q, 4 * cf * e);
    if (!rc) rc = alloc(&a->grad, cl * d * nf * e);
    if (!rc) rc = alloc(&a->W, nf * cf * e);
    if (!rc) rc = alloc(&a->M, 2 * cf * e);
    if (!rc) rc = alloc(&a->S, cf * e);
    a->pblocks = std::max((long)a->ops->amr_blocks(cf), (long)a->ops->amr_blocks(cl));
    if (!rc) rc = alloc((void**)&a->partials, 4 * a->pblocks * fvl::NSLOT * sizeof(double));
    if (rc) {
        release_buffers(a);
        return rc;
    }
    a->cap_leaf = cl;
    a->cap_face = cf;
    a->cap_entry = ce;
    return TORCH_OK;
}

fvl::Amr mesh_of(const torch_amr* a) {
    fvl::Amr m;
    memset(&m, 0, sizeof(m));
    m.nleaf = a->nleaf;
    m.nface = a->nface;
    m.left = a->left;
    m.right = a->right;
    m.dir = a->dir;
    m.side = a->side;
    m.area = a->area;
    m.offl = a->offl;
    m.offr = a->offr;
    m.disp = a->disp;
    m.weight = a->weight;
    m.volume = a->volume;
    m.div = a->div;
    m.inverse = a->inverse;
    m.start = a->start;
    m.cface = a->cface;
    m.crole = a->crole;
    for (int f = 0; f < 6; ++f) m.face[f] = a->faces[f];
    m.geometry = a->geometry;
    m.FL = a->FL;
    m.FR = a->FR;
    m.G = a->G;
    m.Q = a->Q;
    m.grad = a->grad;
    m.W = a->W;
    m.M = a->M;
    m.S = a->S;
    return m;
}

fvl::Grid flat_grid(const torch_amr* a) {
    fvl::Grid g;
    memset(&g, 0, sizeof(g));
    g.dim = a->cfg.dim;
    g.nx = (int)a->nleaf;
    g.ny = 1;
    g.nz = 1;
    for (int d = 0; d < 3; ++d) g.h[d] = 1.0;
    return g;
}

void* frac(torch_amr* a, int slot, int j) { return j < a->nq ? a->q[j][slot] : nullptr; }

// Face values, gradients and fluxes of slot u; outward sums into scalars[0].
// Nothing to launch when the scratch already holds this evaluation.
int evaluate(torch_amr* a, int u, int reconstruct) {
    if (a->cache_slot == u && a->cache_rec == reconstruct) return TORCH_OK;
    a->cache_slot = -1;
    hipStream_t st = a->ctx->stream;
    const fvl::Amr m = mesh_of(a);
    HIP_TRY(a->ops->amr_eval(a->cfg.dim, m, a->U[u], frac(a, u, 0), frac(a, u, 1), a->nq, a->cfg.gamma,
                             a->cfg.reactive, reconstruct, a->flags, region(a, 0), st));
    HIP_TRY(fvl::finalize(region(a, 0), a->ops->amr_blocks(a->nface), a->scalars, st));
    return TORCH_OK;
}

void remember(torch_amr* a, int u, int reconstruct) {
    a->cache_slot = u;
    a->cache_rec = reconstruct;
}

// A slot is about to be overwritten: its cached evaluation goes stale.
void overwrite(torch_amr* a, int slot) {
    if (a->cache_slot == slot) a->cache_slot = -1;
}

// cfl (torchAdaptive.py:149-166) of slot u.
int leaf_dt(torch_amr* a, int u, double* dt) {
    hipStream_t st = a->ctx->stream;
    HIP_TRY(clear(a));
    int rc = evaluate(a, u, a->order == 2);
    if (rc) return rc;
    HIP_TRY(a->ops->amr_cfl(a->cfg.dim, mesh_of(a), region(a, 2), st));
    HIP_TRY(fvl::finalize(region(a, 2), a->ops->amr_blocks(a->nleaf), a->scalars + 2 * fvl::NSLOT, st));
    if ((rc = sync_read(a))) return rc;
    if ((rc = boundary_error(host_flags(a)))) return rc;
    remember(a, u, a->order == 2);
    *dt = a->cfg.cfl / a->host[2 * fvl::NSLOT];
    return TORCH_OK;
}

// One SSPRK2 stage (mode 0: the raw change) of slot S into slot out; leaves
// the boundary sums in scalars[0] and the geometric source in scalars[1].
int stage(torch_amr* a, int mode, int S, int U0, int out, double dt, int first_order, int* flags) {
    fvl::MStage m;
    memset(&m, 0, sizeof(m));
    m.S = a->U[S];
    m.U0 = a->U[U0];
    m.out = a->U[out];
    for (int j = 0; j < 2; ++j) {
        m.qS[j] = frac(a, S, j);
        m.q0[j] = frac(a, U0, j);
        m.qout[j] = frac(a, out, j);
    }
    m.dt = dt;
    m.gamma = a->cfg.gamma;
    m.mode = mode;
    m.first_order = first_order;
    m.nq = a->nq;
    m.reactive = a->cfg.reactive;
    m.flags = a->flags;
    m.partials = region(a, 1);
    hipStream_t st = a->ctx->stream;
    const int rec = a->order == 2 && !first_order;
    HIP_TRY(clear(a));
    int rc = evaluate(a, S, rec);
    if (rc) return rc;
    overwrite(a, out);
    HIP_TRY(a->ops->amr_update(a->cfg.dim, mesh_of(a), m, st));
    HIP_TRY(fvl::finalize(region(a, 1), a->ops->amr_blocks(a->nleaf), a->scalars + fvl::NSLOT, st));
    if ((rc = sync_read(a))) return rc;
    *flags = host_flags(a);
    if ((rc = boundary_error(*flags))) return rc;
    remember(a, S, rec);
    return TORCH_OK;
}

// advance (torchAdaptive.py:169-206): slot `in` -> `out` through `mid`.
int advance(torch_amr* a, int in, int mid, int out, double dt, double* boundary, double* source, int* first_used) {
    const int nf = a->nf;
    for (int first = 0; first < 2; ++first) {
        int flags = 0, rc;
        if ((rc = stage(a, 1, in, in, mid, dt, first, &flags))) return rc;
        if (flags & (fvl::F_M_UNPHYS | fvl::F_M_SPEC)) continue;
        double b1[5], s1[5];
        for (int f = 0; f < nf; ++f) {
            b1[f] = a->host[fvl::S_OUT + f];
            s1[f] = a->host[fvl::NSLOT + fvl::S_MSRC + f];
        }
        if ((rc = stage(a, 2, mid, in, out, dt, first, &flags))) return rc;
        if (flags & (fvl::F_M_UNPHYS | fvl::F_M_SPEC)) continue;
        for (int f = 0; f < nf; ++f) {
            boundary[f] = (0.5 * dt) * (b1[f] + a->host[fvl::S_OUT + f]);
            source[f] = (0.5 * dt) * (s1[f] + a->host[fvl::NSLOT + fvl::S_MSRC + f]);
        }
        *first_used = first;
        return TORCH_OK;
    }
    return fail(TORCH_NONPHYSICAL, "Adaptive update lost positivity");
}

// Chemistry half-step of slot `in` into `out` (torchAdaptive.py:418-434):
// the leaves as one flat array of cells.
int react(torch_amr* a, int in, int out, double dt) {
    overwrite(a, out);
    fvl::React r;
    memset(&r, 0, sizeof(r));
    r.U = a->U[in];
    r.Uout = a->U[out];
    r.lam = a->q[0][in];
    r.lamout = a->q[0][out];
    r.beta = a->nq > 1 ? a->q[1][in] : nullptr;
    r.betaout = a->nq > 1 ? a->q[1][out] : nullptr;
    r.model = a->cfg.kinetic_model;
    r.dt = dt;
    r.q_heat = a->cfg.q_heat;
    r.activation_energy = a->cfg.activation_energy;
    r.pre_exponential = a->cfg.pre_exponential;
    r.flags = a->flags;
    r.partials = region(a, 3);
    HIP_TRY(a->ops->react(flat_grid(a), r, a->ctx->stream));
    return TORCH_OK;
}

__global__ void k_amr_dt(const double* max_rate, double cfl, double time, double t_end, double* dt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double d = cfl / max_rate[0];
        const double rest = t_end - time;
        dt[0] = rest < d ? rest : d;
    }
}

// The common inert step without host round trips (torchAdaptive.py:436-481
// when nothing is retried): the evaluation of the committed state, cfl, dt
// on the device, the first stage on that evaluation, the second stage's
// evaluation and update, one readback. Returns TORCH_OK with the step
// committed, kSlowPath when a stage lost positivity (nothing committed; the
// host-driven step with its retries takes over), or an error.
const int kSlowPath = -1;

int fast_step(torch_amr* a, double time, double t_end, torch_step_info* info) {
    hipStream_t st = a->ctx->stream;
    const int A = a->cur, B = (A + 1) % kSlots, C = (A + 2) % kSlots;
    const int rec = a->order == 2;
    const fvl::Amr mesh = mesh_of(a);
    const int lb = a->ops->amr_blocks(a->nleaf), fb = a->ops->amr_blocks(a->nface);
    double* row[kRows];
    for (int r = 0; r < kRows; ++r) row[r] = a->scalars + r * fvl::NSLOT;
    double* dtp = row[2] + fvl::NSLOT - 1;
    a->cache_slot = -1;
    HIP_TRY(clear(a));
    HIP_TRY(a->ops->amr_eval(a->cfg.dim, mesh, a->U[A], nullptr, nullptr, 0, a->cfg.gamma, 0, rec, a->flags,
                             region(a, 0), st));
    HIP_TRY(fvl::finalize(region(a, 0), fb, row[0], st));
    HIP_TRY(a->ops->amr_cfl(a->cfg.dim, mesh, region(a, 2), st));
    HIP_TRY(fvl::finalize(region(a, 2), lb, row[2], st));
    k_amr_dt<<<1, 64, 0, st>>>(row[2], a->cfg.cfl, time, t_end, dtp);
    HIP_TRY(hipGetLastError());
    for (int k = 0; k < 2; ++k) {
        if (k == 1) {
            HIP_TRY(a->ops->amr_eval(a->cfg.dim, mesh, a->U[B], nullptr, nullptr, 0, a->cfg.gamma, 0, rec,
                                     a->flags + 1, region(a, 0), st));
            HIP_TRY(fvl::finalize(region(a, 0), fb, row[4], st));
        }
        fvl::MStage m;
        memset(&m, 0, sizeof(m));
        m.S = a->U[k == 0 ? A : B];
        m.U0 = a->U[A];
        m.out = a->U[k == 0 ? B : C];
        m.dtp = dtp;
        m.gamma = a->cfg.gamma;
        m.mode = k + 1;
        m.flags = a->flags + k;
        m.partials = region(a, 1);
        HIP_TRY(a->ops->amr_update(a->cfg.dim, mesh, m, st));
        HIP_TRY(fvl::finalize(region(a, 1), lb, row[k == 0 ? 1 : 5], st));
    }
    int rc = sync_read(a);
    if (rc) return rc;
    int flags[2];
    memcpy(flags, a->host + kRows * fvl::NSLOT, sizeof(flags));
    if ((rc = boundary_error(flags[0] | flags[1]))) return rc;
    const double* h = a->host;
    const double dt = h[2 * fvl::NSLOT + fvl::NSLOT - 1];
    if (!isfinite(dt) || dt <= 0 || time + dt == time) return fail(TORCH_STALLED, "Adaptive integration stalled");
    if ((flags[0] | flags[1]) & (fvl::F_M_UNPHYS | fvl::F_M_SPEC)) return kSlowPath;
    memset(info, 0, sizeof(*info));
    info->dt = dt;
    for (int f = 0; f < a->nf; ++f) {
        info->outward[f] = (0.5 * dt) * (h[fvl::S_OUT + f] + h[4 * fvl::NSLOT + fvl::S_OUT + f]);
        info->source[f] = (0.5 * dt) * (h[fvl::NSLOT + fvl::S_MSRC + f] + h[5 * fvl::NSLOT + fvl::S_MSRC + f]);
    }
    a->cur = C;
    return TORCH_OK;
}

int ready(torch_amr* a) {
    if (!a->has_mesh) return fail(TORCH_INVALID, "adaptive solver has no mesh (torch_amr_set_mesh)");
    if (!a->has_state) return fail(TORCH_INVALID, "adaptive solver has no state for this mesh (torch_amr_upload)");
    return TORCH_OK;
}

bool finite_all(const double* p, long n) {
    for (long i = 0; i < n; ++i)
        if (!isfinite(p[i])) return false;
    return true;
}

}  // namespace

extern "C" {

int torch_amr_create(torch_ctx* ctx, const torch_config* cfg, int geometry, const torch_face* faces, int order,
                     torch_amr** out) {
    if (!ctx || !cfg || !faces || !out) return fail(TORCH_INVALID, "null argument");
    const torch_config& c = *cfg;
    if (c.dim < 1 || c.dim > 3) return fail(TORCH_INVALID, "Invalid dimension");
    if (geometry < TORCH_CARTESIAN || geometry > TORCH_SPHERICAL) return fail(TORCH_INVALID, "Unknown geometry");
    if (order != 1 && order != 2) return fail(TORCH_INVALID, "Invalid AMR levels, interval, capacity or order");
    if (!(c.gamma > 1) || !isfinite(c.gamma)) return fail(TORCH_INVALID, "gamma must be finite and greater than one");
    if (!(c.cfl > 0) || !isfinite(c.cfl)) return fail(TORCH_INVALID, "CFL must be positive");
    if (c.kinetic_model < 0 || c.kinetic_model > 1) return fail(TORCH_INVALID, "Unknown limiter or chemistry model");
    if (c.precision != TORCH_FP64 && c.precision != TORCH_FP32) return fail(TORCH_INVALID, "unknown precision");
    if (c.reactive && (!isfinite(c.q_heat) || c.q_heat < 0 || !isfinite(c.activation_energy) ||
                       c.activation_energy < 0 || !isfinite(c.pre_exponential) || c.pre_exponential < 0))
        return fail(TORCH_INVALID, "Reaction time and constants must be finite and nonnegative");
    for (int f = 0; f < 2 * c.dim; ++f) {
        const int t = faces[f].type;
        if (t < TORCH_FACE_OUTFLOW || t > TORCH_FACE_STAGNATION_INLET)
            return fail(TORCH_INVALID, "Unknown face boundary type");
        if ((t == TORCH_FACE_PERIODIC) != (faces[f ^ 1].type == TORCH_FACE_PERIODIC))
            return fail(TORCH_INVALID, "Periodic boundaries must be paired on an axis");
    }
    torch_amr* a = new torch_amr();
    a->ctx = ctx;
    a->cfg = c;
    a->ops = c.precision == TORCH_FP64 ? &kOps64 : &kOps32;
    a->nf = c.dim + 2;
    a->nq = c.reactive ? (c.kinetic_model == 1 ? 2 : 1) : 0;
    a->order = order;
    a->geometry = geometry;
    memset(a->faces, 0, sizeof(a->faces));
    for (int f = 0; f < 2 * c.dim; ++f) {
        fvl::MFace& o = a->faces[f];
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
    std::lock_guard<std::mutex> lock(ctx->mu);
    int rc = TORCH_OK;
    if (hipSetDevice(ctx->device) != hipSuccess) rc = fail(TORCH_HIP_ERROR, "hipSetDevice failed");
    if (!rc) rc = alloc((void**)&a->scalars, kRows * fvl::NSLOT * sizeof(double));
    if (!rc) rc = alloc((void**)&a->flags, 16 * sizeof(int));
    if (!rc && hipHostMalloc((void**)&a->host, (kRows * fvl::NSLOT + 2) * sizeof(double)) != hipSuccess)
        rc = fail(TORCH_HIP_ERROR, "hipHostMalloc failed");
    if (rc) {
        release_buffers(a);
        release_t(&a->scalars);
        release_t(&a->flags);
        if (a->host) (void)hipHostFree(a->host);
        delete a;
        return rc;
    }
    *out = a;
    return TORCH_OK;
}

int torch_amr_destroy(torch_amr* a) {
    if (!a) return TORCH_OK;
    std::lock_guard<std::mutex> lock(a->ctx->mu);
    (void)hipSetDevice(a->ctx->device);
    (void)hipStreamSynchronize(a->ctx->stream);
    release_buffers(a);
    release_t(&a->scalars);
    release_t(&a->flags);
    if (a->host) (void)hipHostFree(a->host);
    delete a;
    return TORCH_OK;
}

int torch_amr_set_mesh(torch_amr* a, long nleaf, long nface, const int* left, const int* right, const int* dir,
                       const int* side, const double* area, const double* weight, const double* offl,
                       const double* offr, const double* disp, const double* volume, const double* div,
                       const double* inverse, long nentry, const long* start, const int* cface, const int* crole) {
    if (!a || !left || !right || !dir || !side || !area || !weight || !offl || !offr || !disp || !volume || !div ||
        !inverse || !start || !cface || !crole)
        return fail(TORCH_INVALID, "null argument");
    const int d = a->cfg.dim;
    if (nleaf < 1 || nface < 1 || nentry < 1 || nleaf > (1L << 30) || nface > (1L << 30) || nentry > (1L << 31) - 1)
        return fail(TORCH_INVALID, "adaptive mesh sizes out of range");
    for (long f = 0; f < nface; ++f) {
        const int k = dir[f], s = side[f], l = left[f], r = right[f];
        if (k < 0 || k >= d) return fail(TORCH_INVALID, "adaptive face direction out of range");
        const bool ok = s == -1   ? (l == -1 && r >= 0 && r < nleaf)
                        : s == 1  ? (r == -1 && l >= 0 && l < nleaf)
                        : s == 0  ? (l >= 0 && l < nleaf && r >= 0 && r < nleaf)
                                  : false;
        if (!ok) return fail(TORCH_INVALID, "adaptive face leaves do not match its side");
        if (s != 0 && a->faces[2 * k + (s > 0)].type == TORCH_FACE_PERIODIC)
            return fail(TORCH_INVALID, "adaptive boundary face on a periodic axis");
    }
    if (!finite_all(area, nface) || !finite_all(weight, nface) || !finite_all(offl, nface * d) ||
        !finite_all(offr, nface * d) || !finite_all(disp, nface * d) || !finite_all(div, nleaf * d) ||
        !finite_all(inverse, nleaf * d * d) || !finite_all(volume, nleaf))
        return fail(TORCH_INVALID, "adaptive mesh arrays must be finite");
    for (long i = 0; i < nleaf; ++i)
        if (!(volume[i] > 0)) return fail(TORCH_INVALID, "adaptive leaf volumes must be positive");
    if (start[0] != 0 || start[nleaf] != nentry) return fail(TORCH_INVALID, "adaptive incidence starts do not match");
    // Every entry names a face of its leaf in that role, and every side of
    // every face that has a leaf appears exactly once.
    std::vector<unsigned char> seen(nface, 0);
    for (long i = 0; i < nleaf; ++i) {
        if (start[i + 1] < start[i]) return fail(TORCH_INVALID, "adaptive incidence starts must not decrease");
        for (long e = start[i]; e < start[i + 1]; ++e) {
            const int f = cface[e], role = crole[e];
            if (f < 0 || f >= nface || (role != -1 && role != 1) || (role < 0 ? left[f] : right[f]) != i)
                return fail(TORCH_INVALID, "adaptive incidence entry does not match the face list");
            const unsigned char bit = role < 0 ? 1 : 2;
            if (seen[f] & bit) return fail(TORCH_INVALID, "adaptive incidence lists a face side twice");
            seen[f] |= bit;
        }
    }
    for (long f = 0; f < nface; ++f)
        if (seen[f] != ((left[f] >= 0 ? 1 : 0) | (right[f] >= 0 ? 2 : 0)))
            return fail(TORCH_INVALID, "adaptive incidence misses a face side");
    std::lock_guard<std::mutex> lock(a->ctx->mu);
    HIP_TRY(hipSetDevice(a->ctx->device));
    HIP_TRY(hipStreamSynchronize(a->ctx->stream));
    a->has_mesh = a->has_state = false;
    int rc = reserve(a, nleaf, nface, nentry);
    if (rc) return rc;
    hipStream_t st = a->ctx->stream;
    auto put = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    HIP_TRY(put(a->left, left, nface * sizeof(int)));
    HIP_TRY(put(a->right, right, nface * sizeof(int)));
    HIP_TRY(put(a->dir, dir, nface * sizeof(int)));
    HIP_TRY(put(a->side, side, nface * sizeof(int)));
    HIP_TRY(put(a->area, area, nface * sizeof(double)));
    HIP_TRY(put(a->weight, weight, nface * sizeof(double)));
    HIP_TRY(put(a->offl, offl, nface * d * sizeof(double)));
    HIP_TRY(put(a->offr, offr, nface * d * sizeof(double)));
    HIP_TRY(put(a->disp, disp, nface * d * sizeof(double)));
    HIP_TRY(put(a->volume, volume, nleaf * sizeof(double)));
    HIP_TRY(put(a->div, div, nleaf * d * sizeof(double)));
    HIP_TRY(put(a->inverse, inverse, nleaf * d * d * sizeof(double)));
    HIP_TRY(put(a->start, start, (nleaf + 1) * sizeof(long)));
    HIP_TRY(put(a->cface, cface, nentry * sizeof(int)));
    HIP_TRY(put(a->crole, crole, nentry * sizeof(int)));
    HIP_TRY(hipStreamSynchronize(st));
    a->nleaf = nleaf;
    a->nface = nface;
    a->nentry = nentry;
    a->cur = 0;
    a->cache_slot = -1;
    a->has_mesh = true;
    return TORCH_OK;
}

int torch_amr_upload(torch_amr* a, const double* U, const double* lam, const double* beta) {
    if (!a || !U) return fail(TORCH_INVALID, "null argument");
    if (!a->has_mesh) return fail(TORCH_INVALID, "adaptive solver has no mesh (torch_amr_set_mesh)");
    if (a->nq > 0 && !lam) return fail(TORCH_INVALID, "reactive leaves need a lambda array");
    std::lock_guard<std::mutex> lock(a->ctx->mu);
    HIP_TRY(hipSetDevice(a->ctx->device));
    hipStream_t st = a->ctx->stream;
    overwrite(a, a->cur);
    int rc = upload_as(a->ops, st, a->U[a->cur], U, a->nf * a->nleaf);
    if (!rc && a->nq > 0) rc = upload_as(a->ops, st, a->q[0][a->cur], lam, a->nleaf);
    if (!rc && a->nq > 1) {
        if (beta) {
            rc = upload_as(a->ops, st, a->q[1][a->cur], beta, a->nleaf);
        } else {
            std::vector<double> ones(a->nleaf, 1.0);
            rc = upload_as(a->ops, st, a->q[1][a->cur], ones.data(), a->nleaf);
        }
    }
    a->has_state = rc == TORCH_OK;
    return rc;
}

int torch_amr_download(torch_amr* a, double* U, double* lam, double* beta) {
    if (!a) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(a->ctx->mu);
    int rc = ready(a);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(a->ctx->device));
    hipStream_t st = a->ctx->stream;
    if (U) rc = download_as(a->ops, st, U, a->U[a->cur], a->nf * a->nleaf);
    if (!rc && lam && a->nq > 0) rc = download_as(a->ops, st, lam, a->q[0][a->cur], a->nleaf);
    if (!rc && beta && a->nq > 1) rc = download_as(a->ops, st, beta, a->q[1][a->cur], a->nleaf);
    return rc;
}

int torch_amr_faces(torch_amr* a, double* left, double* right, double* gl, double* gr) {
    if (!a) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(a->ctx->mu);
    int rc = ready(a);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(a->ctx->device));
    HIP_TRY(clear(a));
    if ((rc = evaluate(a, a->cur, 0))) return rc;
    if ((rc = sync_read(a))) return rc;
    if ((rc = boundary_error(host_flags(a)))) return rc;
    remember(a, a->cur, 0);
    hipStream_t st = a->ctx->stream;
    const long F = a->nface;
    const size_t e = a->ops->elem;
    if (left && (rc = download_as(a->ops, st, left, a->FL, a->nf * F))) return rc;
    if (right && (rc = download_as(a->ops, st, right, a->FR, a->nf * F))) return rc;
    if (gl && (rc = download_as(a->ops, st, gl, a->G, F))) return rc;
    if (gr && (rc = download_as(a->ops, st, gr, (const char*)a->G + F * e, F))) return rc;
    return TORCH_OK;
}

int torch_amr_dt(torch_amr* a, double* dt) {
    if (!a || !dt) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(a->ctx->mu);
    int rc = ready(a);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(a->ctx->device));
    return leaf_dt(a, a->cur, dt);
}

int torch_amr_rhs(torch_amr* a, int first_order, double* change, double* species, double* outward, double* source) {
    if (!a) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(a->ctx->mu);
    int rc = ready(a);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(a->ctx->device));
    const int A = a->cur, B = (A + 1) % kSlots;
    int flags = 0;
    if ((rc = stage(a, 0, A, A, B, 0.0, first_order != 0, &flags))) return rc;
    for (int f = 0; f < a->nf; ++f) {
        if (outward) outward[f] = a->host[fvl::S_OUT + f];
        if (source) source[f] = a->host[fvl::NSLOT + fvl::S_MSRC + f];
    }
    hipStream_t st = a->ctx->stream;
    if (change && (rc = download_as(a->ops, st, change, a->U[B], a->nf * a->nleaf))) return rc;
    for (int j = 0; j < a->nq && species; ++j)
        if ((rc = download_as(a->ops, st, species + j * a->nleaf, a->q[j][B], a->nleaf))) return rc;
    return TORCH_OK;
}

int torch_amr_advance(torch_amr* a, double dt, torch_step_info* info) {
    if (!a || !info) return fail(TORCH_INVALID, "null argument");
    if (!isfinite(dt) || dt < 0) return fail(TORCH_INVALID, "dt must be finite and nonnegative");
    std::lock_guard<std::mutex> lock(a->ctx->mu);
    memset(info, 0, sizeof(*info));
    int rc = ready(a);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(a->ctx->device));
    const int A = a->cur, B = (A + 1) % kSlots, C = (A + 2) % kSlots;
    if ((rc = advance(a, A, B, C, dt, info->outward, info->source, &info->fallback))) return rc;
    info->dt = dt;
    a->cur = C;
    return TORCH_OK;
}

int torch_amr_step(torch_amr* a, double time, double t_end, torch_step_info* info) {
    if (!a || !info) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(a->ctx->mu);
    memset(info, 0, sizeof(*info));
    int rc = ready(a);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(a->ctx->device));
    const bool reactive = a->cfg.reactive != 0;
    if (!reactive) {
        rc = fast_step(a, time, t_end, info);
        if (rc != kSlowPath) return rc;
        memset(info, 0, sizeof(*info));
    }
    const int A = a->cur;
    const int B = (A + 1) % kSlots, C = (A + 2) % kSlots, D = (A + 3) % kSlots;
    hipStream_t st = a->ctx->stream;
    double dt;
    if ((rc = leaf_dt(a, A, &dt))) return rc;
    dt = py_min(dt, t_end - time);
    for (int retry = 0; retry < 24; ++retry) {
        if (!isfinite(dt) || dt <= 0 || time + dt == time) return fail(TORCH_STALLED, "Adaptive integration stalled");
        int old = A;
        if (reactive) {
            HIP_TRY(clear(a));
            if ((rc = react(a, A, B, dt / 2))) return rc;
            if ((rc = sync_read(a))) return rc;
            if (host_flags(a) & fvl::F_REACT_BAD)
                return fail(TORCH_INVALID, "Reaction input outside its domain (state, lambda or beta)");
            double allowed;
            if ((rc = leaf_dt(a, B, &allowed))) return rc;
            if (dt > allowed * (1 + 1e-12)) {
                dt = py_min(dt / 2, allowed);
                info->retries += 1;
                continue;
            }
            old = B;
        }
        // inert: stage B, hydro C; reactive: stage C, hydro D, final C
        const int mid = reactive ? C : B, hydro = reactive ? D : C;
        double boundary[5] = {0, 0, 0, 0, 0}, source[5] = {0, 0, 0, 0, 0};
        int first = 0;
        rc = advance(a, old, mid, hydro, dt, boundary, source, &first);
        if (rc == TORCH_NONPHYSICAL) {
            dt /= 2;
            info->retries += 1;
            continue;
        }
        if (rc) return rc;
        int final_slot = hydro;
        if (reactive) {
            HIP_TRY(clear(a));
            if ((rc = react(a, hydro, C, dt / 2))) return rc;
            HIP_TRY(a->ops->amr_source(a->cfg.dim, mesh_of(a), a->U[A], a->U[B], a->U[C], a->U[hydro],
                                       region(a, 1), st));
            HIP_TRY(fvl::finalize(region(a, 1), a->ops->amr_blocks(a->nleaf), a->scalars + fvl::NSLOT, st));
            if ((rc = sync_read(a))) return rc;
            const int flags = host_flags(a);
            if (flags & fvl::F_REACT_BAD) return fail(TORCH_INVALID, "Reaction input outside its domain");
            if (flags & fvl::F_STATE_BAD) return fail(TORCH_REACTION_STATE, "Nonphysical adaptive reaction");
            const int e = a->nf - 1;
            source[e] = source[e] + a->host[fvl::NSLOT + fvl::S_MSRC + e];
            final_slot = C;
        }
        info->dt = dt;
        info->fallback = first;
        for (int f = 0; f < a->nf; ++f) {
            info->outward[f] = boundary[f];
            info->source[f] = source[f];
        }
        a->cur = final_slot;
        return TORCH_OK;
    }
    return fail(TORCH_NONPHYSICAL, "Could not find a positive adaptive step");
}

}  // extern "C"
