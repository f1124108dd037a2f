// Batch bookkeeping kernels: the per-instance logic of Simulation.step
// (python/cfdSolver/torchSession.py:374-522) done on the device, so a step of
// every instance of a batch is a fixed sequence of launches that a HIP graph
// can replay.
//
// k_batch_check (reactive instances, after the first chemistry half-step):
// the CFL test of the heated state and the dt halving of the retry loop
// (:398-411), on the maxima folded like the post kernel's (below). A replay holds a few attempt/check pairs; an accepted attempt
// opens the hydro gate and later attempts of the replay skip the instance.
// An instance still rejected after the last pair commits nothing in this
// replay and continues the loop in the next one from the same committed
// buffers, so the arithmetic is the loop's.
//
// k_batch_post (every instance, last): folds the instance's per-block
// partials in the order of the single-session finalize, checks the step's
// flags, commits (time, step count, boundary and source integrals, buffer
// rotation) and derives the next dt and step coefficients from the new wave
// maxima. An instance gets kPostSlots blocks, one per partial slot, so the
// folds run side by side on as many CUs; each block stores its slot's total
// in BatchCtl.folded and takes a ticket, and the block that takes the last
// ticket commits.
#include "fv_device.h"
#include "fv_launch.h"

#include "../../include/torch_hip.h"

namespace FV_NS {
using namespace fv;

// dt from the CFL maxima in the reference's order of operations (the same
// expressions as the host driver's): 3D torchSession.py:356-360, 2D
// :346-355, 1D torchSolver.py:684-689 with the 1e-6 default.
__device__ __forceinline__ double dt_of(const fvl::BatchCtl& c, const double* mx) {
    if (c.dim == 3) {
        double sum = mx[0] / c.h[0];
        sum = sum + mx[1] / c.h[1];
        sum = sum + mx[2] / c.h[2];
        return c.cfl / sum;
    }
    if (c.dim == 2) {
        double total = mx[0] / c.h[0] + mx[1] / c.h[1];
        if (total <= 1e-12) return 1e-3;
        return c.cfl / total;
    }
    double a = mx[0] > 0 ? (c.cfl * c.h[0]) / mx[0] : 1e-6;
    double b = c.cfl * c.h[0] / mx[1];
    return b < a ? b : a;
}

__device__ __forceinline__ double py_min(double a, double b) { return b < a ? b : a; }

// One slot of nb per-block partials folded by a 256-thread block in the
// single-session finalize's order (k_finalize): lane t accumulates the rows
// t, t+256, ... in increasing order, then a tree over the 256 lanes.
// Every thread of the block must call it; every thread gets the total.
__device__ double fold_slot(const double* part, int nb, int slot, double* buf) {
    const int t = threadIdx.x;
    const bool is_max = slot < 3;
    double v = 0.0;
#pragma unroll 8
    for (int b = t; b < nb; b += 256) {
        const double p = part[(long)b * fvl::NSLOT + slot];
        v = is_max ? fmax(v, p) : v + p;
    }
    buf[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            double a = buf[t], b = buf[t + s];
            buf[t] = is_max ? fmax(a, b) : a + b;
        }
        __syncthreads();
    }
    return buf[0];
}

// Step coefficients for dt (the host driver's base_step expressions).
__device__ __forceinline__ void set_coef(fvl::BatchCtl& c, double dt) {
    for (int d = 0; d < 3; ++d) {
        c.coef[d] = dt / c.h[d];
        c.coef[3 + d] = 0.5 * dt / c.h[d];
        c.coef[6 + d] = dt * c.volume / c.h[d];
    }
}

__device__ __forceinline__ void stop(fvl::BatchCtl& c, int status) {
    c.status = status;
    c.active = 0;
    c.coef[fvl::K_HYDRO] = 0.0;
    c.coef[fvl::K_CHEM] = 0.0;
}

// Three blocks per instance, one per wave-speed maximum, with the post
// kernel's ticket: the block that takes the last ticket decides.
__global__ void __launch_bounds__(256) k_batch_check(fvl::BatchItem* items, fvl::BatchCtl* ctl) {
#pragma clang fp contract(off)
    __shared__ double buf[256];
    __shared__ int last;
    const int i = blockIdx.x / 3, k = blockIdx.x % 3, t = threadIdx.x;
    fvl::BatchCtl& c = ctl[i];
    // Stopped, or this step's dt already accepted by an earlier attempt
    // (every block of the instance takes the same exit).
    if (c.active != 1 || c.coef[fvl::K_CHEM] == 0.0 || c.coef[fvl::K_HYDRO] != 0.0) return;
    fvl::BatchItem& it = items[i];
    const double v = fold_slot(it.step.partials + (long)it.r1_first * fvl::NSLOT, c.nreact, k, buf);
    if (t == 0) {
        c.folded[k] = v;
        last = __hip_atomic_fetch_add(&c.ticket, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 2;
    }
    __syncthreads();
    if (!last || t != 0) return;
    c.ticket = 0;
    int* flags = it.step.flags;
    if (*flags & fvl::F_REACT_BAD) {
        stop(c, TORCH_INVALID);
        return;
    }
    const double mx[3] = {c.folded[0], c.folded[1], c.folded[2]};
    const double dt = c.dt, allowed = dt_of(c, mx);
    *flags = 0;  // the hydro pass (or the next attempt) starts from clear flags
    if (dt <= allowed * (1 + 1e-12)) {
        set_coef(c, dt);
        c.coef[fvl::K_HYDRO] = 1.0;
        c.retries += c.retry;
        return;
    }
    c.retry += 1;
    if (c.retry == 60) {
        stop(c, TORCH_NO_STABLE_STEP);
        return;
    }
    const double next = py_min(dt / 2, allowed);
    c.dt = next;
    c.coef[fvl::K_HALF] = next / 2;
    c.coef[fvl::K_HYDRO] = 0.0;
}

__global__ void __launch_bounds__(256) k_batch_post(fvl::BatchItem* items, fvl::BatchCtl* ctl, int prime) {
#pragma clang fp contract(off)
    __shared__ double buf[256];
    __shared__ int last;
    const int i = blockIdx.x / fvl::kPostSlots, k = blockIdx.x % fvl::kPostSlots, t = threadIdx.x;
    fvl::BatchCtl& c = ctl[i];
    // Every block of an instance takes the same exit here (nothing below
    // changes these fields before the last ticket is taken).
    if (prime ? c.active != 2 : c.active != 1) return;
    // A reactive attempt rejected by k_batch_check: nothing to commit yet.
    if (!prime && c.reactive && c.coef[fvl::K_HYDRO] == 0.0) return;
    fvl::BatchItem& it = items[i];
    const int flags = prime ? 0 : *it.step.flags;
    const bool fb = (flags & (fvl::F_PRED_BAD | fvl::F_UPD_BAD)) != 0;
    // Slot k's pass: the CFL pass (first dt, 1D), else the second chemistry
    // half-step for the maxima and the source integral of a reactive
    // instance, else the hydro pass whose state was kept.
    const double* part = it.step.partials;
    long first = -1;
    int nb = 0;
    if (prime || c.dim == 1) {
        if (k < 3) first = it.cfl_first, nb = c.ncfl;
    } else if (c.reactive && (k < 3 || k == fvl::S_SRC)) {
        first = it.r2_first, nb = c.nreact;
    } else if (k < fvl::S_SRC) {
        first = fb ? it.llf_first : 0, nb = fb ? c.nllf : c.nblk;
    }
    if (first >= 0) {
        const double v = fold_slot(part + first * fvl::NSLOT, nb, k, buf);
        if (t == 0) c.folded[k] = v;
    }
    if (t == 0)
        last = __hip_atomic_fetch_add(&c.ticket, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
               fvl::kPostSlots - 1;
    __syncthreads();
    if (!last || t != 0) return;
    c.ticket = 0;
    double mx[3], out[5] = {0, 0, 0, 0, 0}, src = 0;
    for (int m = 0; m < 3; ++m) mx[m] = c.folded[m];
    if (!prime && c.dim >= 2) {
        for (int j = 0; j < 5; ++j) out[j] = c.folded[fvl::S_OUT + j];
        if (c.reactive) src = c.folded[fvl::S_SRC];
    }
    if (c.dim == 2 && it.grid.bc[1] == BC_INFLOW) {
        // The inlet ghost rows belong to the padded state the 2D CFL covers.
        real u[4];
        for (int k = 0; k < 4; ++k) u[k] = (real)it.grid.inlet[k];
        Prim2 p = prim2(u, (real)1.4);
        mx[0] = fmax(mx[0], (double)(fabs(p.ux) + p.a));
        mx[1] = fmax(mx[1], (double)(fabs(p.vy) + p.a));
    }
    if (!prime) {
        // Error order of the host driver (torch_state_step); the committed
        // state stays as it was.
        if (flags & fvl::F_LLF_BAD) return stop(c, TORCH_NONPHYSICAL);
        if (c.dim == 3 && (flags & (fb ? fvl::F_SPEC_BAD_LLF : fvl::F_SPEC_BAD))) return stop(c, TORCH_SPECIES_RANGE);
        if (flags & fvl::F_REACT_BAD) return stop(c, TORCH_INVALID);
        if (flags & fvl::F_STATE_BAD) return stop(c, TORCH_REACTION_STATE);
        const int nf = c.dim + 2;
        if (c.dim == 1) {
            const real* Fs = (const real*)it.scratch;
            const int n = it.grid.nx;
            for (int k = 0; k < 3; ++k)
                out[k] = c.coef[6] * (double)(Fs[k * (n + 1) + n] - Fs[k * (n + 1)]);
        }
        for (int k = 0; k < nf; ++k) c.outward[k] = c.outward[k] + out[k];
        if (c.reactive) c.source[nf - 1] = c.source[nf - 1] + c.volume * src;
        c.time = py_min(c.time + c.dt, c.t_end);
        c.steps += 1;
        c.fallbacks += fb ? 1 : 0;
        if (c.reactive) {
            c.cur = (c.cur + 2) % fvl::kReactSlots;  // C of this step is A of the next
            fvl::set_reactive_slots(it, c.cur);
        } else {
            const void* u = it.step.U;
            it.step.U = it.step.Uout;
            it.step.Uout = const_cast<void*>(u);
            c.cur ^= 1;
        }
    }
    for (int m = 0; m < 3; ++m) c.maxima[m] = mx[m];
    if (!prime && (c.time == c.t_end || c.steps >= c.max_steps)) return stop(c, TORCH_OK);
    const double dt = py_min(dt_of(c, mx), c.t_end - c.time);
    if (!isfinite(dt) || dt <= 0 || c.time + dt == c.time) return stop(c, TORCH_STALLED);
    c.dt = dt;
    c.active = 1;
    c.retry = 0;
    *it.step.flags = 0;
    if (c.reactive) {
        // The first half-step decides whether dt stands (k_batch_check).
        c.coef[fvl::K_HALF] = dt / 2;
        c.coef[fvl::K_CHEM] = 1.0;
        c.coef[fvl::K_HYDRO] = 0.0;
    } else {
        set_coef(c, dt);
        c.coef[fvl::K_HYDRO] = 1.0;
    }
}

hipError_t batch_post(fvl::BatchItem* items, fvl::BatchCtl* ctl, int n, int prime, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_batch_post<<<n * fvl::kPostSlots, 256, 0, st>>>(items, ctl, prime);
    return hipGetLastError();
}

hipError_t batch_check(fvl::BatchItem* items, fvl::BatchCtl* ctl, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_batch_check<<<n * 3, 256, 0, st>>>(items, ctl);
    return hipGetLastError();
}

}  // namespace FV_NS
