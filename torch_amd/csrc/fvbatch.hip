// Batch post kernel: the per-instance bookkeeping of Simulation.step
// (python/cfdSolver/torchSession.py:374-522) done on the device, so a step of
// every instance of a batch is a fixed sequence of launches that a HIP graph
// can replay. One block per instance folds that instance's per-block
// partials (maxima and boundary sums, in the same order as the
// single-session finalize), checks the step's flags, commits (time,
// step count, boundary integral, buffer swap) and derives the next dt and
// step coefficients from the new wave maxima.
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

__global__ void __launch_bounds__(256) k_batch_post(fvl::BatchItem* items, fvl::BatchCtl* ctl, int prime) {
#pragma clang fp contract(off)
    __shared__ double red[8][256];
    const int i = blockIdx.x, t = threadIdx.x;
    fvl::BatchCtl& c = ctl[i];
    if (prime ? c.active != 2 : c.active != 1) return;
    fvl::BatchItem& it = items[i];
    const int flags = prime ? 0 : *it.step.flags;
    const bool fb = (flags & (fvl::F_PRED_BAD | fvl::F_UPD_BAD)) != 0;
    int first = 0, nb = c.nblk;
    if (prime || c.dim == 1) {
        first = it.cfl_first;
        nb = c.ncfl;
    } else if (fb) {
        first = it.llf_first;
        nb = c.nllf;
    }
    const double* part = it.step.partials + (long)first * fvl::NSLOT;
    for (int slot = 0; slot < 8; ++slot) {
        const bool is_max = slot < 3;
        double v = 0.0;
        for (int b = t; b < nb; b += 256) {
            double p = part[(long)b * fvl::NSLOT + slot];
            v = is_max ? fmax(v, p) : v + p;
        }
        red[slot][t] = v;
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s)
            for (int slot = 0; slot < 8; ++slot) {
                double a = red[slot][t], b = red[slot][t + s];
                red[slot][t] = slot < 3 ? fmax(a, b) : a + b;
            }
        __syncthreads();
    }
    if (t != 0) return;
    double mx[3] = {red[0][0], red[1][0], red[2][0]};
    if (c.dim == 2 && it.grid.bc[1] == BC_INFLOW) {
        // The inlet ghost rows belong to the padded state the 2D CFL covers.
        real u[4];
        for (int k = 0; k < 4; ++k) u[k] = (real)it.grid.inlet[k];
        Prim2 p = prim2(u, (real)1.4);
        mx[0] = fmax(mx[0], (double)(fabs(p.ux) + p.a));
        mx[1] = fmax(mx[1], (double)(fabs(p.vy) + p.a));
    }
    if (!prime) {
        if (flags & fvl::F_LLF_BAD) {
            c.status = TORCH_NONPHYSICAL;  // the committed state stays as it was
            c.active = 0;
            c.coef[9] = 0.0;
            return;
        }
        const int nf = c.dim + 2;
        double out[5] = {0, 0, 0, 0, 0};
        if (c.dim == 1) {
            const real* Fs = (const real*)it.scratch;
            const int n = it.grid.nx;
            for (int k = 0; k < 3; ++k) out[k] = c.coef[6] * (double)(Fs[k * (n + 1) + n] - Fs[k * (n + 1)]);
        } else {
            for (int k = 0; k < nf; ++k) out[k] = red[fvl::S_OUT + k][0];
        }
        for (int k = 0; k < nf; ++k) c.outward[k] = c.outward[k] + out[k];
        c.time = py_min(c.time + c.dt, c.t_end);
        c.steps += 1;
        c.fallbacks += fb ? 1 : 0;
        const void* u = it.step.U;
        it.step.U = it.step.Uout;
        it.step.Uout = const_cast<void*>(u);
        c.cur ^= 1;
    }
    for (int m = 0; m < 3; ++m) c.maxima[m] = mx[m];
    if (!prime && (c.time == c.t_end || c.steps >= c.max_steps)) {
        c.active = 0;
        c.coef[9] = 0.0;
        return;
    }
    const double dt = py_min(dt_of(c, mx), c.t_end - c.time);
    if (!isfinite(dt) || dt <= 0 || c.time + dt == c.time) {
        c.status = TORCH_STALLED;
        c.active = 0;
        c.coef[9] = 0.0;
        return;
    }
    c.dt = dt;
    for (int d = 0; d < 3; ++d) {
        c.coef[d] = dt / c.h[d];
        c.coef[3 + d] = 0.5 * dt / c.h[d];
        c.coef[6 + d] = dt * c.volume / c.h[d];
    }
    c.coef[9] = 1.0;
    c.active = 1;
    *it.step.flags = 0;
}

hipError_t batch_post(fvl::BatchItem* items, fvl::BatchCtl* ctl, int n, int prime, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_batch_post<<<n, 256, 0, st>>>(items, ctl, prime);
    return hipGetLastError();
}

}  // namespace FV_NS
