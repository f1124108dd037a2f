// Host-side plumbing shared by the C-ABI translation units (capi.hip,
// batch.hip): error reporting, the per-precision kernel tables and the
// device context.
#pragma once

#include <mutex>
#include <string>

#include "../../include/torch_hip.h"
#include "fv_launch.h"

namespace tcapi {

// Records msg as the calling thread's last error and returns code.
int fail(int code, const std::string& msg);
// hipMalloc with the error mapped to TORCH_HIP_ERROR.
int alloc(void** p, size_t bytes);
// Validation of a session configuration (the reference's ValueErrors).
int check_config(const torch_config& c);
// Grid geometry, cell volume and 2D inlet state of a configuration.
void make_grid(const torch_config& c, fvl::Grid* g, double* volume);

struct Ops;
// Host float64 <-> device storage of the precision build (synchronous).
int upload_as(const Ops* ops, hipStream_t st, void* dst, const double* src, long n);
int download_as(const Ops* ops, hipStream_t st, double* dst, const void* src, long n);

// Default z chunk of the 3D step for a grid nz planes deep and a tile
// block_y threads tall, shared by sessions and batches so that a batch of
// one folds its partials in the same blocks as the session: 24 planes where
// the grid is deep enough to fill the chip with such chunks (fastest in a
// 10-32 plane sweep of the 256^3 Sod step in both precisions), else 16;
// the 32x24 FP32 tile (one block per CU) takes 24 at every depth (faster
// than 16 on the 32^3..160^3 ignition mix).
inline int default_zchunk(int nz, int block_y) { return (nz >= 128 || block_y == 24) ? 24 : 16; }

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::tcapi::fail(TORCH_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Kernel entry points of one precision build.
struct Ops {
    size_t elem;
    int (*step3d_blocks)(const fvl::Grid&, const fvl::Launch&, int);
    hipError_t (*step3d)(const fvl::Grid&, const fvl::Step&, const fvl::Launch&);
    hipError_t (*llf3d)(const fvl::Grid&, const fvl::Step&, hipStream_t);
    int (*llf3d_blocks)(const fvl::Grid&);
    hipError_t (*step2d)(const fvl::Grid&, const fvl::Step&, const fvl::Launch&);
    int (*step2d_blocks)(const fvl::Grid&, const fvl::Launch&);
    hipError_t (*llf2d)(const fvl::Grid&, const fvl::Step&, hipStream_t);
    int (*llf2d_blocks)(const fvl::Grid&);
    hipError_t (*step1d)(const fvl::Grid&, const fvl::Step&, void*, hipStream_t);
    size_t (*step1d_scratch)(const fvl::Grid&);
    int (*step1d_blocks)(const fvl::Grid&);
    hipError_t (*cfl_cells)(const fvl::Grid&, const void*, const void*, double, double*, int, const void*,
                            hipStream_t);
    int (*cfl_blocks)(const fvl::Grid&);
    hipError_t (*react)(const fvl::Grid&, const fvl::React&, hipStream_t);
    int (*react_blocks)(const fvl::Grid&);
    hipError_t (*hllc_faces)(int, int, long, const void*, const void*, const void*, const void*, void*,
                             hipStream_t);
    hipError_t (*totals)(const fvl::Grid&, const void*, double*, hipStream_t);
    hipError_t (*convert_in)(const double*, void*, long, hipStream_t);
    hipError_t (*convert_out)(const void*, double*, long, hipStream_t);
    // batches
    hipError_t (*batch3d)(const fvl::BatchItem*, const int*, int, int, int, int, int, int, hipStream_t);
    hipError_t (*batch_llf3d)(const fvl::BatchItem*, const int*, int, int, hipStream_t);
    hipError_t (*batch_cfl3d)(const fvl::BatchItem*, const int*, int, int, hipStream_t);
    hipError_t (*batch2d)(const fvl::BatchItem*, const int*, int, int, int, hipStream_t);
    hipError_t (*batch_llf2d)(const fvl::BatchItem*, const int*, int, int, hipStream_t);
    hipError_t (*batch_cfl2d)(const fvl::BatchItem*, const int*, int, int, hipStream_t);
    hipError_t (*batch1d)(const fvl::BatchItem*, const int*, int, int, int, hipStream_t);
    hipError_t (*batch_llf1d)(const fvl::BatchItem*, const int*, int, int, hipStream_t);
    hipError_t (*batch_cfl1d)(const fvl::BatchItem*, const int*, int, int, int, hipStream_t);
    hipError_t (*batch_post)(fvl::BatchItem*, fvl::BatchCtl*, int, int, hipStream_t);
    hipError_t (*batch_react)(const fvl::BatchItem*, const int*, int, int, int, hipStream_t);
    hipError_t (*batch_check)(fvl::BatchItem*, fvl::BatchCtl*, int, hipStream_t);
    int (*strip2d_rows)(const fvl::Grid&);
    int (*step1d_tiles)(const fvl::Grid&);
    // metric/face scheme
    hipError_t (*metric_stage)(const fvl::Grid&, const fvl::Metric&, const fvl::MStage&, hipStream_t);
    hipError_t (*metric_cfl)(const fvl::Grid&, const fvl::Metric&, const void*, const void*, double, int, double*,
                             int*, const void*, hipStream_t);
    hipError_t (*metric_source)(const fvl::Grid&, const fvl::Metric&, const void*, const void*, const void*,
                                const void*, double*, hipStream_t);
    hipError_t (*rusanov_faces)(int, int, long, const void*, const void*, const void*, const void*, void*,
                                hipStream_t);
    int (*metric_blocks)(const fvl::Grid&);
    // adaptive leaves
    hipError_t (*amr_eval)(int, const fvl::Amr&, const void*, const void*, const void*, int, double, int, int, int*,
                           double*, hipStream_t);
    hipError_t (*amr_update)(int, const fvl::Amr&, const fvl::MStage&, hipStream_t);
    hipError_t (*amr_cfl)(int, const fvl::Amr&, double*, hipStream_t);
    hipError_t (*amr_source)(int, const fvl::Amr&, const void*, const void*, const void*, const void*, double*,
                             hipStream_t);
    int (*amr_blocks)(long);
};

extern Ops kOps64;
extern Ops kOps32;

}  // namespace tcapi

struct torch_ctx {
    int device;
    hipStream_t stream;
    std::mutex mu;
};
