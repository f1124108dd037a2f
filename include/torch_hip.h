/* C ABI of libtorchhip.so: TORCH's Cartesian finite-volume step on MI355X.
 *
 * The reference (BFG508/TORCH, python/) has no native interface; its drop-in
 * boundary for this path is the CPU flux-worker backend
 * `Execution(workers).map(fn, batches)` (python/cfdSolver/torchExecution.py:17-68)
 * that `hllc_flux_grid` (torchFlux.py:66-74), `_hydro` (torchSolver3D.py:182-189),
 * `muscl_hancock_step_2d` (torchSolver2D.py:388-410) and the 1D HLLC
 * (torchSolver.py:542-560) hand their face batches to, under the step driver
 * `Simulation.step` (torchSession.py:374-522). Each entry point below names
 * the reference interface it replaces. All arrays crossing the boundary are
 * host float64 in the reference layout: C order (dim+2, nz, ny, nx), x fastest.
 *
 * Every call returns a status (TORCH_OK, or one of the codes below); the
 * message of the last failure on the calling thread is available from
 * torch_last_error. No C++ exception crosses this boundary. A context
 * serialises its calls with one mutex; each state owns its device buffers.
 */
#ifndef TORCH_HIP_H
#define TORCH_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. Python maps 2, 5, 6, 7, 8 to RuntimeError, 3 to ValueError. */
#define TORCH_OK 0
#define TORCH_NONPHYSICAL 2       /* "...hydrodynamic update produced a nonphysical state" */
#define TORCH_INVALID 3           /* invalid argument or chemistry input outside its domain */
#define TORCH_HIP_ERROR 4         /* device/runtime failure */
#define TORCH_STALLED 5           /* "Integration stalled before requested final time" */
#define TORCH_NO_STABLE_STEP 6    /* "Could not find stable reactive step" */
#define TORCH_SPECIES_RANGE 7     /* "3D scalar transport left the physical interval" */
#define TORCH_REACTION_STATE 8    /* "Reaction produced a nonphysical state" */
#define TORCH_BUDGET 9            /* a batch instance does not fit the batch's HBM budget */

/* Enumerations shared with the host mirror (torch_amd/_abi.py). */
enum torch_limiter { TORCH_VANLEER = 0, TORCH_ALBADA = 1, TORCH_MINBEE = 2, TORCH_SUPERBEE = 3, TORCH_UNOS = 4 };
enum torch_boundary { TORCH_OPEN = 0, TORCH_PERIODIC = 1, TORCH_REFLECTIVE = 2, TORCH_INFLOW = 3 };
enum torch_precision { TORCH_FP64 = 0, TORCH_FP32 = 1 };

typedef struct torch_ctx torch_ctx;
typedef struct torch_state torch_state;

/* Session configuration: the numeric subset of Simulation.config
 * (torchSession.py:66-92) for the Cartesian kernels. */
typedef struct torch_config {
    int dim;                   /* 1, 2 or 3 */
    int n[3];                  /* counts x, y, z (unused axes 1) */
    double lengths[3];         /* lengths x, y, z */
    double gamma;              /* inert gamma */
    int limiter;               /* enum torch_limiter */
    int bc[3];                 /* enum torch_boundary per axis x, y, z */
    double inlet_primitive[4]; /* 2D Inflow [rho, u, v, p] */
    int reactive;              /* 0/1 */
    int kinetic_model;         /* 0 arrhenius, 1 znd */
    double q_heat, activation_energy, pre_exponential;
    int controller;            /* 1D: 0 "Sound Speed Controlled", 1 "w=0" */
    int variant;               /* 1D: 0 "Variante 1 - Original", 1 "Variante_2" */
    int wavespeed;             /* 1D: 0 "Pressure Based", 1 "Enthalpy Based" */
    double cfl;
    int precision;             /* enum torch_precision */
    int block_x, block_y, zchunk; /* 3D tile (64x4, 64x8, 32x8, 32x16; 32x24 in
                                     FP32 only); 0 selects the tuned default */
    double spacing[3];         /* cell sizes x, y, z; 0 derives lengths / n */
    int gamma_mode;            /* 0 scalar gamma; 1 gamma(lambda) per cell (reactive);
                                  2 per-cell gamma array uploaded in place of lambda */
} torch_config;

/* Per-step results (the values Simulation.step folds into its integrals). */
typedef struct torch_step_info {
    double dt;
    int fallback;        /* whole-step first-order fallback was used */
    int retries;         /* reactive dt halvings */
    double outward[5];   /* dt*vol/h * sum(F_last - F_first), per field */
    double source[5];    /* reactive source integral increment */
    double wave_max[3];  /* CFL maxima of the committed state */
} torch_step_info;

/* Execution(workers) construction (torchExecution.py:17-24): one device
 * context with its own HIP stream. */
int torch_ctx_create(int device, torch_ctx** out);
int torch_ctx_destroy(torch_ctx* ctx);
/* Message of the calling thread's last failure. */
int torch_last_error(char* buffer, size_t size);

/* Device-resident session state (Simulation.state/lam/beta,
 * torchSession.py:145-158), double-buffered so a failed step leaves the
 * committed data intact (:374-377, 518-522). */
int torch_state_create(torch_ctx* ctx, const torch_config* config, torch_state** out);
int torch_state_destroy(torch_state* state);
int torch_state_upload(torch_state* state, const double* U, const double* lam, const double* beta);
int torch_state_download(torch_state* state, double* U, double* lam, double* beta);

/* Simulation._dt on the committed state (torchSession.py:333-360). */
int torch_state_dt(torch_state* state, double* dt_limit);
/* The CFL maxima behind torch_state_dt, reduced on the device: 3D/2D
 * max(|u_m| + a) per axis; 1D [max |S| of the wave-speed estimate,
 * max(|u| + a), 0], whose second entry gives hydro_dt_1d = cfl dx / max
 * (torchSolver.py:1049-1051) for the legacy reactive driver's CFL test
 * (torchSolver.py:1319-1331). Right after torch_state_react on a 1D or 2D
 * Cartesian state of the FP32 form: the maxima of the state the half-step
 * started from raised by the exact change its heating makes, which the
 * FP32 storage rounds away when it is below half an ulp (the FP64 method's
 * test then takes the same dt halvings). */
int torch_state_wave_maxima(torch_state* state, double* maxima);

/* Simulation.step (torchSession.py:374-522): CFL dt clipped to t_end - time,
 * reactive R(dt/2) with <= 60 halvings, hydro with fallback, species
 * transport, R(dt/2), balances; commits only on success. */
int torch_state_step(torch_state* state, double time, double t_end, torch_step_info* info);

/* One hydro update of the committed state with a given dt, committing the
 * result: _hydro / muscl_hancock_step_3d (torchSolver3D.py:122-222),
 * muscl_hancock_step_2d (torchSolver2D.py:274-466), the 1D session hydro
 * (torchSession.py:413-437). When want_fluxes is set the face fluxes are kept
 * for torch_state_fluxes. */
int torch_state_hydro(torch_state* state, double dt, int want_fluxes, torch_step_info* info);
/* Download the face fluxes of the last torch_state_hydro, reference layout:
 * x faces (dim+2, nz, ny, nx+1), y faces (dim+2, nz, ny+1, nx), z faces
 * (dim+2, nz+1, ny, nx); unused pointers may be NULL. */
int torch_state_fluxes(torch_state* state, double* fx, double* fy, double* fz);

/* One chemistry half-step of the committed state, committing the result:
 * reaction_step_3d / reaction_step_2d / new_lambda_n_e_calc
 * (torchSolver3D.py:225-243, torchSolver2D.py:474-509, torchSolver.py:708-789). */
int torch_state_react(torch_state* state, double dt);

/* Conserved totals sum(U)*volume per field (Simulation.integral,
 * torchSession.py:285-295), reduced on the device in double. */
int torch_state_totals(torch_state* state, double* out);

/* Measurement hook, no counterpart in the reference: when enabled, HIP events
 * bracket the hydro kernel launch of every step on the state's stream. Returns
 * the accumulated kernel milliseconds and launch count since the last call,
 * then resets them. */
int torch_state_profile(torch_state* state, int enable, double* kernel_ms, long* launches);

/* hllc_flux_grid on host face arrays (torchFlux.py:6-77): the flux-worker
 * batch itself. left/right are (nv, m) in Cartesian component order, gl/gr
 * have m entries, out is (nv, m). */
int torch_hllc_faces(torch_ctx* ctx, int precision, int nv, int direction, long m, const double* left,
                     const double* right, const double* gl, const double* gr, double* out);

/* torch_hllc_faces without any device allocation: no counterpart in the
 * reference. For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments and
 * results as torch_hllc_faces, for at most 65536 faces per call (TORCH_INVALID
 * beyond); the staging lives in a fixed 10.25 MiB device array per GPU that
 * the library holds until the process ends. Calls on one GPU are serialised. */
int torch_hllc_faces_static(torch_ctx* ctx, int precision, int nv, int direction, long m, const double* left,
                            const double* right, const double* gl, const double* gr, double* out);

/* ---------------------------------------------------------------------------
 * Metric/face scheme: the reference's second finite-volume family, used by
 * sessions with a curvilinear geometry or explicit face boundaries
 * (Simulation._metric_step, torchSession.py:523-576; torchFiniteVolume.py:16-276):
 * primitive MUSCL with Rusanov faces on logically uniform Cartesian,
 * cylindrical, annular or spherical grids, per-face boundaries, SSPRK2 with a
 * first-order retry, area-weighted divergence and geometric sources.
 * ------------------------------------------------------------------------- */
enum torch_geometry { TORCH_CARTESIAN = 0, TORCH_CYLINDRICAL = 1, TORCH_ANNULAR = 2, TORCH_SPHERICAL = 3 };
enum torch_face_type {
    TORCH_FACE_OUTFLOW = 0, TORCH_FACE_WALL = 1, TORCH_FACE_PERIODIC = 2, TORCH_FACE_INFLOW = 3,
    TORCH_FACE_PRESSURE_OUTLET = 4, TORCH_FACE_STAGNATION_INLET = 5
};
/* One face boundary, normalized like torchBoundaries.normalize_faces
 * (torchBoundaries.py:17-174): type, and the parameters of its type. */
typedef struct torch_face {
    int type;                 /* enum torch_face_type */
    int has_backflow;         /* pressure outlet: a backflow state is given */
    double primitive[5];      /* inflow primitive, or the pressure outlet's backflow primitive (dim+2 used) */
    double pressure;          /* pressure outlet */
    double total_pressure, total_temperature, gas_constant;  /* stagnation inlet */
    double lambda, beta;      /* boundary fractions (1 when not given) */
} torch_face;

/* Attaches a metric to a state (Metrics, torchGeometry.py:93-182). factors
 * holds the per-axis factors whose outer products (in axis order x, y, z)
 * are the cell volumes and face areas: the volume factor of each axis (n_a
 * values), then for each direction d the area factors of each axis a (n_a
 * values, n_d + 1 when a == d); faces holds x_min, x_max, y_min, ... */
int torch_state_set_metric(torch_state* state, int geometry, const double* factors, long nfactors,
                           const torch_face* faces);
/* cfl_dt of the committed state (torchFiniteVolume.py:66-98). */
int torch_state_metric_dt(torch_state* state, double* dt);
/* rhs of the committed state (torchFiniteVolume.py:125-229): change
 * (dim+2, cells), species change (one row per fraction), outward flux and
 * integrated geometric source (dim+2 each); NULL skips an output. */
int torch_state_metric_rhs(torch_state* state, int first_order, double* change, double* species, double* outward,
                           double* source);
/* advance (torchFiniteVolume.py:232-276) of the committed state by dt,
 * committing the result: outward = boundary, source = geometric source,
 * fallback = the first-order retry was taken. */
int torch_state_metric_advance(torch_state* state, double dt, torch_step_info* info);
/* Simulation._metric_step (torchSession.py:523-576): cfl_dt clipped to
 * t_end - time, reactive R(dt/2) with the CFL test, advance, R(dt/2), up to
 * 20 halvings; commits only on success. */
int torch_state_metric_step(torch_state* state, double time, double t_end, torch_step_info* info);
/* rusanov_flux_grid on host face arrays (torchFlux.py:80-122): left/right
 * are (nv, m) primitive face states, gl/gr m gammas, out (nv, m). */
int torch_rusanov_faces(torch_ctx* ctx, int precision, int nv, int direction, long m, const double* left,
                        const double* right, const double* gl, const double* gr, double* out);

/* ---------------------------------------------------------------------------
 * Adaptive leaves: the reference's dyadic AMR family (AdaptiveSimulation,
 * torchAdaptive.py:14-206 and 400-481; LeafMesh, torchAdaptiveGeometry.py:
 * 125-372). The leaf tree, its face list and remeshing stay host data
 * structures (torch_amd/torchAdaptiveGeometry.py); the device holds one leaf
 * mesh at a time as flat arrays and runs the shared-face Rusanov SSPRK2
 * update, the CFL rate and the split chemistry over it. Arrays crossing the
 * boundary are host float64 (dim+2, nleaf) states and nleaf fractions.
 * ------------------------------------------------------------------------- */
typedef struct torch_amr torch_amr;
/* A leaf solver with the numeric options of config (dim, gamma, cfl,
 * reactive, kinetic model and its constants, precision; counts and
 * boundaries are unused), the geometry, the face boundaries (x_min, x_max,
 * y_min, ...) and the reconstruction order (1 or 2) of
 * AdaptiveSimulation.config (torchAdaptive.py:210-217, torchAdaptiveGeometry.py:12-44). */
int torch_amr_create(torch_ctx* ctx, const torch_config* config, int geometry, const torch_face* faces, int order,
                     torch_amr** out);
int torch_amr_destroy(torch_amr* amr);
/* Installs a leaf mesh (the LeafMesh attributes of torchAdaptiveGeometry.py:
 * 170-248): per face its left and right leaf (-1 outside the domain),
 * direction, side (-1 low or +1 high domain boundary, 0 interior or periodic
 * seam), area, gradient weight, and (nface, dim) offsets of the face centre
 * from the left and right leaf centres and the gradient displacement; per
 * leaf its volume, area divergence (dim, nleaf) and inverse normal matrix
 * (nleaf, dim, dim); the incidence lists in the order the reference
 * accumulates (nleaf + 1 starts, then face and role, -1 left / +1 right, of
 * each of the nentry entries). The stored state is undefined until the next
 * upload. */
int torch_amr_set_mesh(torch_amr* amr, long nleaf, long nface, const int* left, const int* right,
                       const int* direction, const int* side, const double* area, const double* weight,
                       const double* offset_left, const double* offset_right, const double* displacement,
                       const double* volume, const double* divergence, const double* inverse, long nentry,
                       const long* start, const int* entry_face, const int* entry_role);
/* State (dim+2, nleaf) and, reactive, lambda and beta (NULL beta: ones). */
int torch_amr_upload(torch_amr* amr, const double* U, const double* lam, const double* beta);
int torch_amr_download(torch_amr* amr, double* U, double* lam, double* beta);
/* face_values without reconstruction (torchAdaptive.py:14-57) of the
 * committed state, as remesh reads them: left/right (dim+2, nface)
 * primitives, gl/gr the nface gammas; NULL skips an output. */
int torch_amr_faces(torch_amr* amr, double* left, double* right, double* gl, double* gr);
/* cfl (torchAdaptive.py:149-166) of the committed state. */
int torch_amr_dt(torch_amr* amr, double* dt);
/* rhs (torchAdaptive.py:106-146) of the committed state: change (dim+2,
 * nleaf), species change (one row per fraction), boundary and integrated
 * geometric source (dim+2 each); NULL skips an output. */
int torch_amr_rhs(torch_amr* amr, int first_order, double* change, double* species, double* outward,
                  double* source);
/* advance (torchAdaptive.py:169-206) of the committed state by dt,
 * committing the result: outward = boundary, source = geometric source,
 * fallback = the first-order retry was taken. */
int torch_amr_advance(torch_amr* amr, double dt, torch_step_info* info);
/* AdaptiveSimulation.step up to its regrid (torchAdaptive.py:436-481): cfl
 * clipped to t_end - time, reactive R(dt/2) with the CFL test, advance,
 * R(dt/2), up to 24 halvings; commits only on success. */
int torch_amr_step(torch_amr* amr, double time, double t_end, torch_step_info* info);

/* ---------------------------------------------------------------------------
 * Batches. No counterpart in the reference, where every Simulation steps on
 * its own (torchSession.py:374-522) and a batch of problems is a Python loop
 * over sessions. Here many instances (inert of any dimension, reactive 3D;
 * any size, limiter, boundaries, gamma, CFL and end time) live on the device
 * together and one torch_batch_run step advances all of them: a fixed
 * sequence of work-list launches, captured once in a HIP graph and replayed.
 * Each instance keeps the single-session arithmetic and its own dt, time,
 * step count, boundary and source integrals and reactive dt halvings (the
 * bookkeeping of Simulation.step, done on the device), so its trajectory is
 * the one torch_state_step gives alone. A reactive step whose first
 * chemistry half-step fails the CFL test takes one more replay per halving.
 * Instances may be added and removed between runs (continuous batching);
 * the batch caps its device memory at hbm_budget bytes (0: no cap).
 * ------------------------------------------------------------------------- */
typedef struct torch_batch torch_batch;

typedef struct torch_batch_progress {
    double time;         /* committed time */
    double dt;           /* dt of the next step */
    int steps;           /* committed steps */
    int fallbacks;       /* steps that used the first-order fallback */
    int status;          /* TORCH_OK, or the code that stopped the instance */
    int active;          /* still stepping */
    int completed;       /* time == t_end */
    long cells;          /* cells of the instance */
    int retries;         /* reactive dt halvings so far */
    double outward[5];   /* accumulated boundary integral */
    double source[5];    /* accumulated source integral (reactive) */
    double wave_max[3];  /* CFL maxima of the committed state */
} torch_batch_progress;

/* Device bytes an instance of this configuration takes in a batch. */
int torch_batch_footprint(const torch_config* config, int precision, size_t* bytes);
int torch_batch_create(torch_ctx* ctx, int precision, size_t hbm_budget, torch_batch** out);
int torch_batch_destroy(torch_batch* batch);
/* Admits an instance (Simulation(...) followed by run(t_end) with at most
 * max_steps steps, torchSession.py:19-183, 563-600) with state U in the
 * reference layout at `time`: inert with a scalar gamma in 1D, 2D or 3D, or
 * reactive in 3D (gamma_mode 1; lam and, for ZND, beta of the cell shape;
 * beta NULL means ones). TORCH_BUDGET when it does not fit now. */
int torch_batch_add(torch_batch* batch, const torch_config* config, const double* U, const double* lam,
                    const double* beta, double time, double t_end, int max_steps, int* id);
/* Starts an admitted instance's boundary and source integrals (dim+2 values
 * each; NULL leaves zeros) at a session's running totals, before its first
 * step, so the device adds each step's terms in the session's order
 * (torchSession.py:518-521) and a session resumed in a batch keeps its
 * balances bit for bit. */
int torch_batch_seed(torch_batch* batch, int id, const double* outward, const double* source);
int torch_batch_remove(torch_batch* batch, int id);
/* Advances every active instance by up to `steps` steps, then waits. An
 * instance stops at t_end, at max_steps, or on the error its status names
 * (its committed state is then the last good one). */
int torch_batch_run(torch_batch* batch, int steps);
int torch_batch_info(torch_batch* batch, int id, torch_batch_progress* info);
/* Committed state (and, reactive, lambda and beta; NULL skips) of an instance. */
int torch_batch_download(torch_batch* batch, int id, double* U, double* lam, double* beta);
/* Autotuner: times the 3D second-order work list over the batch's current
 * instances for every supported tile shape and z chunk, keeps the fastest. */
int torch_batch_tune(torch_batch* batch, int reps, int* block_x, int* block_y, int* zchunk, double* best_ms);
/* Measurement hook, no counterpart in the reference: advance the batch by
 * `steps` replays of its step graph, captured with a HIP event pair around
 * its one 3D work list (TORCH_INVALID if it has none or several), one replay
 * at a time, and return that kernel's mean duration in ms (`ms`). The
 * replays are ordinary steps; only the host waits between them. */
int torch_batch_time_step_kernel(torch_batch* batch, int steps, double* ms);
/* Fix the tile instead of tuning it: a shape listed at torch_config.block_x
 * (32x24 only for an FP32 batch; TORCH_INVALID otherwise) and any z chunk
 * >= 1. */
int torch_batch_set_tiles(torch_batch* batch, int block_x, int block_y, int zchunk);
/* Measurement hook: average milliseconds of `reps` launches of the 3D
 * second-order work list on the current states (HIP events on the batch's
 * stream); writes only the spare state buffers, so nothing advances. */
int torch_batch_time_pass(torch_batch* batch, int reps, double* ms);
/* Measurement hook: HIP events around the graph replays of each run. Returns
 * the accumulated milliseconds and steps since the last call, then resets. */
int torch_batch_profile(torch_batch* batch, int enable, double* ms, long* steps);
/* Measurement hook: how often the step graph was instantiated and how often
 * an existing one was updated in place (same work lists) since the batch
 * was created. */
int torch_batch_graph_counts(torch_batch* batch, long* instantiated, long* updated);

#ifdef __cplusplus
}
#endif

#endif /* TORCH_HIP_H */
