// Batched sessions (torch_batch_*): many inert instances of any dimension,
// size, limiter and boundary advance together. Each step of the whole batch
// is a fixed launch sequence (per dimension: second-order work list,
// first-order work list gated per instance on the device, 1D CFL pass; then
// one post block per instance) captured once in a HIP graph and replayed.
// Every instance runs exactly the single-session arithmetic of
// Simulation.step (python/cfdSolver/torchSession.py:374-522) with its own
// dt, so its trajectory does not depend on what else is in the batch.
//
// Device layout: instances are ordered [3D | 2D | 1D]; BatchItem and
// BatchCtl arrays follow that order; each work-list launch gets a block
// table (first block of every instance of its segment). Partials of an
// instance are [second-order | first-order | CFL] slots.
#include <math.h>
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "capi_common.h"

using namespace tcapi;

namespace {

constexpr int kFlagStride = 16;  // ints per instance flag word (one cache line each)
constexpr int kAttempts = 3;     // chemistry attempts of a reactive step per graph replay
constexpr int kSide = 8;         // side streams of a captured step's branches
constexpr int kMarks = 64;       // fork/join events, reused round-robin within a capture

struct Inst {
    int id = 0;
    torch_config cfg;
    fvl::Grid grid;
    long ncell = 0;
    int nf = 0;
    void* U[2] = {nullptr, nullptr};   // inert: the two state buffers
    void* slot[3][4] = {};             // reactive: [U, lambda, beta] x kReactSlots rotating buffers
    void* scratch = nullptr;  // 1D interface fluxes
    size_t bytes = 0;
    double volume = 0;
    fvl::BatchCtl ctl;  // host copy, current whenever no run is in flight
    bool fresh = true;  // first dt not derived yet
    // layout of the current build
    int tiles[3] = {0, 0, 0};
    int zchunk = 16;  // z chunk of its 3D work-list blocks
    int nblk = 0, nllf = 0, ncfl = 0, nreact = 0;  // blocks of the passes
    long pslot = 0;                    // first partial slot
};

struct Table {
    int offset = 0;  // into the concatenated int buffer
    int count = 0;   // instances
    int blocks = 0;  // sum of their blocks
};

size_t state_bytes(const torch_config& c, size_t elem) {
    long ncell = (long)c.n[0] * (c.dim >= 2 ? c.n[1] : 1) * (c.dim >= 3 ? c.n[2] : 1);
    size_t b = 2 * (size_t)(c.dim + 2) * ncell * elem;
    if (c.reactive) b = fvl::kReactSlots * (size_t)(c.dim + 2 + 1 + (c.kinetic_model == 1)) * ncell * elem;
    if (c.dim == 1) b += elem * (size_t)(3 * (c.n[0] + 1) + 64);
    // partial slots (default tiles) and bookkeeping
    size_t slots = (size_t)(ncell / 256 + 8) * (c.reactive ? 5 : 3);
    return b + slots * fvl::NSLOT * sizeof(double) + sizeof(fvl::BatchItem) + sizeof(fvl::BatchCtl);
}

}  // namespace

struct torch_batch {
    torch_ctx* ctx = nullptr;
    const Ops* ops = nullptr;
    int precision = 0;
    size_t budget = 0, used = 0;
    std::vector<Inst*> inst;
    int next_id = 1;
    bool dirty = true;
    int seg[4] = {0, 0, 0, 0};  // [3D | 2D | 1D] segment starts, seg[3] = n
    int lim[3][6] = {};         // limiter sub-segment starts within each dimension segment
    int rseg = 0;               // first reactive instance of the 3D segment (reactive ones last)
    int rlim[2][6] = {};        // limiter sub-segments of the reactive 3D instances, per kinetic model
    fvl::BatchItem* items = nullptr;
    fvl::BatchItem* pitems = nullptr;  // copies of the items of new instances (first dt)
    fvl::BatchCtl* ctl = nullptr;
    int* flags = nullptr;
    double* partials = nullptr;
    int* tables = nullptr;
    fvl::BatchCtl* host_ctl = nullptr;  // pinned
    size_t cap_n = 0, cap_partials = 0, cap_tables = 0, cap_pitems = 0;
    Table hyd[3][5];          // second-order tables per dimension and limiter (inert)
    Table hydr[2][5];         // second-order tables of reactive 3D instances per model and limiter
    Table rx;                 // chemistry tables of reactive 3D instances
    Table l3, l2, l1, c1;     // first-order tables per dimension, 1D CFL
    Table p3, p2, p1;         // first-dt tables over pitems
    int pseg[4] = {0, 0, 0, 0};
    // The step graph. A membership or tile change marks it stale (graph_ok
    // false) but keeps the executable: the next capture first tries to
    // update it in place, which is far cheaper than instantiating a new one
    // when the work lists are the same. A stale executable may name freed
    // buffers; it is never launched before that update or its destruction.
    hipGraphExec_t graph = nullptr;
    bool graph_ok = false;
    long graph_builds = 0, graph_updates = 0;  // instantiations, in-place updates
    // Side streams and events that fork the captured step into parallel
    // branches (work lists of different instances); capture only.
    hipStream_t side[kSide] = {};
    hipEvent_t mark[kMarks] = {};
    int bx = 32, by = 16, zchunk = 16;  // tile of the 3D work lists (by: see create)
    bool tiles_set = false;  // zchunk chosen by set_tiles or tune; else default_zchunk per instance
    bool profile = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double prof_ms = 0;
    long prof_steps = 0;
    // torch_batch_time_step_kernel: while it runs, the captured step records
    // kev[0] and kev[1] around the batch's one 3D work list (external event
    // nodes: a plain record during capture only orders streams).
    bool kev_capture = false;
    hipEvent_t kev[2] = {nullptr, nullptr};
};

namespace {

int seg_of(int dim) { return dim == 3 ? 0 : (dim == 2 ? 1 : 2); }

void free_inst(Inst* p) {
    (void)hipFree(p->U[0]);
    (void)hipFree(p->U[1]);
    for (auto& kind : p->slot)
        for (void*& q : kind) {
            (void)hipFree(q);
            q = nullptr;
        }
    (void)hipFree(p->scratch);
    p->U[0] = p->U[1] = p->scratch = nullptr;
    delete p;
}

Inst* find(torch_batch* b, int id) {
    for (Inst* p : b->inst)
        if (p->id == id) return p;
    return nullptr;
}

void drop_graph(torch_batch* b) { b->graph_ok = false; }

void free_graph(torch_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_ok = false;
}

// Block counts of each pass for the current tile choice.
void layout(torch_batch* b, Inst* p) {
    const fvl::Grid& g = p->grid;
    const Ops* o = b->ops;
    if (g.dim == 3) {
        p->tiles[0] = (g.nx + b->bx - 3) / (b->bx - 2);
        p->tiles[1] = (g.ny + b->by - 3) / (b->by - 2);
        p->zchunk = b->tiles_set ? b->zchunk : default_zchunk(g.nz, b->by);
        p->tiles[2] = (g.nz + p->zchunk - 1) / p->zchunk;
        p->nblk = p->tiles[0] * p->tiles[1] * p->tiles[2];
        p->nllf = o->llf3d_blocks(g);
        p->ncfl = o->cfl_blocks(g);
        p->nreact = p->cfg.reactive ? o->react_blocks(g) : 0;
    } else if (g.dim == 2) {
        const int rows = o->strip2d_rows(g);
        p->tiles[0] = (g.nx + 61) / 62;
        p->tiles[1] = (g.ny + rows - 1) / rows;
        p->tiles[2] = rows;
        p->nblk = p->tiles[0] * p->tiles[1];
        p->nllf = o->llf2d_blocks(g);
        p->ncfl = o->cfl_blocks(g);
    } else {
        p->tiles[0] = o->step1d_tiles(g);
        p->tiles[1] = p->tiles[2] = 1;
        p->nblk = p->tiles[0];  // step and first-order tiles; no partials
        p->nllf = p->tiles[0];
        p->ncfl = o->cfl_blocks(g);
    }
}

// Appends one block table (count+1 entries) over instances [a, e).
Table table(std::vector<int>& buf, const std::vector<Inst*>& v, int a, int e, int Inst::*what) {
    Table t;
    t.offset = (int)buf.size();
    t.count = e - a;
    int acc = 0;
    buf.push_back(0);
    for (int i = a; i < e; ++i) {
        acc += v[i]->*what;
        buf.push_back(acc);
    }
    t.blocks = t.count > 0 ? acc : 0;
    return t;
}

int grow(void** p, size_t* cap, size_t bytes) {
    if (bytes <= *cap && *p) return TORCH_OK;
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t want = std::max(bytes, (size_t)4096);
    want += want / 2;
    int rc = alloc(p, want);
    if (!rc) *cap = want;
    return rc;
}

// The second-order pass of every 3D instance (one launch per limiter, the
// reactive ones with species transport).
hipError_t enqueue_3d(torch_batch* b, hipStream_t st) {
    for (int L = 0; L < 5; ++L) {
        const Table& t = b->hyd[0][L];
        const bool timed = b->kev_capture && t.blocks > 0;
        hipError_t e = timed ? hipEventRecordWithFlags(b->kev[0], st, hipEventRecordExternal) : hipSuccess;
        if (e == hipSuccess)
            e = b->ops->batch3d(b->items + b->lim[0][L], b->tables + t.offset, t.count, t.blocks, b->bx, b->by, L, 0,
                                st);
        if (e == hipSuccess && timed) e = hipEventRecordWithFlags(b->kev[1], st, hipEventRecordExternal);
        if (e != hipSuccess) return e;
    }
    // Arrhenius instances transport one fraction, ZND ones two.
    for (int m = 0; m < 2; ++m)
        for (int L = 0; L < 5; ++L) {
            const Table& t = b->hydr[m][L];
            hipError_t e = b->ops->batch3d(b->items + b->rlim[m][L], b->tables + t.offset, t.count, t.blocks, b->bx,
                                           b->by, L, m + 1, st);
            if (e != hipSuccess) return e;
        }
    return hipSuccess;
}

// Fork/join of a captured step: a branch starts on a side stream after the
// work already enqueued on `from`, and `into` continues after the branch.
// Captured, these events become graph edges, so work lists of different
// instances become parallel graph branches (no shared buffers: each
// instance owns its state, partial slots, flags and control record).
struct Fork {
    torch_batch* b;
    int stream = 0, marks = 0;
    hipError_t err = hipSuccess;
    hipStream_t open(hipStream_t from) {
        hipStream_t s = b->side[stream++ % kSide];
        link(from, s);
        return s;
    }
    void link(hipStream_t from, hipStream_t to) {
        if (err != hipSuccess) return;
        // Each event is recorded and waited on back to back, so the pool is
        // reused round-robin (a capture takes the event's nodes at the wait).
        hipEvent_t ev = b->mark[marks++ % kMarks];
        err = hipEventRecord(ev, from);
        if (err == hipSuccess) err = hipStreamWaitEvent(to, ev, 0);
    }
};

// One step with each work list in a branch of its own: inert 3D lists from
// the start, reactive 3D lists after the chemistry attempts, the 2D and 1D
// pipelines beside the 3D one; every branch joins before the pass that
// reads all of its segment (first-order 3D after the 3D lists, the post
// kernel after everything).
hipError_t enqueue_branched(torch_batch* b, hipStream_t st) {
    const Ops* o = b->ops;
    const int* T = b->tables;
    fvl::BatchItem* I3 = b->items + b->seg[0];
    fvl::BatchItem* I2 = b->items + b->seg[1];
    fvl::BatchItem* I1 = b->items + b->seg[2];
    fvl::BatchItem* IR = b->items + b->rseg;
    const int nr = b->seg[1] - b->rseg;
    Fork f{b};
    hipError_t e;
    std::vector<hipStream_t> open3;
    for (int L = 0; L < 5; ++L) {
        const Table& t = b->hyd[0][L];
        if (t.blocks <= 0) continue;
        hipStream_t s = f.open(st);
        open3.push_back(s);
        if (b->kev_capture && (e = hipEventRecordWithFlags(b->kev[0], s, hipEventRecordExternal))) return e;
        if ((e = o->batch3d(b->items + b->lim[0][L], T + t.offset, t.count, t.blocks, b->bx, b->by, L, 0, s)))
            return e;
        if (b->kev_capture && (e = hipEventRecordWithFlags(b->kev[1], s, hipEventRecordExternal))) return e;
    }
    // 2D and 1D pipelines: their lists in parallel, then the segment's
    // first-order (and 1D CFL) passes after the join.
    std::vector<hipStream_t> tails;
    for (int seg = 1; seg <= 2; ++seg) {
        if (b->seg[seg + 1] == b->seg[seg]) continue;
        hipStream_t head = f.open(st);
        std::vector<hipStream_t> lists;
        for (int L = 0; L < 5; ++L) {
            const Table& t = b->hyd[seg][L];
            if (t.blocks <= 0) continue;
            hipStream_t s = f.open(head);
            lists.push_back(s);
            e = seg == 1 ? o->batch2d(b->items + b->lim[1][L], T + t.offset, t.count, t.blocks, L, s)
                         : o->batch1d(b->items + b->lim[2][L], T + t.offset, t.count, t.blocks, L, s);
            if (e) return e;
        }
        for (hipStream_t s : lists) f.link(s, head);
        if (seg == 1) {
            if ((e = o->batch_llf2d(I2, T + b->l2.offset, b->l2.count, b->l2.blocks, head))) return e;
        } else {
            if ((e = o->batch_llf1d(I1, T + b->l1.offset, b->l1.count, b->l1.blocks, head))) return e;
            if ((e = o->batch_cfl1d(I1, T + b->c1.offset, b->c1.count, b->c1.blocks, 0, head))) return e;
        }
        tails.push_back(head);
    }
    for (int attempt = 0; attempt < (nr ? kAttempts : 0); ++attempt) {
        if ((e = o->batch_react(IR, T + b->rx.offset, b->rx.count, b->rx.blocks, 1, st))) return e;
        if ((e = o->batch_check(IR, b->ctl + b->rseg, nr, st))) return e;
    }
    for (int m = 0; m < 2; ++m)
        for (int L = 0; L < 5; ++L) {
            const Table& t = b->hydr[m][L];
            if (t.blocks <= 0) continue;
            hipStream_t s = f.open(st);
            open3.push_back(s);
            if ((e = o->batch3d(b->items + b->rlim[m][L], T + t.offset, t.count, t.blocks, b->bx, b->by, L, m + 1,
                                s)))
                return e;
        }
    for (hipStream_t s : open3) f.link(s, st);
    if ((e = o->batch_llf3d(I3, T + b->l3.offset, b->l3.count, b->l3.blocks, st))) return e;
    if ((e = o->batch_react(IR, T + b->rx.offset, b->rx.count, b->rx.blocks, 2, st))) return e;
    for (hipStream_t s : tails) f.link(s, st);
    if (f.err != hipSuccess) return f.err;
    return o->batch_post(b->items, b->ctl, b->seg[3], 0, st);
}

hipError_t enqueue_step(torch_batch* b, hipStream_t st) {
    const Ops* o = b->ops;
    const int* T = b->tables;
    fvl::BatchItem* I3 = b->items + b->seg[0];
    fvl::BatchItem* I2 = b->items + b->seg[1];
    fvl::BatchItem* I1 = b->items + b->seg[2];
    fvl::BatchItem* IR = b->items + b->rseg;
    const int nr = b->seg[1] - b->rseg;
    hipError_t e;
    // Reactive 3D: R(dt/2), then the CFL test of the heated state; a few
    // attempts per replay (an accepted instance skips the later ones), so a
    // step that needs dt halvings usually still commits in one replay.
    for (int attempt = 0; attempt < (nr ? kAttempts : 0); ++attempt) {
        if ((e = o->batch_react(IR, T + b->rx.offset, b->rx.count, b->rx.blocks, 1, st))) return e;
        if ((e = o->batch_check(IR, b->ctl + b->rseg, nr, st))) return e;
    }
    if ((e = enqueue_3d(b, st))) return e;
    if ((e = o->batch_llf3d(I3, T + b->l3.offset, b->l3.count, b->l3.blocks, st))) return e;
    if ((e = o->batch_react(IR, T + b->rx.offset, b->rx.count, b->rx.blocks, 2, st))) return e;
    for (int L = 0; L < 5; ++L) {
        const Table& t = b->hyd[1][L];
        if ((e = o->batch2d(b->items + b->lim[1][L], T + t.offset, t.count, t.blocks, L, st))) return e;
    }
    if ((e = o->batch_llf2d(I2, T + b->l2.offset, b->l2.count, b->l2.blocks, st))) return e;
    for (int L = 0; L < 5; ++L) {
        const Table& t = b->hyd[2][L];
        if ((e = o->batch1d(b->items + b->lim[2][L], T + t.offset, t.count, t.blocks, L, st))) return e;
    }
    if ((e = o->batch_llf1d(I1, T + b->l1.offset, b->l1.count, b->l1.blocks, st))) return e;
    if ((e = o->batch_cfl1d(I1, T + b->c1.offset, b->c1.count, b->c1.blocks, 0, st))) return e;
    return o->batch_post(b->items, b->ctl, b->seg[3], 0, st);
}

int capture(torch_batch* b) {
    drop_graph(b);
    if (b->seg[3] == 0) return TORCH_OK;
    hipStream_t st = b->ctx->stream;
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = b->side[0] ? enqueue_branched(b, st) : enqueue_step(b, st);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(st, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(TORCH_HIP_ERROR, std::string("graph capture: ") + hipGetErrorString(e ? e : e2));
    }
    if (b->graph) {
        // Same work lists (nodes and edges): the kernels' grids and
        // arguments are rewritten in place. Otherwise a new executable.
        hipGraphNode_t refused = nullptr;
        hipGraphExecUpdateResult how = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(b->graph, g, &refused, &how) == hipSuccess && how == hipGraphExecUpdateSuccess) {
            (void)hipGraphDestroy(g);
            b->graph_ok = true;
            b->graph_updates += 1;
            return TORCH_OK;
        }
        (void)hipGetLastError();
        free_graph(b);
    }
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return fail(TORCH_HIP_ERROR, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    b->graph_ok = true;
    b->graph_builds += 1;
    return TORCH_OK;
}

int read_ctl(torch_batch* b) {
    const int n = b->seg[3];
    if (n == 0) return TORCH_OK;
    hipStream_t st = b->ctx->stream;
    HIP_TRY(hipMemcpyAsync(b->host_ctl, b->ctl, n * sizeof(fvl::BatchCtl), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i) b->inst[i]->ctl = b->host_ctl[i];
    return TORCH_OK;
}

// Host-side lifetime check of a rebuild: every device pointer an item holds
// must lie inside memory the batch holds right now (an instance's buffers or
// the batch's own arrays), so no item, and no graph captured from the items,
// refers to a freed allocation. Costs one pass over the items' pointers.
struct Span {
    const char* lo;
    const char* hi;
};

int check_items(const torch_batch* b, const std::vector<fvl::BatchItem>& items) {
    std::vector<Span> live;
    auto add = [&](const void* q, size_t bytes) {
        if (q) live.push_back({(const char*)q, (const char*)q + bytes});
    };
    const size_t e = b->ops->elem;
    for (const Inst* p : b->inst) {
        const size_t state = (size_t)p->nf * p->ncell * e, row = (size_t)p->ncell * e;
        add(p->U[0], state);
        add(p->U[1], state);
        for (int j = 0; j < 4; ++j) {
            add(p->slot[0][j], state);
            add(p->slot[1][j], row);
            add(p->slot[2][j], row);
        }
        if (p->scratch) add(p->scratch, b->ops->step1d_scratch(p->grid));
    }
    add(b->partials, b->cap_partials);
    add(b->flags, b->cap_n * kFlagStride * sizeof(int));
    add(b->ctl, b->cap_n * sizeof(fvl::BatchCtl));
    auto ok = [&](const void* q) {
        if (!q) return true;
        const char* c = (const char*)q;
        for (const Span& sp : live)
            if (c >= sp.lo && c < sp.hi) return true;
        return false;
    };
    for (size_t i = 0; i < items.size(); ++i) {
        const fvl::BatchItem& it = items[i];
        const fvl::Step& s = it.step;
        bool good = ok(s.U) && ok(s.Ue) && ok(s.Uout) && ok(s.lam) && ok(s.q[0]) && ok(s.q[1]) && ok(s.qout[0]) &&
                    ok(s.qout[1]) && ok(s.F[0]) && ok(s.F[1]) && ok(s.F[2]) && ok(s.flags) && ok(s.partials) &&
                    ok(s.coef) && ok(it.scratch);
        for (const fvl::React* r : {&it.r1, &it.r2})
            good = good && ok(r->U) && ok(r->Uout) && ok(r->lam) && ok(r->lamout) && ok(r->beta) &&
                   ok(r->betaout) && ok(r->Ustart) && ok(r->Uold) && ok(r->flags) && ok(r->partials);
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 4; ++j) good = good && ok(it.slot[k][j]);
        if (!good)
            return fail(TORCH_HIP_ERROR, "internal: batch item " + std::to_string(i) +
                                             " refers to memory the batch does not hold");
    }
    return TORCH_OK;
}

// Device tables, items and control records for the current membership and
// tile choice; derives the first dt of new instances; captures the graph.
int rebuild(torch_batch* b) {
    drop_graph(b);
    std::vector<Inst*>& v = b->inst;
    std::stable_sort(v.begin(), v.end(), [](const Inst* x, const Inst* y) {
        int sx = seg_of(x->grid.dim), sy = seg_of(y->grid.dim);
        if (sx != sy) return sx < sy;
        if (x->cfg.reactive != y->cfg.reactive) return x->cfg.reactive < y->cfg.reactive;
        if (x->cfg.reactive && x->cfg.kinetic_model != y->cfg.kinetic_model)
            return x->cfg.kinetic_model < y->cfg.kinetic_model;
        return x->cfg.limiter < y->cfg.limiter;
    });
    const int n = (int)v.size();
    for (int s = 0; s < 3; ++s) {
        int i = 0;
        while (i < n && seg_of(v[i]->grid.dim) < s) ++i;
        b->seg[s] = i;
    }
    b->seg[3] = n;
    b->rseg = b->seg[0];
    while (b->rseg < b->seg[1] && !v[b->rseg]->cfg.reactive) ++b->rseg;
    for (int s = 0; s < 3; ++s) {
        const int end = s == 0 ? b->rseg : b->seg[s + 1];
        for (int L = 0; L <= 5; ++L) {
            int i = b->seg[s];
            while (i < end && v[i]->cfg.limiter < L) ++i;
            b->lim[s][L] = i;
        }
    }
    int mseg[3] = {b->rseg, b->rseg, b->seg[1]};  // [Arrhenius | ZND] within the reactive instances
    while (mseg[1] < b->seg[1] && v[mseg[1]]->cfg.kinetic_model == 0) ++mseg[1];
    for (int m = 0; m < 2; ++m)
        for (int L = 0; L <= 5; ++L) {
            int i = mseg[m];
            while (i < mseg[m + 1] && v[i]->cfg.limiter < L) ++i;
            b->rlim[m][L] = i;
        }
    long slots = 0;
    for (Inst* p : v) {
        layout(b, p);
        p->pslot = slots;
        slots += (p->grid.dim == 1 ? 0 : p->nblk + p->nllf) + p->ncfl + 2 * p->nreact;
    }
    std::vector<Inst*> fresh;
    for (Inst* p : v)
        if (p->fresh) fresh.push_back(p);
    for (int s = 0; s < 3; ++s) {
        int i = 0;
        while (i < (int)fresh.size() && seg_of(fresh[i]->grid.dim) < s) ++i;
        b->pseg[s] = i;
    }
    b->pseg[3] = (int)fresh.size();
    std::vector<int> tb;
    for (int s = 0; s < 3; ++s)
        for (int L = 0; L < 5; ++L) b->hyd[s][L] = table(tb, v, b->lim[s][L], b->lim[s][L + 1], &Inst::nblk);
    for (int m = 0; m < 2; ++m)
        for (int L = 0; L < 5; ++L) b->hydr[m][L] = table(tb, v, b->rlim[m][L], b->rlim[m][L + 1], &Inst::nblk);
    b->rx = table(tb, v, b->rseg, b->seg[1], &Inst::nreact);
    b->l3 = table(tb, v, b->seg[0], b->seg[1], &Inst::nllf);
    b->l2 = table(tb, v, b->seg[1], b->seg[2], &Inst::nllf);
    b->l1 = table(tb, v, b->seg[2], b->seg[3], &Inst::nllf);
    b->c1 = table(tb, v, b->seg[2], b->seg[3], &Inst::ncfl);
    b->p3 = table(tb, fresh, b->pseg[0], b->pseg[1], &Inst::ncfl);
    b->p2 = table(tb, fresh, b->pseg[1], b->pseg[2], &Inst::ncfl);
    b->p1 = table(tb, fresh, b->pseg[2], b->pseg[3], &Inst::ncfl);
    if (n == 0) return TORCH_OK;

    int rc = TORCH_OK;
    if ((size_t)n > b->cap_n) {
        size_t cap = std::max((size_t)n, b->cap_n * 2);
        (void)hipFree(b->items);
        (void)hipFree(b->ctl);
        (void)hipFree(b->flags);
        (void)hipHostFree(b->host_ctl);
        b->items = nullptr;
        b->ctl = nullptr;
        b->flags = nullptr;
        b->host_ctl = nullptr;
        b->cap_n = 0;
        rc = alloc((void**)&b->items, cap * sizeof(fvl::BatchItem));
        if (!rc) rc = alloc((void**)&b->ctl, cap * sizeof(fvl::BatchCtl));
        if (!rc) rc = alloc((void**)&b->flags, cap * kFlagStride * sizeof(int));
        if (!rc && hipHostMalloc((void**)&b->host_ctl, cap * sizeof(fvl::BatchCtl)) != hipSuccess)
            rc = fail(TORCH_HIP_ERROR, "hipHostMalloc failed");
        if (rc) return rc;
        b->cap_n = cap;
    }
    if ((rc = grow((void**)&b->partials, &b->cap_partials, (size_t)slots * fvl::NSLOT * sizeof(double))))
        return rc;
    if ((rc = grow((void**)&b->tables, &b->cap_tables, tb.size() * sizeof(int)))) return rc;
    if (!fresh.empty() &&
        (rc = grow((void**)&b->pitems, &b->cap_pitems, fresh.size() * sizeof(fvl::BatchItem))))
        return rc;

    std::vector<fvl::BatchItem> items(n);
    for (int i = 0; i < n; ++i) {
        Inst* p = v[i];
        fvl::BatchItem& it = items[i];
        memset(&it, 0, sizeof(it));
        fvl::Step& s = it.step;
        s.U = p->U[p->ctl.cur];
        s.Uout = p->U[p->ctl.cur ^ 1];
        s.gamma = p->cfg.gamma;
        s.flags = b->flags + (long)i * kFlagStride;
        s.partials = b->partials + p->pslot * fvl::NSLOT;
        s.zchunk = p->zchunk;
        s.limiter = p->cfg.limiter;
        s.want_cfl = p->grid.dim >= 2;
        s.controller = p->cfg.controller;
        s.variant = p->cfg.variant;
        s.wavespeed = p->cfg.wavespeed;
        s.cfl = p->cfg.cfl;
        // Device address of this instance's step coefficients.
        s.coef = (const double*)((const char*)(b->ctl + i) + offsetof(fvl::BatchCtl, coef));
        it.grid = p->grid;
        for (int d = 0; d < 3; ++d) it.tiles[d] = p->tiles[d];
        it.llf_first = p->grid.dim == 1 ? 0 : p->nblk;
        it.cfl_first = p->grid.dim == 1 ? 0 : p->nblk + p->nllf;
        it.scratch = p->scratch;
        if (p->cfg.reactive) {
            const torch_config& c = p->cfg;
            s.want_cfl = 0;  // the maxima come from the second chemistry half-step
            s.nq = c.kinetic_model == 1 ? 2 : 1;
            it.reactive = 1;
            it.znd = c.kinetic_model == 1;
            for (int k = 0; k < 3; ++k)
                for (int j = 0; j < 4; ++j) it.slot[k][j] = p->slot[k][j];
            for (fvl::React* r : {&it.r1, &it.r2}) {
                r->model = c.kinetic_model;
                r->q_heat = c.q_heat;
                r->activation_energy = c.activation_energy;
                r->pre_exponential = c.pre_exponential;
                r->flags = s.flags;
                r->partials = s.partials;
                r->want_cfl = 1;
            }
            it.r1_first = p->nblk + p->nllf + p->ncfl;
            it.r2_first = it.r1_first + p->nreact;
            fvl::set_reactive_slots(it, p->ctl.cur);
        }
        p->ctl.nblk = p->nblk;
        p->ctl.nllf = p->nllf;
        p->ctl.ncfl = p->ncfl;
        p->ctl.nreact = p->nreact;
        p->ctl.ticket = 0;
        b->host_ctl[i] = p->ctl;
    }
    if ((rc = check_items(b, items))) return rc;
    hipStream_t st = b->ctx->stream;
    HIP_TRY(hipMemcpyAsync(b->items, items.data(), n * sizeof(fvl::BatchItem), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->ctl, b->host_ctl, n * sizeof(fvl::BatchCtl), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->tables, tb.data(), tb.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(b->flags, 0, n * kFlagStride * sizeof(int), st));
    HIP_TRY(hipStreamSynchronize(st));

    // First dt of new instances: CFL maxima of their initial states, from
    // copies of their items (same buffers and partial slots).
    if (!fresh.empty()) {
        std::vector<fvl::BatchItem> pit;
        for (Inst* p : fresh)
            for (int i = 0; i < n; ++i)
                if (v[i] == p) pit.push_back(items[i]);
        HIP_TRY(hipMemcpyAsync(b->pitems, pit.data(), pit.size() * sizeof(fvl::BatchItem), hipMemcpyHostToDevice,
                               st));
        const Ops* o = b->ops;
        const int* T = b->tables;
        HIP_TRY(o->batch_cfl3d(b->pitems + b->pseg[0], T + b->p3.offset, b->p3.count, b->p3.blocks, st));
        HIP_TRY(o->batch_cfl2d(b->pitems + b->pseg[1], T + b->p2.offset, b->p2.count, b->p2.blocks, st));
        HIP_TRY(o->batch_cfl1d(b->pitems + b->pseg[2], T + b->p1.offset, b->p1.count, b->p1.blocks, 1, st));
        HIP_TRY(o->batch_post(b->items, b->ctl, n, 1, st));
        if ((rc = read_ctl(b))) return rc;
        for (Inst* p : v) p->fresh = false;
    }
    if ((rc = capture(b))) return rc;
    b->dirty = false;
    return TORCH_OK;
}

}  // namespace

extern "C" {

int torch_batch_footprint(const torch_config* config, int precision, size_t* bytes) {
    if (!config || !bytes) return fail(TORCH_INVALID, "null argument");
    int rc = check_config(*config);
    if (rc) return rc;
    *bytes = state_bytes(*config, precision == TORCH_FP32 ? sizeof(float) : sizeof(double));
    return TORCH_OK;
}

int torch_batch_create(torch_ctx* ctx, int precision, size_t hbm_budget, torch_batch** out) {
    if (!ctx || !out) return fail(TORCH_INVALID, "null argument");
    if (precision != TORCH_FP64 && precision != TORCH_FP32) return fail(TORCH_INVALID, "unknown precision");
    torch_batch* b = new torch_batch();
    b->ctx = ctx;
    b->precision = precision;
    b->ops = precision == TORCH_FP64 ? &kOps64 : &kOps32;
    b->budget = hbm_budget;
    b->by = precision == TORCH_FP32 ? 8 : 16;  // the single-session default tile
    hipError_t e = hipSetDevice(ctx->device);
    for (int i = 0; i < kSide && e == hipSuccess; ++i) e = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking);
    for (int i = 0; i < kMarks && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&b->mark[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        torch_batch_destroy(b);
        return fail(TORCH_HIP_ERROR, std::string("batch streams: ") + hipGetErrorString(e));
    }
    *out = b;
    return TORCH_OK;
}

int torch_batch_destroy(torch_batch* b) {
    if (!b) return TORCH_OK;
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    free_graph(b);
    for (Inst* p : b->inst) free_inst(p);
    (void)hipFree(b->items);
    (void)hipFree(b->ctl);
    (void)hipFree(b->flags);
    (void)hipFree(b->partials);
    (void)hipFree(b->tables);
    (void)hipFree(b->pitems);
    (void)hipHostFree(b->host_ctl);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    for (hipEvent_t ev : b->kev)
        if (ev) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : b->mark)
        if (ev) (void)hipEventDestroy(ev);
    for (hipStream_t s : b->side)
        if (s) (void)hipStreamDestroy(s);
    delete b;
    return TORCH_OK;
}

int torch_batch_add(torch_batch* b, const torch_config* config, const double* U, const double* lam,
                    const double* beta, double time, double t_end, int max_steps, int* id) {
    if (!b || !config || !U || !id) return fail(TORCH_INVALID, "null argument");
    const torch_config& c = *config;
    int rc = check_config(c);
    if (rc) return rc;
    if (c.reactive ? c.dim != 3 : c.gamma_mode != 0)
        return fail(TORCH_INVALID, "batches hold inert scalar-gamma instances and reactive 3D instances");
    if (c.reactive && !lam) return fail(TORCH_INVALID, "reactive instance needs a lambda array");
    if (!isfinite(time) || !isfinite(t_end) || !(time < t_end)) return fail(TORCH_INVALID, "require time < t_end");
    if (max_steps <= 0) return fail(TORCH_INVALID, "max_steps must be positive");
    const size_t bytes = state_bytes(c, b->ops->elem);
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    if (b->budget && b->used + bytes > b->budget) return fail(TORCH_BUDGET, "instance does not fit the HBM budget");
    HIP_TRY(hipSetDevice(b->ctx->device));
    Inst* p = new Inst();
    p->cfg = c;
    p->cfg.precision = b->precision;
    make_grid(c, &p->grid, &p->volume);
    p->ncell = (long)p->grid.nx * p->grid.ny * p->grid.nz;
    p->nf = c.dim + 2;
    p->bytes = bytes;
    const size_t e = b->ops->elem;
    hipStream_t st = b->ctx->stream;
    if (c.reactive) {
        const bool znd = c.kinetic_model == 1;
        for (int j = 0; j < fvl::kReactSlots && !rc; ++j) {
            rc = alloc(&p->slot[0][j], p->nf * p->ncell * e);
            if (!rc) rc = alloc(&p->slot[1][j], p->ncell * e);
            if (!rc && znd) rc = alloc(&p->slot[2][j], p->ncell * e);
        }
        if (!rc) rc = upload_as(b->ops, st, p->slot[0][0], U, p->nf * p->ncell);
        if (!rc) rc = upload_as(b->ops, st, p->slot[1][0], lam, p->ncell);
        if (!rc && znd) {
            std::vector<double> ones;
            if (!beta) ones.assign(p->ncell, 1.0);
            rc = upload_as(b->ops, st, p->slot[2][0], beta ? beta : ones.data(), p->ncell);
        }
    } else {
        rc = alloc(&p->U[0], p->nf * p->ncell * e);
        if (!rc) rc = alloc(&p->U[1], p->nf * p->ncell * e);
        if (!rc && c.dim == 1) rc = alloc(&p->scratch, b->ops->step1d_scratch(p->grid));
        if (!rc) rc = upload_as(b->ops, st, p->U[0], U, p->nf * p->ncell);
    }
    if (rc) {
        free_inst(p);
        return rc;
    }
    fvl::BatchCtl& k = p->ctl;
    memset(&k, 0, sizeof(k));
    k.t_end = t_end;
    k.volume = p->volume;
    k.cfl = c.cfl;
    for (int d = 0; d < 3; ++d) k.h[d] = p->grid.h[d];
    k.time = time;
    k.dim = c.dim;
    k.max_steps = max_steps;
    k.active = 2;
    k.reactive = c.reactive != 0;
    p->id = b->next_id++;
    b->inst.push_back(p);
    b->used += bytes;
    b->dirty = true;
    *id = p->id;
    return TORCH_OK;
}

int torch_batch_seed(torch_batch* b, int id, const double* outward, const double* source) {
    if (!b) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    for (Inst* p : b->inst)
        if (p->id == id) {
            // Until its first step the host copy of the control block is the
            // one the next rebuild uploads.
            if (!p->fresh) return fail(TORCH_INVALID, "integrals can be seeded only before an instance's first step");
            for (int f = 0; f < p->nf; ++f)
                if ((outward && !isfinite(outward[f])) || (source && !isfinite(source[f])))
                    return fail(TORCH_INVALID, "integrals must be finite");
            for (int f = 0; f < p->nf; ++f) {
                if (outward) p->ctl.outward[f] = outward[f];
                if (source) p->ctl.source[f] = source[f];
            }
            return TORCH_OK;
        }
    return fail(TORCH_INVALID, "no such batch instance");
}

int torch_batch_remove(torch_batch* b, int id) {
    if (!b) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    for (size_t i = 0; i < b->inst.size(); ++i)
        if (b->inst[i]->id == id) {
            HIP_TRY(hipStreamSynchronize(b->ctx->stream));
            drop_graph(b);
            b->used -= b->inst[i]->bytes;
            free_inst(b->inst[i]);
            b->inst.erase(b->inst.begin() + i);
            b->dirty = true;
            return TORCH_OK;
        }
    return fail(TORCH_INVALID, "no such batch instance");
}

int torch_batch_run(torch_batch* b, int steps) {
    if (!b || steps < 0) return fail(TORCH_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    int rc;
    if (b->dirty && (rc = rebuild(b))) return rc;
    if (!b->graph_ok || steps == 0) return TORCH_OK;
    hipStream_t st = b->ctx->stream;
    if (b->profile) HIP_TRY(hipEventRecord(b->ev0, st));
    for (int k = 0; k < steps; ++k) HIP_TRY(hipGraphLaunch(b->graph, st));
    if (b->profile) HIP_TRY(hipEventRecord(b->ev1, st));
    if ((rc = read_ctl(b))) return rc;
    if (b->profile) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        b->prof_ms += ms;
        b->prof_steps += steps;
    }
    return TORCH_OK;
}

int torch_batch_info(torch_batch* b, int id, torch_batch_progress* info) {
    if (!b || !info) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    Inst* p = find(b, id);
    if (!p) return fail(TORCH_INVALID, "no such batch instance");
    memset(info, 0, sizeof(*info));
    const fvl::BatchCtl& k = p->ctl;
    info->time = k.time;
    info->dt = k.dt;
    info->steps = k.steps;
    info->fallbacks = k.fallbacks;
    info->status = k.status;
    info->active = k.active != 0;
    info->completed = k.time == k.t_end;
    info->cells = p->ncell;
    info->retries = k.retries;
    for (int f = 0; f < 5; ++f) {
        info->outward[f] = k.outward[f];
        info->source[f] = k.source[f];
    }
    for (int m = 0; m < 3; ++m) info->wave_max[m] = k.maxima[m];
    return TORCH_OK;
}

int torch_batch_download(torch_batch* b, int id, double* U, double* lam, double* beta) {
    if (!b || !U) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    Inst* p = find(b, id);
    if (!p) return fail(TORCH_INVALID, "no such batch instance");
    hipStream_t st = b->ctx->stream;
    if (!p->cfg.reactive) return download_as(b->ops, st, U, p->U[p->ctl.cur], p->nf * p->ncell);
    const int cur = p->ctl.cur;
    int rc = download_as(b->ops, st, U, p->slot[0][cur], p->nf * p->ncell);
    if (!rc && lam) rc = download_as(b->ops, st, lam, p->slot[1][cur], p->ncell);
    if (!rc && beta && p->slot[2][cur]) rc = download_as(b->ops, st, beta, p->slot[2][cur], p->ncell);
    return rc;
}

int torch_batch_tune(torch_batch* b, int reps, int* block_x, int* block_y, int* zchunk, double* best_ms) {
    if (!b || reps <= 0) return fail(TORCH_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    int rc = TORCH_OK;
    if (b->dirty && (rc = rebuild(b))) return rc;
    hipStream_t st = b->ctx->stream;
    double best = -1;
    if (b->seg[1] > b->seg[0]) {
        // 32x24 (768 threads, one block per CU) exists in the FP32 form only:
        // FP64 tiles need more than the 168 VGPRs three waves per SIMD allow.
        static const int shapes[5][2] = {{64, 4}, {64, 8}, {32, 8}, {32, 16}, {32, 24}};
        static const int chunks[5] = {8, 16, 24, 32, 64};
        const int nshape = b->precision == TORCH_FP32 ? 5 : 4;
        int bbx = b->bx, bby = b->by, bz = b->zchunk;
        b->tiles_set = true;
        hipEvent_t e0, e1;
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        for (int si = 0; si < nshape && !rc; ++si)
            for (int zc : chunks) {
                const int* sh = shapes[si];
                b->bx = sh[0];
                b->by = sh[1];
                b->zchunk = zc;
                if ((rc = rebuild(b))) break;
                // The second-order pass alone, writing only the spare state
                // buffers and partials: the committed states do not move.
                HIP_TRY(enqueue_3d(b, st));
                HIP_TRY(hipEventRecord(e0, st));
                for (int r = 0; r < reps; ++r) HIP_TRY(enqueue_3d(b, st));
                HIP_TRY(hipEventRecord(e1, st));
                HIP_TRY(hipEventSynchronize(e1));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
                const double per = ms / reps;
                if (best < 0 || per < best) {
                    best = per;
                    bbx = sh[0];
                    bby = sh[1];
                    bz = zc;
                }
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        b->bx = bbx;
        b->by = bby;
        b->zchunk = bz;
        if ((rc = rebuild(b))) return rc;  // also clears the flags the trial passes raised
    }
    if (block_x) *block_x = b->bx;
    if (block_y) *block_y = b->by;
    if (zchunk) *zchunk = b->zchunk;
    if (best_ms) *best_ms = best;
    return TORCH_OK;
}

int torch_batch_time_pass(torch_batch* b, int reps, double* ms) {
    if (!b || reps <= 0 || !ms) return fail(TORCH_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    int rc;
    if (b->dirty && (rc = rebuild(b))) return rc;
    *ms = 0;
    if (b->seg[1] == b->seg[0]) return TORCH_OK;
    hipStream_t st = b->ctx->stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    hipError_t e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = enqueue_3d(b, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(TORCH_HIP_ERROR, hipGetErrorString(e));
    *ms = t / reps;
    // The trial passes may have raised flags the next real step raises anyway;
    // start it from clear ones as the post kernel leaves them.
    HIP_TRY(hipMemsetAsync(b->flags, 0, b->seg[3] * kFlagStride * sizeof(int), st));
    HIP_TRY(hipStreamSynchronize(st));
    return TORCH_OK;
}

int torch_batch_set_tiles(torch_batch* b, int block_x, int block_y, int zchunk) {
    if (!b) return fail(TORCH_INVALID, "null argument");
    const bool ok = (block_x == 64 && (block_y == 4 || block_y == 8)) ||
                    (block_x == 32 && (block_y == 8 || block_y == 16 || (block_y == 24 && b->precision == TORCH_FP32)));
    if (!ok || zchunk < 1) return fail(TORCH_INVALID, "unsupported tile shape");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    HIP_TRY(hipStreamSynchronize(b->ctx->stream));
    b->bx = block_x;
    b->by = block_y;
    b->zchunk = zchunk;
    b->tiles_set = true;
    b->dirty = true;
    drop_graph(b);
    return TORCH_OK;
}

int torch_batch_time_step_kernel(torch_batch* b, int steps, double* ms) {
    if (!b || steps <= 0 || !ms) return fail(TORCH_INVALID, "invalid argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    int rc;
    if (b->dirty && (rc = rebuild(b))) return rc;
    int lists = 0;
    for (int L = 0; L < 5; ++L) lists += b->hyd[0][L].blocks > 0;
    if (lists != 1) return fail(TORCH_INVALID, "step kernel timing needs exactly one 3D work list");
    if (!b->kev[0]) {
        HIP_TRY(hipEventCreate(&b->kev[0]));
        HIP_TRY(hipEventCreate(&b->kev[1]));
    }
    hipStream_t st = b->ctx->stream;
    HIP_TRY(hipStreamSynchronize(st));
    // A step graph with the event pair around the 3D work list, replayed one
    // step at a time so each pair is read before the next replay records it.
    b->kev_capture = true;
    rc = capture(b);
    b->kev_capture = false;
    double total = 0;
    for (int k = 0; !rc && k < steps; ++k) {
        hipError_t e = hipGraphLaunch(b->graph, st);
        if (e == hipSuccess) e = hipEventSynchronize(b->kev[1]);
        float t = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, b->kev[0], b->kev[1]);
        if (e != hipSuccess) rc = fail(TORCH_HIP_ERROR, std::string("step kernel timing: ") + hipGetErrorString(e));
        total += t;
    }
    if (!rc) rc = read_ctl(b);
    // Later runs replay the ordinary step graph (no event nodes).
    const int rc2 = capture(b);
    if (rc) return rc;
    if (rc2) return rc2;
    *ms = total / steps;
    return TORCH_OK;
}

int torch_batch_profile(torch_batch* b, int enable, double* ms, long* steps) {
    if (!b) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    HIP_TRY(hipSetDevice(b->ctx->device));
    if (enable && !b->ev0) {
        HIP_TRY(hipEventCreate(&b->ev0));
        HIP_TRY(hipEventCreate(&b->ev1));
    }
    if (ms) *ms = b->prof_ms;
    if (steps) *steps = b->prof_steps;
    b->prof_ms = 0;
    b->prof_steps = 0;
    b->profile = enable != 0;
    return TORCH_OK;
}

int torch_batch_graph_counts(torch_batch* b, long* instantiated, long* updated) {
    if (!b) return fail(TORCH_INVALID, "null argument");
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    if (instantiated) *instantiated = b->graph_builds;
    if (updated) *updated = b->graph_updates;
    return TORCH_OK;
}

}  // extern "C"
