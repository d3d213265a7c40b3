// pk_batch.hip -- many (ensemble, grid) pairs stepped together (no counterpart in pica).
//
// One launch per stage covers every ensemble of the batch: the particle and insert kernels walk a
// list of (ensemble, supercell) items, the field kernels find their grid from a block prefix.  A
// step is the single-ensemble pk_step sequence (zero J, particles, fold + fill J, B/2, E, B/2 with
// ghost fills, insert), 14 launches for the whole batch instead of 14 per ensemble; with
// PK_BATCH_GRAPH it is captured once into a HIP graph and replayed.
//
// Slack overflow without a host round trip: an insert that finds a supercell's slack exhausted
// leaves that supercell as it is and raises PK_HALT_PENDING; the last kernel of the step turns
// that into PK_HALT_STOP, which every kernel of the following steps checks first.  The
// host, after the queued steps, rebuilds the overflowed ensembles (their stayers are still in
// place and their leavers in the outbox), re-prepares and replays the remaining steps.
#include <algorithm>

#include "pk_internal.h"

namespace {

__global__ void k_batch_step_done(int* halt)
{
    if (threadIdx.x != 0 || halt[PK_HALT_STOP]) return;
    if (halt[PK_HALT_PENDING]) halt[PK_HALT_STOP] = 1;
    else halt[PK_HALT_DONE] += 1;
}

__global__ void k_batch_clear(int* halt)
{
    if (threadIdx.x < PK_HALT_N) halt[threadIdx.x] = 0;
}

constexpr unsigned MASK_J = (1u << PK_JX) | (1u << PK_JY) | (1u << PK_JZ);

// (ensemble, supercell) items: ensemble by ensemble in the XCD-aware supercell order of the
// single-ensemble kernels, or all items by decreasing particle count (largest first); and the
// particle kernel's items split by form factor in the same order
int build_items(pk_batch* b, int shape)
{
    std::vector<int2> items;
    std::vector<int32_t> key;
    const bool by_count = b->knob[PK_BATCH_ORDER] == 1;
    for (int i = 0; i < b->n(); i++) {
        const pk_ens* e = b->ens[i];
        // the counts only order the items (largest first); ensemble order needs none
        std::vector<int32_t> cnt(by_count ? (size_t)e->nsct : 0);
        if (by_count)
            PK_HIP(b->ctx, hipMemcpy(cnt.data(), e->cnt, sizeof(int32_t) * e->nsct, hipMemcpyDeviceToHost));
        const long chunk = (e->nsct + PK_XCD - 1) / PK_XCD;
        for (long w = 0; w < PK_XCD * chunk; w++) {
            const long s = (w % PK_XCD) * chunk + w / PK_XCD;
            if (s >= e->nsct) continue;
            items.push_back(make_int2(i, (int)s));
            if (by_count) key.push_back(cnt[(size_t)s]);
        }
    }
    if (by_count) {
        std::vector<size_t> idx(items.size());
        for (size_t k = 0; k < idx.size(); k++) idx[k] = k;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t c) { return key[a] > key[c]; });
        std::vector<int2> sorted(items.size());
        for (size_t k = 0; k < idx.size(); k++) sorted[k] = items[idx[k]];
        items.swap(sorted);
    }
    b->nitems = (long)items.size();
    std::vector<int2> cls[2];
    for (const int2& it : items) cls[pk_batch_shape_of(b, it.x, shape)].push_back(it);
    for (int k = 0; k < 2; k++) {
        b->cnitems[k] = (long)cls[k].size();
        if (int r = pk_batch_upload(b, (void**)&b->citems[k], cls[k].data(), sizeof(int2) * cls[k].size())) return r;
    }
    return pk_batch_upload(b, (void**)&b->items, items.data(), sizeof(int2) * items.size());
}

// enqueue one step of every ensemble
int enqueue_step(pk_batch* b)
{
    int r;
    if ((r = pk_batch_fields_launch(b, PK_BOP_ZERO_J, 0, 0))) return r;
    r = b->prec == PK_F64 ? pk_batch_f64_launch(b, b->dep) : pk_batch_split_launch(b, b->dep);
    if (r) return r;
    // the periodic step's field forms: J fold gathered per cell, then its fill; Yee updates that
    // write their own ghost images (pk_step's sequence, bit for bit)
    if ((r = pk_batch_fields_launch(b, PK_BOP_FOLD_J_P, 0, 0))) return r;
    if ((r = pk_batch_fields_launch(b, PK_BOP_FILL, MASK_J, 0))) return r;
    if ((r = pk_batch_fields_launch(b, PK_BOP_YEE_B_P, 0, 0))) return r;
    if ((r = pk_batch_fields_launch(b, PK_BOP_YEE_E_P, 0, 0))) return r;
    if ((r = pk_batch_fields_launch(b, PK_BOP_YEE_B_P, 0, 0))) return r;
    if ((r = pk_batch_insert_launch(b))) return r;
    hipLaunchKernelGGL(k_batch_step_done, dim3(1), dim3(64), 0, b->ctx->stream, b->halt);
    PK_HIP(b->ctx, hipGetLastError());
    return pk_batch_interleave_launch(b);
}

void drop_graph(pk_batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
}

// work lists, per-ensemble arguments and (with PK_BATCH_GRAPH) the captured step
int prepare(pk_batch* b, double dt, int shape, int dep)
{
    pk_ctx* c = b->ctx;
    PK_HIP(c, hipStreamSynchronize(c->stream));
    drop_graph(b);
    int r;
    if ((r = build_items(b, shape))) return r;
    if ((r = pk_batch_fields_prepare(b, dt))) return r;
    r = b->prec == PK_F64 ? pk_batch_f64_prepare(b, dt, shape, dep) : pk_batch_split_prepare(b, dt, shape, dep);
    if (r) return r;
    if ((r = pk_batch_insert_prepare(b))) return r;
    if ((r = pk_batch_interleave_prepare(b))) return r;
    b->dt = dt;
    b->shape = shape;
    b->dep = dep;
    if (b->knob[PK_BATCH_GRAPH]) {
        const bool timing = c->timing;
        c->timing = false;      // no event records inside the capture
        PK_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        r = enqueue_step(b);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
        c->timing = timing;
        if (r) {
            if (graph) (void)hipGraphDestroy(graph);
            return r;
        }
        if (ec != hipSuccess) return pk_fail(c, PK_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            b->exec = nullptr;
            return pk_fail(c, PK_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
    }
    b->dirty = false;
    return PK_OK;
}

int check_args(pk_batch* b, double dt, int shape, int dep)
{
    if ((shape != PK_CIC && shape != PK_TSC && shape != PK_SHAPE_PER_INSTANCE) ||
        (dep != PK_DIRECT && dep != PK_ESIRKEPOV))
        return pk_fail(b->ctx, PK_E_ARG, "pk_batch_run: unknown shape or depositor");
    if (shape == PK_SHAPE_PER_INSTANCE)
        for (int s : b->shapes)
            if (s != PK_CIC && s != PK_TSC)
                return pk_fail(b->ctx, PK_E_ARG, "pk_batch_run: PK_SHAPE_PER_INSTANCE needs pk_batch_set_shape for "
                                                 "every ensemble");
    if (!(dt > 0.0)) return pk_fail(b->ctx, PK_E_ARG, "pk_batch_run: dt must be positive");
    return PK_OK;
}

int run(pk_batch* b, double dt, int shape, int dep, int nsteps)
{
    pk_ctx* c = b->ctx;
    int r;
    for (pk_ens* e : b->ens)
        if ((r = pk_ens_resort(e))) return r;
    const int period = b->ens[0]->posmode == PK_POS_SPLIT
                           ? (b->knob[PK_BATCH_INTERLEAVE] >= 0 ? b->knob[PK_BATCH_INTERLEAVE] : c->opt[PK_OPT_INTERLEAVE])
                           : 0;
    if (b->dirty || dt != b->dt || shape != b->shape || dep != b->dep || period != b->ilperiod)
        if ((r = prepare(b, dt, shape, dep))) return r;
    // re-sort counts before this run; each pass tells the interleave kernel where it starts
    std::vector<long> r0(b->n());
    for (int j = 0; j < b->n(); j++) r0[j] = b->ens[j]->resorts;
    auto set_base = [&](int done) -> int {
        if (b->ilperiod <= 0) return PK_OK;
        for (int j = 0; j < b->n(); j++) b->ilbase_host[j] = r0[j] + done;
        PK_HIP(c, hipMemcpyAsync(b->ilbase, b->ilbase_host.data(), sizeof(int64_t) * b->n(), hipMemcpyHostToDevice,
                                 c->stream));
        return PK_OK;
    };
    hipLaunchKernelGGL(k_batch_clear, dim3(1), dim3(64), 0, c->stream, b->halt);
    PK_HIP(c, hipGetLastError());
    bool state_err = false;
    int done = 0;
    if ((r = set_base(0))) return r;
    while (done < nsteps) {
        {
            pk_ktimer tm(c, "batch_steps");
            for (int k = done; k < nsteps; k++) {
                if (b->exec) {
                    PK_HIP(c, hipGraphLaunch(b->exec, c->stream));
                } else if ((r = enqueue_step(b))) {
                    return r;
                }
            }
        }
        int32_t h[PK_HALT_N];
        PK_HIP(c, hipMemcpyAsync(h, b->halt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        PK_HIP(c, hipStreamSynchronize(c->stream));
        if (!h[PK_HALT_STOP]) break;
        // step done + h[PK_HALT_DONE] overflowed in its insert: finish its re-sort by rebuilds
        done += h[PK_HALT_DONE] + 1;
        for (pk_ens* e : b->ens) {
            int32_t fl[PK_NFLAGS];
            if ((r = pk_flags_read(e, fl))) return r;
            if (fl[PK_FLAG_ERR]) state_err = true;
            if (fl[PK_FLAG_OVERFLOW] && (r = pk_ens_rebuild_after_overflow(e))) return r;
            if ((r = pk_flags_clear(e))) return r;
        }
        // the interleave the overflowed step skipped, where its count is due
        for (int j = 0; j < b->n() && b->ilperiod > 0; j++)
            if ((r0[j] + done) % b->ilperiod == 0 && (r = pk_ens_interleave_launch(b->ens[j]))) return r;
        b->dirty = true;
        if (done < nsteps && (r = prepare(b, dt, shape, dep))) return r;
        hipLaunchKernelGGL(k_batch_clear, dim3(1), dim3(64), 0, c->stream, b->halt);
        PK_HIP(c, hipGetLastError());
        if ((r = set_base(done))) return r;
    }
    for (int j = 0; j < b->n(); j++) b->ens[j]->resorts = r0[j] + nsteps;
    for (pk_ens* e : b->ens) {
        int32_t fl[PK_NFLAGS];
        if ((r = pk_flags_read(e, fl))) return r;
        if (fl[PK_FLAG_ERR]) state_err = true;
        if ((r = pk_flags_clear(e))) return r;
    }
    if (state_err)
        return pk_fail(c, PK_E_STATE,
                       "batch step: a particle was not finite or moved further than its tile allows "
                       "(it was left unchanged)");
    return PK_OK;
}

// device copy of every ensemble's particles and every grid, for the autotuner's trials
struct Snapshot {
    struct Ens {
        void* main = nullptr;
        int64_t* off = nullptr;
        int32_t* cnt = nullptr;
        size_t capacity = 0;
        int maxcap = 0;
        int64_t np = 0;
        long resorts = 0;
        int ncps[3] = {0, 0, 0};
    };
    std::vector<Ens> ens;
    std::vector<void*> grids;
    ~Snapshot()
    {
        for (auto& s : ens) {
            if (s.main) (void)hipFree(s.main);
            if (s.off) (void)hipFree(s.off);
            if (s.cnt) (void)hipFree(s.cnt);
        }
        for (void* g : grids)
            if (g) (void)hipFree(g);
    }
};

size_t grid_bytes(const pk_grid* g) { return g->pitch * PK_NCOMP * g->esize(); }

int snapshot_take(pk_batch* b, Snapshot& S)
{
    pk_ctx* c = b->ctx;
    S.ens.resize(b->n());
    S.grids.assign(b->n(), nullptr);
    for (int i = 0; i < b->n(); i++) {
        pk_ens* e = b->ens[i];
        Snapshot::Ens& s = S.ens[i];
        const size_t mb = e->capacity * e->nfield() * e->esize();
        if (hipMalloc(&s.main, mb ? mb : 16) != hipSuccess || hipMalloc((void**)&s.off, sizeof(int64_t) * (e->nsct + 1)) != hipSuccess ||
            hipMalloc((void**)&s.cnt, sizeof(int32_t) * e->nsct) != hipSuccess ||
            hipMalloc(&S.grids[i], grid_bytes(b->grids[i])) != hipSuccess) {
            (void)hipGetLastError();
            return pk_fail(c, PK_E_ALLOC, "pk_batch_autotune: no device memory for the state copy");
        }
        PK_HIP(c, hipMemcpyAsync(s.main, e->main, mb, hipMemcpyDeviceToDevice, c->stream));
        PK_HIP(c, hipMemcpyAsync(s.off, e->off, sizeof(int64_t) * (e->nsct + 1), hipMemcpyDeviceToDevice, c->stream));
        PK_HIP(c, hipMemcpyAsync(s.cnt, e->cnt, sizeof(int32_t) * e->nsct, hipMemcpyDeviceToDevice, c->stream));
        PK_HIP(c, hipMemcpyAsync(S.grids[i], b->grids[i]->data, grid_bytes(b->grids[i]), hipMemcpyDeviceToDevice,
                                 c->stream));
        s.capacity = e->capacity;
        s.maxcap = e->maxcap;
        s.np = e->np;
        s.resorts = e->resorts;
        for (int d = 0; d < 3; d++) s.ncps[d] = e->ncps[d];
    }
    PK_HIP(c, hipStreamSynchronize(c->stream));
    return PK_OK;
}

int snapshot_restore(pk_batch* b, const Snapshot& S)
{
    pk_ctx* c = b->ctx;
    PK_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < b->n(); i++) {
        pk_ens* e = b->ens[i];
        const Snapshot::Ens& s = S.ens[i];
        int r;
        if ((e->ncps[0] != s.ncps[0] || e->ncps[1] != s.ncps[1] || e->ncps[2] != s.ncps[2]) &&
            (r = pk_ens_reshape(e, s.ncps)))
            return r;
        if (e->capacity != s.capacity && (r = pk_ens_alloc(e, s.capacity))) return r;
        const size_t mb = e->capacity * e->nfield() * e->esize();
        PK_HIP(c, hipMemcpyAsync(e->main, s.main, mb, hipMemcpyDeviceToDevice, c->stream));
        PK_HIP(c, hipMemcpyAsync(e->off, s.off, sizeof(int64_t) * (e->nsct + 1), hipMemcpyDeviceToDevice, c->stream));
        PK_HIP(c, hipMemcpyAsync(e->cnt, s.cnt, sizeof(int32_t) * e->nsct, hipMemcpyDeviceToDevice, c->stream));
        PK_HIP(c, hipMemcpyAsync(b->grids[i]->data, S.grids[i], grid_bytes(b->grids[i]), hipMemcpyDeviceToDevice,
                                 c->stream));
        e->maxcap = s.maxcap;
        e->np = s.np;
        e->resorts = s.resorts;
        e->pending_movers = false;
        e->failed = false;
        if ((r = pk_flags_clear(e))) return r;
    }
    PK_HIP(c, hipStreamSynchronize(c->stream));
    b->dirty = true;
    return PK_OK;
}

}  // namespace

int pk_batch_shape_of(const pk_batch* b, int i, int shape)
{
    return shape == PK_SHAPE_PER_INSTANCE ? b->shapes[(size_t)i] : shape;
}

int pk_batch_upload(pk_batch* b, void** dst, const void* src, size_t bytes)
{
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    if (bytes == 0) return PK_OK;
    if (hipMalloc(dst, bytes) != hipSuccess) {
        *dst = nullptr;
        (void)hipGetLastError();
        return pk_fail(b->ctx, PK_E_ALLOC, "batch: hipMalloc of a work list failed");
    }
    PK_HIP(b->ctx, hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return PK_OK;
}

extern "C" {

int pk_batch_create(pk_ctx* c, int n, pk_ens* const* ens, pk_grid* const* grids, pk_batch** out)
{
    if (!c || !out || n <= 0 || !ens || !grids) return pk_fail(c, PK_E_ARG, "pk_batch_create: bad arguments");
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!ens[i] || !grids[i] || ens[i]->ctx != c || grids[i]->ctx != c)
            return pk_fail(c, PK_E_ARG, "pk_batch_create: every ensemble and grid must belong to this context");
        int r = pk_check_pair(ens[i], grids[i]);
        if (r) return r;
        if (ens[i]->prec != ens[0]->prec) return pk_fail(c, PK_E_ARG, "pk_batch_create: mixed precisions");
        for (int d = 0; d < 3; d++)
            if (ens[i]->ncps[d] != ens[0]->ncps[d])
                return pk_fail(c, PK_E_ARG, "pk_batch_create: all ensembles need the same supercell shape");
        for (int j = 0; j < i; j++)
            if (ens[j] == ens[i] || grids[j] == grids[i])
                return pk_fail(c, PK_E_ARG, "pk_batch_create: an ensemble or grid appears twice");
    }
    pk_batch* b = new pk_batch();
    b->ctx = c;
    b->prec = ens[0]->prec;
    b->ens.assign(ens, ens + n);
    b->grids.assign(grids, grids + n);
    b->shapes.assign(n, -1);
    for (int d = 0; d < 3; d++) b->ncps0[d] = ens[0]->ncps[d];
    if (hipMalloc((void**)&b->halt, sizeof(int) * PK_HALT_N) != hipSuccess) {
        (void)hipGetLastError();
        delete b;
        return pk_fail(c, PK_E_ALLOC, "pk_batch_create: hipMalloc failed");
    }
    *out = b;
    return PK_OK;
}

void pk_batch_destroy(pk_batch* b)
{
    if (!b) return;
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    for (void* p : {(void*)b->items, (void*)b->citems[0], (void*)b->citems[1], b->pjobs, b->ijobs, b->fjobs,
                    (void*)b->fstart, (void*)b->halt, b->iljobs, (void*)b->ilbase})
        if (p) (void)hipFree(p);
    delete b;
}

int pk_batch_set(pk_batch* b, int knob, int value)
{
    if (!b) return PK_E_ARG;
    if (knob < 0 || knob >= PK_BATCH_NKNOB) return pk_fail(b->ctx, PK_E_ARG, "pk_batch_set: unknown knob");
    const int lo = knob == PK_BATCH_INTERLEAVE ? -1 : knob == PK_BATCH_BLOCK ? 128 : 0;
    const int hi = knob == PK_BATCH_WG_PER_CU || knob == PK_BATCH_INTERLEAVE ? 64
                   : knob == PK_BATCH_SUPERCELL ? 12 : knob == PK_BATCH_BLOCK ? 256 : 1;
    if (knob == PK_BATCH_BLOCK && value != 128 && value != 256)
        return pk_fail(b->ctx, PK_E_ARG, "pk_batch_set: block must be 128 or 256 threads");
    if (value < lo || value > hi) return pk_fail(b->ctx, PK_E_ARG, "pk_batch_set: value out of range");
    if (knob == PK_BATCH_SUPERCELL) {
        const int ncps[3] = {value ? value : b->ncps0[0], value ? value : b->ncps0[1], value ? value : b->ncps0[2]};
        for (pk_ens* e : b->ens)
            for (int d = 0; d < 3; d++)
                if (e->ncell[d] % ncps[d] != 0)
                    return pk_fail(b->ctx, PK_E_ARG, "pk_batch_set: a grid edge is not a multiple of the supercell size");
        for (pk_ens* e : b->ens)
            if (int r = pk_ens_set_supercells(e, ncps)) return r;
    }
    if (b->knob[knob] != value) b->dirty = true;
    b->knob[knob] = value;
    return PK_OK;
}

int pk_batch_set_shape(pk_batch* b, int index, int shape)
{
    if (!b) return PK_E_ARG;
    if (index < 0 || index >= b->n()) return pk_fail(b->ctx, PK_E_ARG, "pk_batch_set_shape: index out of range");
    if (shape != PK_CIC && shape != PK_TSC) return pk_fail(b->ctx, PK_E_ARG, "pk_batch_set_shape: unknown shape");
    if (b->shapes[(size_t)index] != shape) b->dirty = true;
    b->shapes[(size_t)index] = shape;
    return PK_OK;
}

int pk_batch_get(pk_batch* b, int knob, int* value)
{
    if (!b || !value) return PK_E_ARG;
    if (knob < 0 || knob >= PK_BATCH_NKNOB) return pk_fail(b->ctx, PK_E_ARG, "pk_batch_get: unknown knob");
    *value = b->knob[knob];
    return PK_OK;
}

int pk_batch_run(pk_batch* b, double dt, int shape, int dep, int nsteps)
{
    if (!b) return PK_E_ARG;
    int r = check_args(b, dt, shape, dep);
    if (r) return r;
    if (nsteps < 0) return pk_fail(b->ctx, PK_E_ARG, "pk_batch_run: negative step count");
    if (nsteps == 0) return PK_OK;
    return run(b, dt, shape, dep, nsteps);
}

int pk_batch_autotune(pk_batch* b, double dt, int shape, int dep, int trial_steps, double* best_ms)
{
    if (!b) return PK_E_ARG;
    int r = check_args(b, dt, shape, dep);
    if (r) return r;
    if (trial_steps < 1) return pk_fail(b->ctx, PK_E_ARG, "pk_batch_autotune: trial_steps must be >= 1");
    pk_ctx* c = b->ctx;
    for (pk_ens* e : b->ens)
        if ((r = pk_ens_resort(e))) return r;
    Snapshot S;
    if ((r = snapshot_take(b, S))) return r;
    hipEvent_t t0, t1;
    PK_HIP(c, hipEventCreate(&t0));
    if (hipEventCreate(&t1) != hipSuccess) {
        (void)hipEventDestroy(t0);
        return pk_fail(c, PK_E_HIP, "pk_batch_autotune: hipEventCreate failed");
    }
    int saved[PK_BATCH_NKNOB];
    for (int k = 0; k < PK_BATCH_NKNOB; k++) saved[k] = b->knob[k];
    // supercell edge of the state the autotuner starts from (0 when not cubic)
    const int* n0 = S.ens[0].ncps;
    const int edge0 = n0[0] == n0[1] && n0[1] == n0[2] ? n0[0] : 0;
    // one trial from the saved state: supercell edge sc (0 = the saved one), then the knobs, timed
    // over `steps` steps after one warm-up step; *ms = ms per step
    auto trial = [&](int sc, const int knobs[PK_BATCH_NKNOB], int steps, double* ms) -> int {
        int rr;
        if ((rr = snapshot_restore(b, S))) return rr;
        if (sc > 0 && sc != edge0) {
            const int n3[3] = {sc, sc, sc};
            for (pk_ens* e : b->ens)
                if ((rr = pk_ens_set_supercells(e, n3))) return rr;
        }
        for (int k = 0; k < PK_BATCH_NKNOB; k++)
            if (k != PK_BATCH_SUPERCELL) b->knob[k] = knobs[k];
        b->dirty = true;
        if ((rr = run(b, dt, shape, dep, 1))) return rr;     // warm-up, prepares and captures
        if (hipEventRecord(t0, c->stream) != hipSuccess) return PK_E_HIP;
        if ((rr = run(b, dt, shape, dep, steps))) return rr;
        if (hipEventRecord(t1, c->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess) return PK_E_HIP;
        float t = 0.f;
        (void)hipEventElapsedTime(&t, t0, t1);
        *ms = t / steps;
        return PK_OK;
    };
    int best[PK_BATCH_NKNOB];
    for (int k = 0; k < PK_BATCH_NKNOB; k++) best[k] = saved[k];
    best[PK_BATCH_GRAPH] = 1;
    int best_sc = 0;
    double best_t = 1e300;
    // a candidate that does not fit (supercell past 65,536 particles or the LDS) is skipped.  Each
    // stage compares its own candidates (the current best setting among them) over trials of one
    // length: best_t is reset when a stage starts.
    int steps = trial_steps;
    auto consider = [&](int sc, const int knobs[PK_BATCH_NKNOB]) -> int {
        double ms = 0.0;
        const int rr = trial(sc, knobs, steps, &ms);
        if (rr == PK_E_CAPACITY) return PK_OK;
        if (rr) return rr;
        if (ms < best_t) {
            best_t = ms;
            best_sc = sc;
            for (int k = 0; k < PK_BATCH_NKNOB; k++) best[k] = knobs[k];
        }
        return PK_OK;
    };
    // (1) supercell size
    {
        const int cand[] = {0, 2, 4, 8};
        for (int sc : cand) {
            if (sc > 0 && sc == edge0) continue;
            bool fits = true;
            for (pk_ens* e : b->ens)
                for (int d = 0; d < 3; d++)
                    if (sc > 0 && e->ncell[d] % sc != 0) fits = false;
            if (!fits) continue;
            int kn[PK_BATCH_NKNOB];
            for (int k = 0; k < PK_BATCH_NKNOB; k++) kn[k] = best[k];
            if ((r = consider(sc, kn))) break;
        }
    }
    // (2) interleave period (FP32 split ensembles).  The interleave restores an order that decays
    // over tens of steps, so these trials run at least 24 steps (a short trial would see only its
    // cost)
    if (!r && b->prec == PK_F32) {
        const int per[] = {0, 2, 3, 4, 6, 8};
        steps = trial_steps > 24 ? trial_steps : 24;
        best_t = 1e300;
        const int sc = best_sc;
        int kn0[PK_BATCH_NKNOB];
        for (int k = 0; k < PK_BATCH_NKNOB; k++) kn0[k] = best[k];
        for (int p : per) {
            int kn[PK_BATCH_NKNOB];
            for (int k = 0; k < PK_BATCH_NKNOB; k++) kn[k] = kn0[k];
            kn[PK_BATCH_INTERLEAVE] = p;
            if ((r = consider(sc, kn))) break;
        }
        steps = trial_steps;
    }
    // (3) block shape: four or two wavefronts per workgroup (FP32 split ensembles)
    if (!r && b->prec == PK_F32) {
        best_t = 1e300;
        int kn0[PK_BATCH_NKNOB];
        for (int k = 0; k < PK_BATCH_NKNOB; k++) kn0[k] = best[k];
        for (int bs : {PK_BLOCK, 128}) {
            int kn[PK_BATCH_NKNOB];
            for (int k = 0; k < PK_BATCH_NKNOB; k++) kn[k] = kn0[k];
            kn[PK_BATCH_BLOCK] = bs;
            if ((r = consider(best_sc, kn))) break;
        }
    }
    // (4) launch: graph, item order, persistence
    if (!r) {
        best_t = 1e300;
        const int wgs[] = {0, 1, 2, 4, 8};
        const int sc = best_sc;
        int kn0[PK_BATCH_NKNOB];
        for (int k = 0; k < PK_BATCH_NKNOB; k++) kn0[k] = best[k];
        for (int graph = 0; graph <= 1 && !r; graph++)
            for (int order = 0; order <= 1 && !r; order++)
                for (int wg : wgs) {
                    int kn[PK_BATCH_NKNOB];
                    for (int k = 0; k < PK_BATCH_NKNOB; k++) kn[k] = kn0[k];
                    kn[PK_BATCH_GRAPH] = graph;
                    kn[PK_BATCH_ORDER] = order;
                    kn[PK_BATCH_WG_PER_CU] = wg;
                    if ((r = consider(sc, kn))) break;
                }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    int rr = snapshot_restore(b, S);
    if (!rr && !r && best_sc > 0 && best_sc != edge0) {
        const int n3[3] = {best_sc, best_sc, best_sc};
        for (pk_ens* e : b->ens)
            if ((rr = pk_ens_set_supercells(e, n3))) break;
    }
    for (int k = 0; k < PK_BATCH_NKNOB; k++) b->knob[k] = r ? saved[k] : best[k];
    if (!r) b->knob[PK_BATCH_SUPERCELL] = best_sc > 0 ? best_sc : edge0;
    b->dirty = true;
    if (r) return r == PK_E_HIP ? pk_fail(c, PK_E_HIP, "pk_batch_autotune: event timing failed") : r;
    if (rr) return rr;
    if (best_t >= 1e300) return pk_fail(c, PK_E_CAPACITY, "pk_batch_autotune: no setting fits");
    if (best_ms) *best_ms = best_t;
    return PK_OK;
}

}  // extern "C"
