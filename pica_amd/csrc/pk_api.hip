// pk_api.hip -- the C ABI of libpica_amd (host side): contexts, grids, ensembles, the step.
#include <cmath>
#include <cstdio>

#include "pk_internal.h"

int pk_particles_f64(pk_ens* e, pk_grid* g, double dt, int shape, int dep);
int pk_particles_split(pk_ens* e, pk_grid* g, double dt, int shape, int dep);

int pk_fail(pk_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

pk_ktimer::pk_ktimer(pk_ctx* ctx, const char* nm) : c(ctx), name(nm)
{
    if (!c || !c->timing) return;
    for (hipEvent_t* ev : {&a, &b}) {
        if (!c->evpool.empty()) {
            *ev = c->evpool.back();
            c->evpool.pop_back();
        } else if (hipEventCreate(ev) != hipSuccess) {
            *ev = nullptr;
        }
    }
    if (a) (void)hipEventRecord(a, c->stream);
}

pk_ktimer::~pk_ktimer()
{
    if (!c || !c->timing || !a || !b) return;
    (void)hipEventRecord(b, c->stream);
    c->pend.push_back({name, a, b});
}

namespace {

__global__ void k_fill_flags(int32_t* f, int n)
{
    if ((int)threadIdx.x < n) f[threadIdx.x] = 0;
}

}  // namespace

int pk_check_pair(pk_ens* e, pk_grid* g)
{
    if (!e || !g) return PK_E_ARG;
    if (e->ctx != g->ctx) return pk_fail(e->ctx, PK_E_ARG, "ensemble and grid belong to different contexts");
    if (e->prec != g->prec) return pk_fail(e->ctx, PK_E_ARG, "ensemble and grid precision differ");
    if (g->ghost != 2) return pk_fail(e->ctx, PK_E_ARG, "the particle step needs a grid with 2 periodic ghost layers");
    for (int d = 0; d < 3; d++) {
        if (g->s[d] != e->ncell[d] + 4)
            return pk_fail(e->ctx, PK_E_ARG, "grid size must be ncell + 4 (2 ghost layers per side)");
        if (g->step[d] != e->step[d]) return pk_fail(e->ctx, PK_E_ARG, "grid and ensemble steps differ");
        if (g->origin[d] != -2.0 * g->step[d])
            return pk_fail(e->ctx, PK_E_ARG, "the particle step needs grid origin = -2 step (domain min 0)");
    }
    return PK_OK;
}

int pk_flags_read(pk_ens* e, int32_t out[PK_NFLAGS])
{
    pk_ctx* c = e->ctx;
    PK_HIP(c, hipMemcpyAsync(c->flags_host, e->flags, sizeof(int32_t) * PK_NFLAGS, hipMemcpyDeviceToHost, c->stream));
    PK_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < PK_NFLAGS; i++) out[i] = c->flags_host[i];
    return PK_OK;
}

int pk_ens_usable(pk_ens* e)
{
    if (!e->failed) return PK_OK;
    return pk_fail(e->ctx, PK_E_STATE,
                   "ensemble unusable: a re-sort could not rebuild its storage, so its particles are "
                   "inconsistent; upload or load it again");
}

int pk_ens_rebuild_after_overflow(pk_ens* e)
{
    // The insert kernel has already copied the overflowing supercell's leavers into its
    // neighbours, so a failed rebuild leaves particles duplicated or missing: refuse every later
    // use of the ensemble until new contents arrive.
    const int r = pk_ens_rebuild(e, 0);
    if (r) {
        e->failed = true;
        e->ctx->err += " (the ensemble is unusable until it is uploaded or loaded again)";
    }
    return r;
}

int pk_flags_clear(pk_ens* e)
{
    hipLaunchKernelGGL(k_fill_flags, dim3(1), dim3(64), 0, e->ctx->stream, e->flags, PK_NFLAGS);
    PK_HIP(e->ctx, hipGetLastError());
    return PK_OK;
}

extern "C" {

int pk_ctx_create(int device, pk_ctx** out)
{
    if (!out) return PK_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PK_E_HIP;
    if (device < 0 || device >= ndev) return PK_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return PK_E_HIP;
    pk_ctx* c = new pk_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&c->flags_host, sizeof(int32_t) * PK_NFLAGS, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&c->flags_ev, hipEventDisableTiming) != hipSuccess) {
        if (c->flags_host) (void)hipHostFree(c->flags_host);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return PK_E_HIP;
    }
    // default species table: one electron species (Constants.h:12-13)
    c->mass.push_back(9.10938215e-28);
    c->charge.push_back(-4.80320427e-10);
    // this device's static diagnostics scratch (pk_diag_energies_static), resolved once per device
    if (pk_diag_static_init(c, device) != PK_OK) {
        pk_ctx_destroy(c);
        return PK_E_HIP;
    }
    *out = c;
    return PK_OK;
}

void pk_ctx_destroy(pk_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->pend) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto ev : c->evpool) (void)hipEventDestroy(ev);
    if (c->flags_ev) (void)hipEventDestroy(c->flags_ev);
    if (c->flags_host) (void)hipHostFree(c->flags_host);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* pk_last_error(const pk_ctx* c) { return c ? c->err.c_str() : "null context"; }

int pk_sync(pk_ctx* c)
{
    if (!c) return PK_E_ARG;
    PK_HIP(c, hipStreamSynchronize(c->stream));
    return PK_OK;
}

void* pk_ctx_stream(pk_ctx* c) { return c ? (void*)c->stream : nullptr; }

int pk_timer_enable(pk_ctx* c, int on)
{
    if (!c) return PK_E_ARG;
    c->timing = on != 0;
    return PK_OK;
}

int pk_ctx_set_option(pk_ctx* c, int option, int value)
{
    if (!c) return PK_E_ARG;
    if (option == PK_OPT_PARTICLE_KERNEL) {
        if (value < PK_KERNEL_AUTO || value > PK_KERNEL_CELL_SORTED)
            return pk_fail(c, PK_E_ARG, "pk_ctx_set_option: unknown particle kernel");
        c->opt[option] = value;
        return PK_OK;
    }
    if (option == PK_OPT_CHANGE_QUEUE) {
        if (value < 0) return pk_fail(c, PK_E_ARG, "pk_ctx_set_option: negative queue size");
        c->opt[option] = value;
        return PK_OK;
    }
    if (option == PK_OPT_INTERLEAVE || option == PK_OPT_INTERLEAVE_F64) {
        if (value < 0) return pk_fail(c, PK_E_ARG, "pk_ctx_set_option: negative interleave period");
        c->opt[option] = value;
        return PK_OK;
    }
    return pk_fail(c, PK_E_ARG, "pk_ctx_set_option: unknown option");
}

int pk_ctx_get_option(pk_ctx* c, int option, int* value)
{
    if (!c || !value) return PK_E_ARG;
    if (option < 0 || option >= PK_OPT_N) return pk_fail(c, PK_E_ARG, "pk_ctx_get_option: unknown option");
    *value = c->opt[option];
    return PK_OK;
}

static int timer_collect(pk_ctx* c)
{
    if (c->pend.empty()) return PK_OK;
    PK_HIP(c, hipStreamSynchronize(c->stream));
    for (auto& p : c->pend) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->timers[p.name].ms += ms;
            c->timers[p.name].launches += 1;
        }
        c->evpool.push_back(p.a);
        c->evpool.push_back(p.b);
    }
    c->pend.clear();
    return PK_OK;
}

int pk_timer_get(pk_ctx* c, const char* kernel, double* total_ms, long* launches)
{
    if (!c || !kernel) return PK_E_ARG;
    int r = timer_collect(c);
    if (r) return r;
    auto it = c->timers.find(kernel);
    if (total_ms) *total_ms = it == c->timers.end() ? 0.0 : it->second.ms;
    if (launches) *launches = it == c->timers.end() ? 0 : it->second.launches;
    return PK_OK;
}

int pk_timer_reset(pk_ctx* c)
{
    if (!c) return PK_E_ARG;
    int r = timer_collect(c);
    c->timers.clear();
    return r;
}

int pk_species_set(pk_ctx* c, int ntypes, const double* mass, const double* charge)
{
    if (!c || ntypes <= 0 || !mass || !charge) return pk_fail(c, PK_E_ARG, "pk_species_set: bad arguments");
    c->mass.assign(mass, mass + ntypes);
    c->charge.assign(charge, charge + ntypes);
    return PK_OK;
}

// ---------------------------------------------------------------- grid
int pk_grid_create(pk_ctx* c, int prec, const int size[3], const double origin[3], const double step[3], int ghost,
                   pk_grid** out)
{
    if (!c || !out || !size || !origin || !step) return pk_fail(c, PK_E_ARG, "pk_grid_create: null argument");
    *out = nullptr;
    if (prec != PK_F64 && prec != PK_F32) return pk_fail(c, PK_E_ARG, "pk_grid_create: prec must be PK_F64 or PK_F32");
    for (int d = 0; d < 3; d++) {
        if (size[d] < 2) return pk_fail(c, PK_E_ARG, "pk_grid_create: size must be >= 2 in every dimension");
        if (!(step[d] > 0.0)) return pk_fail(c, PK_E_ARG, "pk_grid_create: step must be positive");
        // at least one physical cell: a 1D / 2D grid is a 3D grid with edges of one cell, whose
        // periodic ghosts all image that cell
        if (ghost < 0 || size[d] - 2 * ghost < 1) return pk_fail(c, PK_E_ARG, "pk_grid_create: bad ghost width");
    }
    pk_grid* g = new pk_grid();
    g->ctx = c;
    g->prec = prec;
    g->ghost = ghost;
    for (int d = 0; d < 3; d++) {
        g->s[d] = size[d];
        g->origin[d] = origin[d];
        g->step[d] = step[d];
    }
    g->n = (size_t)size[0] * size[1] * size[2];
    const size_t align = 256 / g->esize();
    g->pitch = (g->n + align - 1) / align * align;
    if (hipMalloc(&g->data, g->pitch * g->esize() * PK_NCOMP) != hipSuccess) {
        (void)hipGetLastError();
        delete g;
        return pk_fail(c, PK_E_ALLOC, "pk_grid_create: hipMalloc failed");
    }
    for (int comp = 0; comp < PK_NCOMP; comp++) {
        int r = pk_field_zero(g, comp);
        if (r) {
            pk_grid_destroy(g);
            return r;
        }
    }
    *out = g;
    return PK_OK;
}

void pk_grid_destroy(pk_grid* g)
{
    if (!g) return;
    (void)hipStreamSynchronize(g->ctx->stream);
    if (g->data) (void)hipFree(g->data);
    if (g->scratch) (void)hipFree(g->scratch);
    if (g->gauss_ref) (void)hipFree(g->gauss_ref);
    if (g->rho_prev) (void)hipFree(g->rho_prev);
    if (g->proj) (void)hipFree(g->proj);
    delete g;
}

int pk_grid_copy(pk_grid* g, int comp, void* host, int dir)
{
    if (!g || !host || comp < 0 || comp >= PK_NCOMP) return pk_fail(g ? g->ctx : nullptr, PK_E_ARG, "pk_grid_copy");
    pk_ctx* c = g->ctx;
    const size_t bytes = g->n * g->esize();
    if (dir == PK_H2D)
        PK_HIP(c, hipMemcpyAsync(g->comp(comp), host, bytes, hipMemcpyHostToDevice, c->stream));
    else if (dir == PK_D2H)
        PK_HIP(c, hipMemcpyAsync(host, g->comp(comp), bytes, hipMemcpyDeviceToHost, c->stream));
    else
        return pk_fail(c, PK_E_ARG, "pk_grid_copy: dir must be PK_H2D or PK_D2H");
    PK_HIP(c, hipStreamSynchronize(c->stream));
    return PK_OK;
}

int pk_grid_zero(pk_grid* g, int comp)
{
    if (!g || comp < 0 || comp >= PK_NCOMP) return PK_E_ARG;
    return pk_field_zero(g, comp);
}

int pk_yee_update_e(pk_grid* g, double dt)
{
    if (!g) return PK_E_ARG;
    return pk_field_yee_e(g, dt);
}

int pk_yee_update_b(pk_grid* g, double dt)
{
    if (!g) return PK_E_ARG;
    return pk_field_yee_b(g, dt);
}

int pk_ghost(pk_grid* g, int ops)
{
    if (!g) return PK_E_ARG;
    return pk_field_ghost(g, ops);
}

// ---------------------------------------------------------------- ensemble
int pk_ens_create(pk_ctx* c, int prec, int posmode, const int ncell[3], const double step[3], const int ncps[3],
                  int type_index, pk_ens** out)
{
    if (!c || !out || !ncell || !step || !ncps) return pk_fail(c, PK_E_ARG, "pk_ens_create: null argument");
    *out = nullptr;
    if (!((prec == PK_F64 && posmode == PK_POS_ABS) || (prec == PK_F32 && posmode == PK_POS_SPLIT)))
        return pk_fail(c, PK_E_UNSUPPORTED, "pk_ens_create: supported forms are (F64, ABS) and (F32, SPLIT)");
    if (type_index < 0 || type_index >= (int)c->mass.size())
        return pk_fail(c, PK_E_ARG, "pk_ens_create: type index outside the species table");
    long nsct = 1;
    for (int d = 0; d < 3; d++) {
        if (ncell[d] <= 0 || ncps[d] <= 0 || ncell[d] % ncps[d] != 0)
            return pk_fail(c, PK_E_ARG, "pk_ens_create: ncell must be a positive multiple of ncps");
        if (ncps[d] > 12) return pk_fail(c, PK_E_ARG, "pk_ens_create: supercells are limited to 12 cells per side");
        if (!(step[d] > 0.0)) return pk_fail(c, PK_E_ARG, "pk_ens_create: step must be positive");
        nsct *= ncell[d] / ncps[d];
    }
    if (nsct * 28 >= (1L << 31)) return pk_fail(c, PK_E_ARG, "pk_ens_create: too many supercells");
    pk_ens* e = new pk_ens();
    e->ctx = c;
    e->prec = prec;
    e->posmode = posmode;
    e->type = type_index;
    for (int d = 0; d < 3; d++) {
        e->ncell[d] = ncell[d];
        e->ncps[d] = ncps[d];
        e->nsc[d] = ncell[d] / ncps[d];
        e->step[d] = step[d];
    }
    e->nsct = nsct;
    bool ok = hipMalloc((void**)&e->off, sizeof(int64_t) * (nsct + 1)) == hipSuccess &&
              hipMalloc((void**)&e->cnt, sizeof(int32_t) * nsct) == hipSuccess &&
              hipMalloc((void**)&e->nstay, sizeof(int32_t) * nsct) == hipSuccess &&
              hipMalloc((void**)&e->nmove, sizeof(int32_t) * nsct) == hipSuccess &&
              hipMalloc((void**)&e->mvc, sizeof(int32_t) * nsct * 27) == hipSuccess &&
              hipMalloc((void**)&e->flags, sizeof(int32_t) * PK_NFLAGS) == hipSuccess;
    if (!ok) {
        (void)hipGetLastError();
        pk_ens_destroy(e);
        return pk_fail(c, PK_E_ALLOC, "pk_ens_create: hipMalloc failed");
    }
    // empty ensemble
    if (hipMemsetAsync(e->off, 0, sizeof(int64_t) * (nsct + 1), c->stream) != hipSuccess ||
        hipMemsetAsync(e->cnt, 0, sizeof(int32_t) * nsct, c->stream) != hipSuccess ||
        hipMemsetAsync(e->nstay, 0, sizeof(int32_t) * nsct, c->stream) != hipSuccess ||
        hipMemsetAsync(e->nmove, 0, sizeof(int32_t) * nsct, c->stream) != hipSuccess ||
        hipMemsetAsync(e->flags, 0, sizeof(int32_t) * PK_NFLAGS, c->stream) != hipSuccess ||
        pk_ens_alloc(e, 0) != PK_OK) {
        pk_ens_destroy(e);
        return pk_fail(c, PK_E_HIP, "pk_ens_create: initialisation failed");
    }
    *out = e;
    return PK_OK;
}

void pk_ens_destroy(pk_ens* e)
{
    if (!e) return;
    (void)hipStreamSynchronize(e->ctx->stream);
    for (void* p : {(void*)e->main, (void*)e->outbox, (void*)e->odir, (void*)e->off, (void*)e->cnt, (void*)e->nstay,
                    (void*)e->nmove, (void*)e->mvc, (void*)e->flags})
        if (p) (void)hipFree(p);
    delete e;
}

int pk_ens_set_slack(pk_ens* e, double frac, int add)
{
    if (!e || !(frac >= 0.0) || add < 0) return PK_E_ARG;
    e->slack_frac = frac;
    e->slack_add = add;
    return PK_OK;
}

int pk_ens_upload(pk_ens* e, size_t n, const double* x, const double* y, const double* z, const double* ux,
                  const double* uy, const double* uz, const double* w)
{
    if (!e || (n > 0 && (!x || !y || !z || !ux || !uy || !uz || !w))) return PK_E_ARG;
    if (n >= (size_t)1 << 40) return pk_fail(e->ctx, PK_E_ARG, "pk_ens_upload: too many particles");
    const double* h[7] = {x, y, z, ux, uy, uz, w};
    return pk_ens_upload_impl(e, n, h);
}

int pk_ens_resort(pk_ens* e)
{
    if (!e) return PK_E_ARG;
    int r = pk_ens_usable(e);
    if (r) return r;
    if (!e->pending_movers) return PK_OK;
    if ((r = pk_ens_resort_begin(e))) return r;
    return pk_ens_resort_end(e);
}

}  // extern "C"

int pk_ens_resort_begin(pk_ens* e)
{
    if (!e->pending_movers || e->resort_inflight) return PK_OK;
    int r = pk_ens_usable(e);
    if (r) return r;
    if ((r = pk_ens_insert_launch(e))) return r;
    pk_ctx* c = e->ctx;
    if (hipMemcpyAsync(c->flags_host, e->flags, sizeof(int32_t) * PK_NFLAGS, hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipEventRecord(c->flags_ev, c->stream) != hipSuccess) {
        // the insert is queued: running it again would insert its movers twice
        e->failed = true;
        return pk_fail(c, PK_E_HIP, "re-sort: reading the insert's flags failed; the ensemble is unusable until "
                                    "it is uploaded or loaded again");
    }
    e->resort_inflight = true;
    return PK_OK;
}

int pk_ens_resort_end(pk_ens* e)
{
    if (!e->resort_inflight) return PK_OK;
    e->resort_inflight = false;
    pk_ctx* c = e->ctx;
    if (hipEventSynchronize(c->flags_ev) != hipSuccess) {
        // the insert ran or may have run: running it again would insert its movers twice
        e->failed = true;
        return pk_fail(c, PK_E_HIP, "re-sort: waiting for the insert failed; the ensemble is unusable until it "
                                    "is uploaded or loaded again");
    }
    int32_t fl[PK_NFLAGS];
    for (int i = 0; i < PK_NFLAGS; i++) fl[i] = c->flags_host[i];
    int r;
    e->pending_movers = false;
    if (fl[PK_FLAG_OVERFLOW]) {
        r = pk_ens_rebuild_after_overflow(e);
        if (r) return r;
    }
    r = pk_flags_clear(e);
    if (r) return r;
    if (fl[PK_FLAG_ERR])
        return pk_fail(e->ctx, PK_E_STATE,
                       "particle step: a particle was not finite or moved further than its tile allows "
                       "(it was left unchanged)");
    e->resorts++;
    const int period = e->ctx->opt[e->posmode == PK_POS_SPLIT ? PK_OPT_INTERLEAVE : PK_OPT_INTERLEAVE_F64];
    if (period > 0 && e->resorts % period == 0)
        return e->posmode == PK_POS_SPLIT ? pk_ens_interleave_launch(e) : pk_ens_interleave_f64_launch(e);
    return PK_OK;
}

extern "C" {

int pk_ens_download(pk_ens* e, double* x, double* y, double* z, double* ux, double* uy, double* uz, double* w,
                    int64_t* off)
{
    if (!e) return PK_E_ARG;
    int r = pk_ens_resort(e);
    if (r) return r;
    double* h[7] = {x, y, z, ux, uy, uz, w};
    return pk_ens_download_impl(e, h, off);
}

int64_t pk_ens_size(pk_ens* e)
{
    if (!e) return -1;
    if (pk_ens_resort(e)) return -1;
    return e->np;
}

int64_t pk_ens_num_supercells(pk_ens* e) { return e ? e->nsct : -1; }

int pk_ens_cells(pk_ens* e, int32_t* counts, int32_t* cells)
{
    if (!e) return PK_E_ARG;
    int r = pk_ens_resort(e);
    if (r) return r;
    return pk_ens_cells_impl(e, counts, cells);
}

int pk_ens_load_thermal(pk_ens* e, int ppc, double uth, double drift_x, double weight, uint64_t seed)
{
    if (!e || ppc <= 0 || !(uth >= 0.0) || !(weight > 0.0)) return PK_E_ARG;
    return pk_ens_load_impl(e, ppc, uth, drift_x, weight, seed, 0, nullptr);
}

int pk_ens_load_modes(pk_ens* e, int ppc, double uth, double drift_x, double weight, uint64_t seed, int nmodes,
                      const double* modes)
{
    if (!e || ppc <= 0 || !(uth >= 0.0) || !(weight > 0.0) || nmodes < 0 || (nmodes > 0 && !modes)) return PK_E_ARG;
    return pk_ens_load_impl(e, ppc, uth, drift_x, weight, seed, nmodes, modes);
}

// ---------------------------------------------------------------- step
int pk_particle_step(pk_ens* e, pk_grid* g, double dt, int shape, int dep)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if ((shape != PK_CIC && shape != PK_TSC) || (dep != PK_DIRECT && dep != PK_ESIRKEPOV))
        return pk_fail(e->ctx, PK_E_ARG, "pk_particle_step: unknown shape or depositor");
    if (!(dt > 0.0)) return pk_fail(e->ctx, PK_E_ARG, "pk_particle_step: dt must be positive");
    if ((r = pk_ens_usable(e))) return r;
    if (e->pending_movers) {
        r = pk_ens_resort(e);
        if (r) return r;
    }
    r = e->prec == PK_F64 ? pk_particles_f64(e, g, dt, shape, dep) : pk_particles_split(e, g, dt, shape, dep);
    if (r) return r;
    e->pending_movers = true;
    g->last_dt = dt;
    return PK_OK;
}

int pk_step(pk_ens* e, pk_grid* g, double dt, int shape, int dep)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if ((r = pk_field_zero_j(g))) return r;
    if ((r = pk_particle_step(e, g, dt, shape, dep))) return r;
    if ((r = pk_field_fold_fill_j(g, true))) return r;
    // the re-sort's insert goes ahead of the field update (they touch disjoint data), so that the
    // GPU runs the field update while the host waits for the re-sort's flags
    if ((r = pk_ens_resort_begin(e))) return r;
    // B(dt/2), E(dt), B(dt/2) with ghost fills (the leapfrog around YeeSolver.h:28-38)
    if (!(r = pk_field_yee_periodic(g, dt * 0.5, false)) && !(r = pk_field_yee_periodic(g, dt, true)))
        r = pk_field_yee_periodic(g, dt * 0.5, false);
    const int rs = pk_ens_resort_end(e);
    return r ? r : rs;
}

int pk_step_species(pk_ens* const* ens, int n, pk_grid* g, double dt, int shape, int dep)
{
    if (!ens || n < 1 || !g) return PK_E_ARG;
    int r;
    for (int k = 0; k < n; k++)
        if ((r = pk_check_pair(ens[k], g))) return r;
    if ((r = pk_field_zero_j(g))) return r;
    // every species deposits into the same J (the particle stage adds to it)
    for (int k = 0; k < n; k++)
        if ((r = pk_particle_step(ens[k], g, dt, shape, dep))) return r;
    if ((r = pk_field_fold_fill_j(g, true))) return r;
    if ((r = pk_field_yee_periodic(g, dt * 0.5, false))) return r;
    if ((r = pk_field_yee_periodic(g, dt, true))) return r;
    if ((r = pk_field_yee_periodic(g, dt * 0.5, false))) return r;
    for (int k = 0; k < n; k++)
        if ((r = pk_ens_resort(ens[k]))) return r;
    return PK_OK;
}

int pk_run(pk_ens* e, pk_grid* g, double dt, int shape, int dep, int nsteps)
{
    for (int i = 0; i < nsteps; i++) {
        int r = pk_step(e, g, dt, shape, dep);
        if (r) return r;
    }
    return PK_OK;
}

int pk_run_captured(pk_ens* e, pk_grid* g, double dt, int shape, int dep, int nsteps)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if (nsteps < 0) return pk_fail(e->ctx, PK_E_ARG, "pk_run_captured: negative step count");
    if (nsteps == 0) return PK_OK;
    pk_batch* b = nullptr;
    pk_ens* const es[1] = {e};
    pk_grid* const gs[1] = {g};
    if ((r = pk_batch_create(e->ctx, 1, es, gs, &b))) return r;
    if (!(r = pk_batch_set(b, PK_BATCH_GRAPH, 1))) r = pk_batch_run(b, dt, shape, dep, nsteps);
    pk_batch_destroy(b);
    return r;
}

// ---------------------------------------------------------------- diagnostics
int pk_deposit_rho(pk_ens* e, pk_grid* g, int shape)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if ((r = pk_ens_resort(e))) return r;
    return pk_rho_impl(e, g, shape);
}

int pk_diag(pk_ens* e, pk_grid* g, int shape, double* out)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if (!out) return PK_E_ARG;
    if ((r = pk_ens_resort(e))) return r;
    return pk_diag_impl(e, g, shape, out, false);
}

int pk_diag_energies_static(pk_ens* e, pk_grid* g, double* out)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if (!out) return PK_E_ARG;
    if ((r = pk_ens_usable(e))) return r;
    if (e->pending_movers)
        return pk_fail(e->ctx, PK_E_STATE,
                       "pk_diag_energies_static: movers pending (a re-sort may allocate); call pk_ens_resort or "
                       "finish the step with pk_step first");
    return pk_diag_energies_static_impl(e, g, out);
}

int pk_diag_set_reference(pk_ens* e, pk_grid* g, int shape)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if ((r = pk_ens_resort(e))) return r;
    double out[PK_DIAG_N];
    return pk_diag_impl(e, g, shape, out, true);
}

}  // extern "C"
