// pk_internal.h -- internal structures of libpica_amd (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/pica_amd.h"
#include "pk_fastdiv.h"

// pica's constants (Constants.h:10-13), pi deliberately the truncated literal pica uses.
#define PK_PI 3.14159265358
#define PK_C 29979245800.0

#define PK_BLOCK 256           // threads per particle workgroup (4 wavefronts of 64)
#define PK_XCD 8               // MI355X: 8 XCDs, workgroups are dealt round-robin over them
#define PK_MAX_SC_PARTICLES 65536
#define PK_DN_MAX_GAIN 1024     // entries of the denoiser's radial charge-filter gain table

struct pk_timer_rec {
    double ms = 0.0;
    long launches = 0;
};

struct pk_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<double> mass, charge;   // ParticleTypes table
    bool timing = false;
    struct pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<pending> pend;
    std::vector<hipEvent_t> evpool;
    std::map<std::string, pk_timer_rec> timers;
    int32_t* flags_host = nullptr;     // pinned, PK_NFLAGS
    hipEvent_t flags_ev = nullptr;     // recorded after a re-sort's flag copy (pk_ens_resort_begin)
    int opt[PK_OPT_N] = {PK_KERNEL_AUTO, 0, 3, 0};
};

enum { PK_FLAG_ERR = 0, PK_FLAG_OVERFLOW = 1, PK_NFLAGS = 4 };

struct pk_grid {
    pk_ctx* ctx = nullptr;
    int prec = PK_F64;
    int s[3] = {0, 0, 0};
    double origin[3] = {0, 0, 0};
    double step[3] = {1, 1, 1};
    int ghost = 0;
    size_t n = 0;             // elements per component
    size_t pitch = 0;         // elements between components (aligned)
    void* data = nullptr;     // PK_NCOMP components
    double* scratch = nullptr;  // FP64 reduction scratch (partials + results)
    size_t scratch_n = 0;
    double* gauss_ref = nullptr;  // reference residual for PK_DIAG_GAUSS_DRIFT
    double* rho_prev = nullptr;   // saved FP64 charge density (n) of the continuity projection
    double last_dt = 0.0;         // dt of the particle step whose current J holds (continuity diagnostic)
    double* proj = nullptr;       // projection work: complex physical volume + twiddles
    size_t proj_n = 0;
    size_t esize() const { return prec == PK_F64 ? 8 : 4; }
    void* comp(int c) const { return (char*)data + (size_t)c * pitch * esize(); }
};

struct pk_ens {
    pk_ctx* ctx = nullptr;
    int prec = PK_F64;
    int posmode = PK_POS_ABS;
    int type = 0;
    int ncell[3] = {0, 0, 0}, ncps[3] = {0, 0, 0}, nsc[3] = {0, 0, 0};
    double step[3] = {1, 1, 1};
    long nsct = 0;
    double slack_frac = 0.25;
    int slack_add = 32;
    // particle buffers (SoA, nfield arrays of `capacity` elements of esize bytes)
    size_t capacity = 0;
    void* main = nullptr;
    void* outbox = nullptr;
    uint8_t* odir = nullptr;     // direction class of each outbox entry
    int64_t* off = nullptr;      // nsct + 1 (capacity prefix)
    int32_t* cnt = nullptr;      // particles per supercell
    int32_t* nstay = nullptr;    // after the particle step: stayers (compacted)
    int32_t* nmove = nullptr;    // after the particle step: movers in the outbox
    int32_t* mvc = nullptr;      // nsct * 27 movers per direction class
    int32_t* flags = nullptr;    // PK_NFLAGS device flags
    int maxcap = 0;
    int64_t np = 0;
    bool pending_movers = false;  // particle step done, re-sort not yet done
    bool resort_inflight = false; // insert and flag copy launched, flags not yet read
    bool failed = false;          // a re-sort rebuild failed: contents inconsistent until upload / load
    long resorts = 0;             // re-sorts done (the interleave period counts them)
    int nfield() const { return prec == PK_F64 ? 7 : 8; }
    size_t esize() const { return prec == PK_F64 ? 8 : 4; }
    void* field(void* buf, int f) const { return (char*)buf + (size_t)f * capacity * esize(); }
};

// ---- error helpers ----
int pk_fail(pk_ctx* c, int code, const std::string& msg);
#define PK_HIP(ctx, call)                                                                      \
    do {                                                                                       \
        hipError_t _e = (call);                                                                \
        if (_e != hipSuccess)                                                                  \
            return pk_fail((ctx), PK_E_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

// ---- per-kernel event timing (RAII around a launch) ----
struct pk_ktimer {
    pk_ctx* c;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    pk_ktimer(pk_ctx* ctx, const char* nm);
    ~pk_ktimer();
};

// ---- device views passed to kernels by value ----
template <class T>
struct GridView {
    T* f[PK_NCOMP];
    int s0, s1, s2;
    int G;
    double step[3];
    double origin[3];
};

// forward-difference divergence of J at physical node (i, j, k) of an n[0] x n[1] x n[2] periodic
// grid, FP64, from the physical cells alone (the neighbour past the last cell is the first one,
// the value a filled ghost holds, so J's ghosts need not be filled)
template <class T>
__device__ __forceinline__ double pk_div_j(const GridView<T>& g, const int n[3], int i, int j, int k)
{
    const int ip = i + 1 == n[0] ? 0 : i + 1, jp = j + 1 == n[1] ? 0 : j + 1, kp = k + 1 == n[2] ? 0 : k + 1;
    const auto at = [&](int a, int b, int c) { return ((long)(a + g.G) * g.s1 + (b + g.G)) * g.s2 + (c + g.G); };
    const long c = at(i, j, k);
    double r = ((double)g.f[PK_JX][at(ip, j, k)] - (double)g.f[PK_JX][c]) / g.step[0];
    r += ((double)g.f[PK_JY][at(i, jp, k)] - (double)g.f[PK_JY][c]) / g.step[1];
    r += ((double)g.f[PK_JZ][at(i, j, kp)] - (double)g.f[PK_JZ][c]) / g.step[2];
    return r;
}

template <class T>
GridView<T> grid_view(const pk_grid* g)
{
    GridView<T> v;
    for (int c = 0; c < PK_NCOMP; c++) v.f[c] = (T*)g->comp(c);
    v.s0 = g->s[0]; v.s1 = g->s[1]; v.s2 = g->s[2];
    v.G = g->ghost;
    for (int d = 0; d < 3; d++) { v.step[d] = g->step[d]; v.origin[d] = g->origin[d]; }
    return v;
}

// particle arrays of one buffer (main, outbox or staging)
struct PartF64 {
    double *x, *y, *z, *ux, *uy, *uz, *w;
};
struct PartSplit {
    int32_t* cell;
    float *ox, *oy, *oz, *ux, *uy, *uz, *w;
};

inline PartF64 part_f64(const pk_ens* e, void* buf)
{
    PartF64 p;
    p.x = (double*)e->field(buf, 0); p.y = (double*)e->field(buf, 1); p.z = (double*)e->field(buf, 2);
    p.ux = (double*)e->field(buf, 3); p.uy = (double*)e->field(buf, 4); p.uz = (double*)e->field(buf, 5);
    p.w = (double*)e->field(buf, 6);
    return p;
}
inline PartSplit part_split(const pk_ens* e, void* buf)
{
    PartSplit p;
    p.cell = (int32_t*)e->field(buf, 0);
    p.ox = (float*)e->field(buf, 1); p.oy = (float*)e->field(buf, 2); p.oz = (float*)e->field(buf, 3);
    p.ux = (float*)e->field(buf, 4); p.uy = (float*)e->field(buf, 5); p.uz = (float*)e->field(buf, 6);
    p.w = (float*)e->field(buf, 7);
    return p;
}

// supercell geometry shared by the ensemble kernels
struct EnsGeom {
    int nsc[3], ncps[3], ncell[3];
    long nsct;
    double step[3];
    int xcd_chunk;   // ceil(nsct / PK_XCD)
    FastDiv div12, div2;   // by ncell[1] * ncell[2] and by ncell[2]
};

// 3D cell of a linear physical cell index 0 <= c < 2^31 (z fastest)
// position of the k-th (from 0) set bit of m, k < popcount(m): popcount halving, 6 steps
__device__ __forceinline__ int pk_select_bit(uint64_t m, int k)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const int c = __popcll(m & ((1ull << w) - 1ull));
        if (k >= c) {
            k -= c;
            m >>= w;
            pos += w;
        }
    }
    return pos;
}

__device__ __forceinline__ void cell_coords(int c, const EnsGeom& eg, int cc[3])
{
    const int q12 = (int)fastdiv((uint32_t)c, eg.div12), q2 = (int)fastdiv((uint32_t)c, eg.div2);
    cc[0] = q12;
    cc[1] = q2 - q12 * eg.ncell[1];
    cc[2] = c - q2 * eg.ncell[2];
}
EnsGeom ens_geom(const pk_ens* e);

// XCD-aware supercell of a workgroup: the workgroups one XCD receives (b % 8 == x) get one
// contiguous slab of supercells, so tile halos and J flushes stay in that XCD's L2.
__device__ __forceinline__ long pk_xcd_supercell(long b, int chunk)
{
    return (b % PK_XCD) * (long)chunk + b / PK_XCD;
}

// ---- many ensembles stepped together (pk_batch_*) ----
// Work of one stage for every ensemble goes into one launch: the particle and insert kernels
// walk a list of (ensemble, supercell) items, the field kernels a per-ensemble block prefix.
// work ranges of the batched field kernels: all cells, ghost cells, physical cells, physical cells
// within the ghost width of a face
enum { PK_BSTAGE_CELLS = 0, PK_BSTAGE_FILL = 1, PK_BSTAGE_PHYS = 2, PK_BSTAGE_FACE = 3, PK_BSTAGES = 4 };
// the periodic step's field operations: J zero, ghost fill, Yee updates that write their own ghost
// images, the J fold gathered per cell
enum { PK_BOP_ZERO_J = 0, PK_BOP_FILL, PK_BOP_YEE_E_P, PK_BOP_YEE_B_P, PK_BOP_FOLD_J_P };
// device halt words of a batch: every batched kernel returns at once while STOP is set; an insert
// overflow sets PENDING, which the step's last kernel turns into STOP (else it counts the step in DONE)
enum { PK_HALT_STOP = 0, PK_HALT_DONE = 1, PK_HALT_PENDING = 2, PK_HALT_N = 4 };

struct pk_batch {
    pk_ctx* ctx = nullptr;
    int prec = PK_F64;
    std::vector<pk_ens*> ens;
    std::vector<pk_grid*> grids;
    int knob[PK_BATCH_NKNOB] = {0, 1, 1, 0, -1, PK_BLOCK};
    int ncps0[3] = {0, 0, 0};  // supercell shape the ensembles had when the batch was created
    std::vector<int> shapes;   // per-ensemble form factor for PK_SHAPE_PER_INSTANCE (-1: unset)
    // prepared launch state, valid while !dirty and (dt, shape, dep) are unchanged
    bool dirty = true;
    double dt = 0.0;
    int shape = -1, dep = -1;
    int2* items = nullptr;     // (ensemble, supercell) in launch order
    long nitems = 0;
    void* pjobs = nullptr;     // particle-kernel arguments, one entry per ensemble
    void* ijobs = nullptr;     // insert-kernel arguments
    void* fjobs = nullptr;     // field-kernel arguments
    long* fstart = nullptr;    // PK_BSTAGES x (n + 1) block prefixes of the field stages
    long fblocks[PK_BSTAGES] = {};
    // particle-kernel items of each form factor (PK_CIC, PK_TSC), one launch per non-empty class,
    // and the dynamic LDS of each class (its largest ensemble)
    int2* citems[2] = {nullptr, nullptr};
    long cnitems[2] = {0, 0};
    size_t cplds[2] = {0, 0};
    int* halt = nullptr;       // device halt words (PK_HALT_*)
    void* iljobs = nullptr;    // interleave-kernel arguments (FP32 split ensembles)
    int64_t* ilbase = nullptr; // per ensemble: re-sort count at the start of the current pass
    std::vector<int64_t> ilbase_host;
    int ilperiod = 0;          // interleave period captured by the prepared step (0: none)
    hipGraphExec_t exec = nullptr;
    int n() const { return (int)ens.size(); }
};
// replace a device buffer by a copy of host data (outside any capture)
int pk_batch_upload(pk_batch* b, void** dst, const void* src, size_t bytes);
int pk_batch_fields_prepare(pk_batch* b, double dt);
int pk_batch_fields_launch(pk_batch* b, int op, unsigned mask, int d);
int pk_batch_shape_of(const pk_batch* b, int i, int shape);   // form factor of ensemble i in a run
int pk_batch_split_prepare(pk_batch* b, double dt, int shape, int dep);
int pk_batch_split_launch(pk_batch* b, int dep);
int pk_batch_f64_prepare(pk_batch* b, double dt, int shape, int dep);
int pk_batch_f64_launch(pk_batch* b, int dep);
int pk_ens_reshape(pk_ens* e, const int ncps[3]);   // new supercell arrays and geometry, no particles moved
int pk_batch_insert_prepare(pk_batch* b);
int pk_batch_insert_launch(pk_batch* b);
int pk_batch_interleave_prepare(pk_batch* b);
int pk_batch_interleave_launch(pk_batch* b);
int pk_ens_interleave_launch(pk_ens* e);   // reorder now, no re-sort first
int pk_check_pair(pk_ens* e, pk_grid* g);   // the periodic step's grid/ensemble requirements
int pk_flags_read(pk_ens* e, int32_t out[PK_NFLAGS]);
// pk_ens_resort in two halves, so that independent work can be queued between them:
// _begin launches the insert and the copy of the flags (no wait); _end waits for that copy,
// rebuilds after an overflow, reports a failed particle and launches the interleave when due
int pk_ens_resort_begin(pk_ens* e);
// the FP64 ensemble's periodic reorder (PK_OPT_INTERLEAVE_F64)
int pk_ens_interleave_f64_launch(pk_ens* e);
int pk_ens_resort_end(pk_ens* e);
int pk_flags_clear(pk_ens* e);
int pk_ens_usable(pk_ens* e);               // PK_E_STATE after a failed re-sort rebuild
int pk_ens_rebuild_after_overflow(pk_ens* e);   // rebuild of an overflowed re-sort; marks failure

// ---- launchers implemented in the .hip files ----
int pk_field_zero(pk_grid* g, int comp);
// Jx, Jy, Jz zeroed in one launch (they are consecutive components of the grid's block)
int pk_field_zero_j(pk_grid* g);
// pk_dn_apply; with fill = false J's ghosts are left as they were
int pk_dn_apply_impl(pk_dn* d, pk_grid* g, bool fill, double* div0 = nullptr);
int pk_field_yee_e(pk_grid* g, double dt);
int pk_field_yee_b(pk_grid* g, double dt);
int pk_field_ghost(pk_grid* g, int ops);
// the periodic step's Yee update (E when e_update, else B): the whole-array update followed by the
// fill of that field's ghosts, in one launch over the physical cells
#define PK_MAX_PERIODIC_GHOST 2
int pk_field_yee_periodic(pk_grid* g, double dt, bool e_update);
// the periodic step's J fold (physical cells bit-identical to PK_FOLD_J, in one launch; the ghosts
// then hold stale values), followed by PK_FILL_J when fill
int pk_field_fold_fill_j(pk_grid* g, bool fill);
int pk_particles_launch(pk_ens* e, pk_grid* g, double dt, int shape, int dep);
int pk_ens_insert_launch(pk_ens* e);
int pk_ens_rebuild(pk_ens* e, int source);   // source: 0 = main+outbox, 1 = staging (upload)
int pk_ens_upload_impl(pk_ens* e, size_t n, const double* const* host);
int pk_ens_download_impl(pk_ens* e, double* const* host, int64_t* off);
int pk_ens_cells_impl(pk_ens* e, int32_t* counts, int32_t* cells);
int pk_ens_load_impl(pk_ens* e, int ppc, double uth, double drift, double weight, uint64_t seed, int nmodes,
                     const double* modes);
int pk_ens_alloc(pk_ens* e, size_t capacity);
int pk_check_flags(pk_ens* e, bool* overflow);
int pk_diag_impl(pk_ens* e, pk_grid* g, int shape, double* out, bool set_ref);
int pk_diag_static_init(pk_ctx* c, int device);
int pk_diag_energies_static_impl(pk_ens* e, pk_grid* g, double out[2]);
int pk_rho_impl(pk_ens* e, pk_grid* g, int shape);
int pk_rho_fp64(pk_ens* e, pk_grid* g, int shape);   // folded, filled FP64 rho in g->scratch[0, n)
