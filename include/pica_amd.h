/*
 * pica_amd.h -- C ABI of the MI355X (gfx950) particle-in-cell step.
 *
 * Drop-in boundary for pica's hot path.  pica itself is a header-only C++ template library
 * whose per-particle functors (FieldInterpolatorCIC::get, BorisPusher::push,
 * CurrentDepositorCIC::deposit) cannot cross a host->device boundary one particle at a time,
 * so this ABI works at grid x ensemble granularity.  The C++ surface in
 * include/pica_amd/pica.hpp keeps pica's type and method names on top of it.
 *
 * Conventions
 *  - Every call returns PK_OK (0) or a negative PK_E_* code; pk_last_error() gives the text.
 *    No exception crosses the ABI.
 *  - Arrays are pica's Array3d layout: z fastest, data[(i*ny + j)*nz + k] (Array.h:140-144).
 *  - Units are pica's CGS units (Constants.h:8-19); pi is pica's truncated literal.
 *  - One HIP stream per context; calls on one context are ordered; a context is not
 *    thread-safe.  Device buffers are owned by the library.
 */
#ifndef PICA_AMD_H
#define PICA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define PK_OK 0
#define PK_E_ARG (-1)         /* invalid argument or mismatched geometry */
#define PK_E_HIP (-2)         /* HIP runtime error */
#define PK_E_ALLOC (-3)       /* device or host allocation failed */
#define PK_E_STATE (-4)       /* a particle left its tile (non-finite or faster than the tile allows) */
#define PK_E_UNSUPPORTED (-5) /* combination not implemented */
#define PK_E_CAPACITY (-6)    /* fixed-capacity buffer too small */

/* ---- enums ---- */
enum { PK_F64 = 0, PK_F32 = 1 };                  /* arithmetic type of grid / ensemble */
enum { PK_POS_ABS = 0, PK_POS_SPLIT = 1 };        /* absolute positions, or int32 cell + float offset */
enum { PK_CIC = 0, PK_TSC = 1 };                  /* particle form factor */
enum { PK_DIRECT = 0, PK_ESIRKEPOV = 1 };         /* current depositor */
enum { PK_EX = 0, PK_EY, PK_EZ, PK_BX, PK_BY, PK_BZ, PK_JX, PK_JY, PK_JZ, PK_RHO, PK_NCOMP };
enum { PK_H2D = 0, PK_D2H = 1 };
enum { PK_FILL_E = 1, PK_FILL_B = 2, PK_FILL_J = 4, PK_FOLD_J = 8, PK_FOLD_RHO = 16 };

/* pk_diag output slots (all FP64) */
enum {
    PK_DIAG_FIELD_ENERGY = 0,   /* sum_cells (E^2 + B^2) V / (8 pi) */
    PK_DIAG_KINETIC_ENERGY,     /* sum_p w m c^2 (gamma - 1) */
    PK_DIAG_CHARGE,             /* sum_cells rho V */
    PK_DIAG_GAUSS_MAX,          /* max |div E - 4 pi rho| over cells */
    PK_DIAG_GAUSS_DRIFT,        /* max |(div E - 4 pi rho) - reference residual| (0 without a reference) */
    PK_DIAG_RHO_MAX,            /* max |4 pi rho| (scale for the residuals) */
    PK_DIAG_CONTINUITY_MAX,     /* max |(rho - rho_saved)/dt + div J| over cells: rho_saved is the density
                                   kept by the last pk_rho_save, dt and J those of the last particle step
                                   on the grid (0 without a saved density) */
    PK_DIAG_DIVJ_MAX,           /* max |div J| (scale for the continuity residual) */
    PK_DIAG_N
};

typedef struct pk_ctx pk_ctx;
typedef struct pk_grid pk_grid;
typedef struct pk_ens pk_ens;

/* ---- context ---- */
int pk_ctx_create(int device, pk_ctx** out);
void pk_ctx_destroy(pk_ctx* ctx);
const char* pk_last_error(const pk_ctx* ctx);
int pk_sync(pk_ctx* ctx);
/* the hipStream_t all work of this context is issued on (for events / interop) */
void* pk_ctx_stream(pk_ctx* ctx);
/* per-kernel HIP-event timing: when enabled, every kernel launch of this context is bracketed
 * by events on its stream; pk_timer_get returns the accumulated ms and launch count for a
 * kernel name ("particles", "insert", "yee_e", "yee_b", "ghost", "rebuild", ...). */
int pk_timer_enable(pk_ctx* ctx, int on);
int pk_timer_get(pk_ctx* ctx, const char* kernel, double* total_ms, long* launches);
int pk_timer_reset(pk_ctx* ctx);
/* tuning options of a context (no counterpart in pica, whose loops are the caller's):
 * PK_OPT_PARTICLE_KERNEL selects the FP32 Esirkepov particle kernel: PK_KERNEL_AUTO (the
 * per-particle kernel), PK_KERNEL_PER_PARTICLE (one lane per particle, every contribution an LDS
 * atomic) or PK_KERNEL_CELL_SORTED (particles sorted by cell in LDS, per-cell register sums;
 * 2^3 or 4^3 supercells).
 * PK_OPT_CHANGE_QUEUE caps the cell-sorted kernel's LDS queue of cell-changing particles
 * (0 = as many entries as the LDS budget allows; beyond the queue a lane deposits at once).
 * PK_OPT_INTERLEAVE (default 3): every that many re-sorts of an FP32 ensemble, pk_ens_resort
 * reorders each supercell's particles: pairs of particles of one cell in adjacent slots, ordered
 * by (pair rank within the cell, cell), then the odd particle of every odd-count cell.  A lane of
 * the particle kernel takes two adjacent slots and deposits a same-cell pair with one set of LDS
 * atomics, and the lanes of a wavefront sit in different cells.  0 turns it off.  Supercells of
 * more than 64 cells or 4096 particles, or with a cell of more than 128 particles, are left as
 * they are.  Batched stepping (pk_batch_run) reorders at the same re-sort counts.
 * PK_OPT_INTERLEAVE_F64 (default 0, off): the same for FP64 ensembles, whose kernel takes one
 * particle per lane: rank-major order (the r-th particles of every cell, by cell, for r = 0, 1,
 * ...), as the loader leaves them.  Off by default so that FP64 runs keep the oracle's re-sort
 * order (pica's own arithmetic does not depend on the order); batched stepping ignores it. */
enum { PK_OPT_PARTICLE_KERNEL = 0, PK_OPT_CHANGE_QUEUE = 1, PK_OPT_INTERLEAVE = 2, PK_OPT_INTERLEAVE_F64 = 3,
       PK_OPT_N };
enum { PK_KERNEL_AUTO = 0, PK_KERNEL_PER_PARTICLE = 1, PK_KERNEL_CELL_SORTED = 2 };
int pk_ctx_set_option(pk_ctx* ctx, int option, int value);
int pk_ctx_get_option(pk_ctx* ctx, int option, int* value);

/* ParticleTypes table (Particle.h:21-31): species masses and charges (CGS). */
int pk_species_set(pk_ctx* ctx, int ntypes, const double* mass, const double* charge);

/* ---- grid: replaces YeeGrid<Three, Real>(origin, step, size), YeeGrid.h:191-256 ----
 * size = array size incl. ghosts; ghost = periodic ghost width used by pk_ghost and the
 * particle stage (2 for the periodic PIC step, origin = -2 step; 0 for a plain pica grid). */
int pk_grid_create(pk_ctx* ctx, int prec, const int size[3], const double origin[3], const double step[3],
                   int ghost, pk_grid** out);
void pk_grid_destroy(pk_grid* g);
/* copy one component (PK_EX..PK_RHO) between host and device; host element type = grid prec */
int pk_grid_copy(pk_grid* g, int comp, void* host, int dir);
int pk_grid_zero(pk_grid* g, int comp);
/* YeeSolver::updateE / updateB (YeeSolver.h:14-18,132-213): pica's semantics on the whole
 * array, including its open-edge rules. */
int pk_yee_update_e(pk_grid* g, double dt);
int pk_yee_update_b(pk_grid* g, double dt);
/* periodic ghost fill / fold (PK_FILL_* | PK_FOLD_*), width = the grid's ghost */
int pk_ghost(pk_grid* g, int ops);

/* ---- ensemble: replaces Ensemble<ParticleArraySoA<Particle3d>, Supercells>::Type,
 * EnsembleSupercells(min, max, numCells, numCellsPerSupercell) (EnsembleSupercells.h:25-34).
 * The domain is [0, ncell*step); ncell must be a multiple of ncps.  One species per ensemble.
 * prec/posmode: (PK_F64, PK_POS_ABS) or (PK_F32, PK_POS_SPLIT). */
int pk_ens_create(pk_ctx* ctx, int prec, int posmode, const int ncell[3], const double step[3],
                  const int ncps[3], int type_index, pk_ens** out);
void pk_ens_destroy(pk_ens* e);
/* slack policy of the fixed per-supercell capacity: cap = count + count*frac + add */
int pk_ens_set_slack(pk_ens* e, double frac, int add);
/* EnsembleSupercells::add for n particles (any order); replaces the current contents.
 * Positions absolute (CGS), u = p/(mc) as ParticleRef::setP takes it, w = factor. */
int pk_ens_upload(pk_ens* e, size_t n, const double* x, const double* y, const double* z, const double* ux,
                  const double* uy, const double* uz, const double* w);
/* particles in supercell order; arrays must hold pk_ens_size() entries; supercell offsets
 * (nsupercells + 1 entries) written when off != NULL */
int pk_ens_download(pk_ens* e, double* x, double* y, double* z, double* ux, double* uy, double* uz,
                    double* w, int64_t* off);
int64_t pk_ens_size(pk_ens* e);
int64_t pk_ens_num_supercells(pk_ens* e);
/* per-cell particle counts (physical cells, z fastest) and per-particle cell keys in
 * supercell order (linear physical cell index), either may be NULL */
int pk_ens_cells(pk_ens* e, int32_t* cell_counts, int32_t* cell_of_particle);
/* reorder every supercell of an FP32 ensemble now, as PK_OPT_INTERLEAVE does periodically
 * (no counterpart in pica); a no-op for FP64 ensembles */
int pk_ens_interleave(pk_ens* e);
/* re-bucket every particle into supercells of ncps cells (ncell must be a multiple; at most 12
 * per side), as if the ensemble had been created with them (EnsembleSupercells.h:25-34 takes the
 * supercell size at construction; pica cannot change it).  Order inside a new supercell: by old
 * supercell, then old slot (FP32 ensembles are then interleaved).  Used by pk_batch_autotune. */
int pk_ens_set_supercells(pk_ens* e, const int ncps[3]);
/* seeded thermal loader on the device: ppc particles per cell at uniform random positions,
 * u ~ N(0, uth) per component, plus +drift_x on even and -drift_x on odd particle ids. */
int pk_ens_load_thermal(pk_ens* e, int ppc, double uth, double drift_x, double weight, uint64_t seed);
/* the same plus nmodes velocity modes, modes[7 m .. 7 m + 6] = (kx, ky, kz, ax, ay, az, phase):
 * u += (ax, ay, az) sin(2 pi (kx x / Lx + ky y / Ly + kz z / Lz) + phase), L = ncell * step */
int pk_ens_load_modes(pk_ens* e, int ppc, double uth, double drift_x, double weight, uint64_t seed, int nmodes,
                      const double* modes);

/* ---- snapshots (no counterpart in pica): raw files for restart, fixtures and training data.
 * One 160-byte little-endian header ("PKSNAP1\0", kind 1 grid / 2 ensemble, version 1, prec,
 * posmode, dims[3], ghost or species, ncps[3], nfield, origin[3], step[3], count, nsct, esize,
 * re-sort count)
 * followed by raw arrays: a grid's PK_NCOMP components over its whole storage (ghosts included);
 * an ensemble's nsct int32 supercell counts, then its fields in supercell order (F64: x y z ux uy
 * uz w as float64; F32 split: cell as int32, ox oy oz ux uy uz w as float32).  A load restores
 * the bits and the order exactly; the target must have the same geometry, species and
 * precision (else PK_E_ARG). */
int pk_grid_save(pk_grid* g, const char* path);
int pk_grid_load(pk_grid* g, const char* path);
int pk_ens_save(pk_ens* e, const char* path);   /* re-sorts pending movers first */
int pk_ens_load(pk_ens* e, const char* path);

/* ---- many ensembles stepped together (no counterpart in pica, whose loops are the caller's) ----
 * A batch steps n (ensemble, grid) pairs of one precision and one supercell shape, with uneven
 * grid sizes and particle counts, doing each stage of pk_step for all of them in one launch
 * (particles and re-sort over a list of (ensemble, supercell) items), optionally captured once
 * into a HIP graph and replayed every step.  The ensembles and grids stay owned by the caller
 * and must outlive the batch; they can still be used one by one between batch calls. */
typedef struct pk_batch pk_batch;
enum {
    PK_BATCH_WG_PER_CU = 0,   /* particle kernel: 0 = one workgroup per item, k = persistent, k per CU */
    PK_BATCH_ORDER = 1,       /* items: 0 = ensemble by ensemble, 1 = most particles first */
    PK_BATCH_GRAPH = 2,       /* 1 = replay a captured HIP graph per step, 0 = launch eagerly */
    PK_BATCH_SUPERCELL = 3,   /* k > 0: every ensemble re-bucketed into k^3 supercells (pk_ens_set_supercells)
                                 when set; 0 = the supercells the ensembles had when the batch was created */
    PK_BATCH_INTERLEAVE = 4,  /* FP32 interleave period of the batched step: -1 = the context's
                                 PK_OPT_INTERLEAVE, 0 = off, k = every k-th re-sort */
    PK_BATCH_BLOCK = 5,       /* FP32 particle kernel workgroup: 256 threads (four wavefronts take the
                                 128-slot groups of a supercell in turn) or 128 (two wavefronts) */
    PK_BATCH_NKNOB
};
/* shape argument of pk_batch_run / pk_batch_autotune: every ensemble with its own form factor, set
 * by pk_batch_set_shape (the particle stage then takes one launch per form factor present) */
enum { PK_SHAPE_PER_INSTANCE = 2 };
int pk_batch_create(pk_ctx* ctx, int n, pk_ens* const* ens, pk_grid* const* grids, pk_batch** out);
void pk_batch_destroy(pk_batch* b);
int pk_batch_set(pk_batch* b, int knob, int value);
/* form factor (PK_CIC / PK_TSC) of ensemble `index` of the batch, used when a run passes
 * PK_SHAPE_PER_INSTANCE */
int pk_batch_set_shape(pk_batch* b, int index, int shape);
int pk_batch_get(pk_batch* b, int knob, int* value);
/* nsteps periodic PIC steps (pk_step's sequence) of every ensemble of the batch */
int pk_batch_run(pk_batch* b, double dt, int shape, int dep, int nsteps);
/* autotuner: from the current state, restored after every trial, times trial_steps steps (after
 * one warm-up step) of (1) every supercell size 2^3, 4^3, 8^3 that divides all grids (and the
 * batch's original one), (2) FP32: interleave periods 0, 2, 3, 4, 6, 8 (over at least 24 steps:
 * its gain builds up over tens of steps), (3) FP32: workgroups of 256 and 128 threads, (4) every
 * (graph, order, workgroups per CU) of {0,1} x {0,1} x {0,1,2,4,8}; each stage starts from the
 * best setting of the previous ones and keeps its fastest candidate (the last stage's ms per step
 * in *best_ms; the ensembles end re-bucketed to the chosen supercells) */
int pk_batch_autotune(pk_batch* b, double dt, int shape, int dep, int trial_steps, double* best_ms);

/* ---- learned form (no counterpart in pica) ----
 * A current denoiser: periodic conv3d 3 -> 16 -> 16 -> 3 with 3x3x3 taps and ReLU, BF16 MFMA
 * with FP32 accumulation, trained with Adam on pairs of (low-ppc J, high-ppc J) taken from plain
 * runs; applied to the deposited J of a low-ppc run, followed by a projection that restores
 * discrete charge continuity against the particles' own charge density.
 * Parameters (FP32, pk_dn_num_params() = 9539 values, in this order): W1[27][3][16], b1[16],
 * W2[27][16][16], b2[16], W3[27][16][3], b3[3]; tap t = (dx+1)*9 + (dy+1)*3 + (dz+1) reads the
 * input at (x+dx, y+dy, z+dz).  Host volumes are [3][n0][n1][n2] FP32, z fastest, periodic, in
 * network units (J times the denoiser's scale). */
typedef struct pk_dn pk_dn;
int pk_dn_create(pk_ctx* ctx, pk_dn** out);
void pk_dn_destroy(pk_dn* d);
int pk_dn_num_params(void);
int pk_dn_set_params(pk_dn* d, const float* params);   /* also resets the Adam state */
int pk_dn_init(pk_dn* d, uint64_t seed);                 /* He-normal weights, zero biases */
int pk_dn_get_params(pk_dn* d, float* params);
int pk_dn_set_scale(pk_dn* d, double scale);            /* network units = scale x J (CGS) */
/* charge filter S of pk_step_denoised (Esirkepov): a radial filter applied in Fourier space with
 * gain g(|k|), |k| in radians per cell, given on nbins points evenly spaced over [0, kmax] (linear
 * interpolation; the last value beyond kmax); gain[0] must be 1 (the filter keeps the total
 * charge).  The projection then makes D J = -S (rho1 - rho0) / dt: continuity and Gauss's law hold
 * for the filtered charge S rho.  nbins = 0 removes the filter (the particles' own charge). */
int pk_dn_set_charge_gain(pk_dn* d, int nbins, double kmax, const double* gain);
/* one forward pass on a host volume; y = network output [3][n] */
int pk_dn_forward_host(pk_dn* d, const int n[3], const float* x, float* y);
/* one convolution layer (1..3) on host input [cin][n]: pre-activation output [cout][n] */
int pk_dn_layer_host(pk_dn* d, int layer, const int n[3], const float* in, float* out);
/* gradient of the mean squared error of forward(x) against y; grad has pk_dn_num_params() values */
int pk_dn_grad_host(pk_dn* d, const int n[3], const float* x, const float* y, float* grad, double* loss);
/* the same over nsamples volumes of size n at once: x, y are [nsamples][3][n], the loss is the mean
   over all of them (the mini-batch form that pk_dn_train runs) */
int pk_dn_grad_batch_host(pk_dn* d, const int n[3], int nsamples, const float* x, const float* y, float* grad,
                          double* loss);
/* training pairs: the physical J of two grids of equal size (low-ppc input, high-ppc target) */
int pk_dn_add_pair(pk_dn* d, pk_grid* low, pk_grid* high);
/* ncrops training pairs cut from two periodic grids of equal size: each an edge^3 crop of both
   grids' physical J at one offset (drawn from seed and the crop's index, wrapping periodically;
   written to offsets[3 i .. 3 i + 2] when offsets is not NULL). The crop is not periodic, so the
   training loss leaves out its voxels within 3 (the network's receptive radius) of a face.
   7 <= edge <= every grid edge. Lets a mini-batch draw small volumes from large runs. */
int pk_dn_add_pair_crops(pk_dn* d, pk_grid* low, pk_grid* high, int ncrops, int edge, uint64_t seed, int* offsets);
int pk_dn_num_pairs(pk_dn* d);
int pk_dn_clear_pairs(pk_dn* d);
/* iters Adam steps on mini-batches of `batch` pairs drawn with `seed`; per-step mean loss to losses */
int pk_dn_train(pk_dn* d, int iters, int batch, double lr, uint64_t seed, float* losses);
/* replace the physical J of g by the denoised J (ghosts refilled) */
int pk_dn_apply(pk_dn* d, pk_grid* g);
/* charge continuity: pk_rho_save keeps the charge density of the ensemble's particles now (FP64);
 * pk_project_continuity deposits it again and subtracts grad(phi) from J, lap(phi) = div J +
 * (rho_now - rho_saved)/dt (7-point periodic Laplacian, FFT; power-of-two physical edges), so that
 * (rho_now - rho_saved)/dt + div J = 0; rho_now becomes the saved density. */
int pk_rho_save(pk_ens* e, pk_grid* g, int shape);
int pk_project_continuity(pk_ens* e, pk_grid* g, int shape, double dt);
/* pk_step with the denoiser and the projection applied to J before the field update (with
 * Esirkepov deposition, continuity for the denoiser's filtered charge, pk_dn_set_charge_filter) */
int pk_step_denoised(pk_ens* e, pk_grid* g, pk_dn* d, double dt, int shape, int dep);

/* ---- the step ----
 * Fused interpolation + Boris push + deposition for every particle (one workgroup per
 * supercell: FieldInterpolatorCICSupercell, BorisPusher::push, CurrentDepositorCICSupercell,
 * FieldInterpolatorYeeGrid.h:187-283, BorisPusher.h:32-47, CurrentDepositorYeeGrid.h:168-264).
 * J is accumulated into the grid (not zeroed); particles that leave their supercell are
 * extracted for pk_ens_resort. */
int pk_particle_step(pk_ens* e, pk_grid* g, double dt, int shape, int dep);
/* moves extracted particles into their new supercells (deterministic order) */
int pk_ens_resort(pk_ens* e);
/* a whole periodic PIC step: J = 0, particle step, fold J, B(dt/2), E(dt), B(dt/2) with ghost
 * fills, re-sort (its insert is queued ahead of the field update, which touches no particle, and
 * the host reads its overflow flags while the GPU runs the field update) */
int pk_step(pk_ens* e, pk_grid* g, double dt, int shape, int dep);
/* pk_step for several species on one grid (pica's ParticleTypes table, Particle.h:26-31, with
 * one ensemble per type set by pk_species_set): every ensemble's particle stage adds to J, then
 * one field update; ensembles may be electrons and mobile ions alike */
int pk_step_species(pk_ens* const* ens, int n, pk_grid* g, double dt, int shape, int dep);
int pk_run(pk_ens* e, pk_grid* g, double dt, int shape, int dep, int nsteps);
/* pk_run without a host synchronisation between steps (no pica counterpart: pica's loop is the
 * caller's): the step is captured once into a HIP graph and replayed nsteps times, a slack overflow
 * stops the replay on the device and is rebuilt once the run returns (pk_batch_run on a batch of
 * this one ensemble and grid).  For grids small enough that launches and per-step synchronisation
 * dominate; results equal pk_run's to the batched form's tolerance (atomic order). */
int pk_run_captured(pk_ens* e, pk_grid* g, double dt, int shape, int dep, int nsteps);

/* ---- diagnostics (FP64 reductions on the device) ----
 * rho of the ensemble with the given shape at cell centres into the grid's PK_RHO component
 * (folded and filled), then energies, charge and Gauss-law residual; out has PK_DIAG_N slots */
int pk_deposit_rho(pk_ens* e, pk_grid* g, int shape);
int pk_diag(pk_ens* e, pk_grid* g, int shape, double* out);
/* store the current div E - 4 pi rho as the reference for PK_DIAG_GAUSS_DRIFT */
int pk_diag_set_reference(pk_ens* e, pk_grid* g, int shape);
/* No counterpart in pica: for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  out[0] = PK_DIAG_FIELD_ENERGY, out[1] = PK_DIAG_KINETIC_ENERGY,
 * bit-identical to pk_diag's, computed without any allocation: the reduction partials live in a
 * static device array of 8.1 MB that each GPU holds for the life of the process.  Limits: at most
 * 1,048,576 supercells (else PK_E_CAPACITY); the ensemble must have no pending movers (a step
 * finished by pk_step or pk_ens_resort; else PK_E_STATE).  Calls share the GPU's array: calls
 * that overlap on one GPU (several contexts or threads on the same device) race. */
int pk_diag_energies_static(pk_ens* e, pk_grid* g, double* out);

#ifdef __cplusplus
}
#endif

#endif /* PICA_AMD_H */
