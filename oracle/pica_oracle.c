/*
 * pica_oracle.c -- CPU oracle for the pica particle-in-cell step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load this library, and only as the checker; the product path
 * (pica_amd/) never links, loads or calls it.
 *
 * What it restates, operation by operation and in the same floating-point order
 * (compile with -ffp-contract=off so no FMA is formed):
 *   - YeeSolver::Implementation<Three>::updateE / updateB   YeeSolver.h:132-213
 *   - FieldInterpolatorCIC<YeeGrid<Three>>::get             FieldInterpolatorYeeGrid.h:29-39,150-182
 *   - FieldInterpolatorCICSupercell (gather + get)          FieldInterpolatorYeeGrid.h:197-283
 *   - CurrentDepositorCIC<YeeGrid<Three>>::deposit          CurrentDepositorYeeGrid.h:30-40,138-163
 *   - CurrentDepositorCICSupercell (zero, deposit, flush)   CurrentDepositorYeeGrid.h:178-262
 *   - BorisPusher::push (+ ParticleRef::setP/getVelocity)   BorisPusher.h:23-47, ParticleArray.h:135-141,209
 *   - EnsembleSupercells::getSupercellIndex                 EnsembleSupercells.h:25-34,53-57
 *   - Grid::getCellIndex                                    Grid.h:36
 *   - truncate / floor / inverse / cross / norm2            Vectors.h:388,569-574,652-697
 *   - constants (truncated pi, c)                           Constants.h:10-11
 *
 * Parts with no pica counterpart are this project's own definitions, written from the
 * published methods and pinned by invariant tests (tests/test_oracle_invariants.py):
 *   - TSC (quadratic B-spline) interpolation and direct deposition (Hockney & Eastwood),
 *   - Esirkepov charge-conserving deposition, CIC and TSC shapes (CPC 135 (2001) 144),
 *   - periodic ghost fill / J fold with two ghost layers (grid origin = -2 h),
 *   - the deterministic supercell re-sort rule (stayers first, then incoming movers by
 *     direction class), and rho deposition for the diagnostics;
 *   - the seeded thermal loader (or_load_thermal), the twin of the library's device loader,
 *     with its ln / sincos pinned against libm and closed forms (tests/test_oracle_loader.py).
 *
 * Storage follows pica's Array3d: z fastest, data[(i*ny + j)*nz + k] (Array.h:143).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358          /* Constants.h:10 (truncated literal, as pica uses it) */
#define OR_C 29979245800.0           /* Constants.h:11 */

enum { EX = 0, EY, EZ, BX, BY, BZ, JX, JY, JZ };
enum { SHAPE_CIC = 0, SHAPE_TSC = 1 };
enum { DEP_DIRECT = 0, DEP_ESIRKEPOV = 1 };

typedef struct {
    int s[3];            /* array size incl. ghosts (pica YeeGrid size) */
    double origin[3];    /* position of array index 0 */
    double step[3];
    double* f[9];        /* EX..JZ */
} or_grid;

typedef struct {
    int s[3];
    float origin[3];
    float step[3];
    float* f[9];
} or_gridf;

/* Supercell-ordered ensemble of one species.  Particles of supercell s occupy
 * [off[s], off[s+1]).  Positions are absolute; the physical domain is [0, ncell*step). */
typedef struct {
    int nsc[3];
    int ncps[3];
    int ncell[3];
    int ghost;
    double step[3];
    long np;
    double* x[3];
    double* u[3];
    double* w;
    long* off;           /* nsc_total + 1 */
    signed char* cls;    /* per particle direction class of the last push (13 = stayed) */
    double q, m;
} or_ens;

#define IDX3(s, i, j, k) ((((long)(i)) * (s)[1] + (j)) * (long)(s)[2] + (k))

/* ------------------------------------------------------------------------------------ */
/* Yee solver, pica semantics (YeeSolver.h:132-213), double and float.                   */
/* Every component is updated wherever its curl stencil is inside the array: this is the */
/* union of pica's main loop and its three edge loops, which use the same formula.      */
/* ------------------------------------------------------------------------------------ */
#define OR_DEFINE_YEE(NAME_E, NAME_B, GRID, T)                                              \
void NAME_E(GRID* g, T dt)                                                                  \
{                                                                                           \
    const T coeffCurrent = -(T)4 * (T)OR_PI * dt;                                           \
    const T cdt = (T)OR_C * dt;                                                             \
    const T kx = cdt / g->step[0], ky = cdt / g->step[1], kz = cdt / g->step[2];            \
    const int nx = g->s[0], ny = g->s[1], nz = g->s[2];                                     \
    T *ex = g->f[EX], *ey = g->f[EY], *ez = g->f[EZ];                                       \
    const T *bx = g->f[BX], *by = g->f[BY], *bz = g->f[BZ];                                 \
    const T *jx = g->f[JX], *jy = g->f[JY], *jz = g->f[JZ];                                 \
    _Pragma("omp parallel for collapse(2)")                                                 \
    for (int i = 0; i < nx; i++)                                                            \
    for (int j = 0; j < ny; j++)                                                            \
    for (int k = 0; k < nz; k++) {                                                          \
        const long c = IDX3(g->s, i, j, k);                                                 \
        if (j < ny - 1 && k < nz - 1)                                                       \
            ex[c] += coeffCurrent * jx[c] + ky * (bz[IDX3(g->s, i, j + 1, k)] - bz[c]) -    \
                kz * (by[IDX3(g->s, i, j, k + 1)] - by[c]);                                 \
        if (i < nx - 1 && k < nz - 1)                                                       \
            ey[c] += coeffCurrent * jy[c] + kz * (bx[IDX3(g->s, i, j, k + 1)] - bx[c]) -    \
                kx * (bz[IDX3(g->s, i + 1, j, k)] - bz[c]);                                 \
        if (i < nx - 1 && j < ny - 1)                                                       \
            ez[c] += coeffCurrent * jz[c] + kx * (by[IDX3(g->s, i + 1, j, k)] - by[c]) -    \
                ky * (bx[IDX3(g->s, i, j + 1, k)] - bx[c]);                                 \
    }                                                                                       \
}                                                                                           \
void NAME_B(GRID* g, T dt)                                                                  \
{                                                                                           \
    const T cdt = (T)OR_C * dt;                                                             \
    const T kx = cdt / g->step[0], ky = cdt / g->step[1], kz = cdt / g->step[2];            \
    const int nx = g->s[0], ny = g->s[1], nz = g->s[2];                                     \
    const T *ex = g->f[EX], *ey = g->f[EY], *ez = g->f[EZ];                                 \
    T *bx = g->f[BX], *by = g->f[BY], *bz = g->f[BZ];                                       \
    _Pragma("omp parallel for collapse(2)")                                                 \
    for (int i = 0; i < nx; i++)                                                            \
    for (int j = 0; j < ny; j++)                                                            \
    for (int k = 0; k < nz; k++) {                                                          \
        const long c = IDX3(g->s, i, j, k);                                                 \
        if (j >= 1 && k >= 1)                                                               \
            bx[c] += kz * (ey[c] - ey[IDX3(g->s, i, j, k - 1)]) -                           \
                ky * (ez[c] - ez[IDX3(g->s, i, j - 1, k)]);                                 \
        if (i >= 1 && k >= 1)                                                               \
            by[c] += kx * (ez[c] - ez[IDX3(g->s, i - 1, j, k)]) -                           \
                kz * (ex[c] - ex[IDX3(g->s, i, j, k - 1)]);                                 \
        if (i >= 1 && j >= 1)                                                               \
            bz[c] += ky * (ex[c] - ex[IDX3(g->s, i, j - 1, k)]) -                           \
                kx * (ey[c] - ey[IDX3(g->s, i - 1, j, k)]);                                 \
    }                                                                                       \
}

OR_DEFINE_YEE(or_yee_update_e, or_yee_update_b, or_grid, double)
OR_DEFINE_YEE(or_yee_update_e_f32, or_yee_update_b_f32, or_gridf, float)

/* ------------------------------------------------------------------------------------ */
/* Periodic ghosts, width G on both sides of every dimension.  Physical array indices     */
/* are [G, G+N), N >= 1.  fill: ghost := its periodic image, dimension by dimension over the full */
/* extent of the other two.  fold: physical += ghost image, low ghosts first, then high.   */
/* ------------------------------------------------------------------------------------ */
/* array index of the periodic image of index i; an edge of N < G cells wraps more than once */
static int or_ghost_image(int i, int G, int N)
{
    if (i >= G && i < G + N) return i;
    int r = (i - G) % N;
    if (r < 0) r += N;
    return G + r;
}

#define OR_DEFINE_GHOST(NAME_FILL, NAME_FOLD, T)                                            \
void NAME_FILL(const int s[3], int G, T* f)                                                 \
{                                                                                           \
    for (int d = 0; d < 3; d++) {                                                           \
        const int N = s[d] - 2 * G;                                                         \
        const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;                                       \
        for (int a = 0; a < G; a++)                                                         \
        for (int p = 0; p < s[o1]; p++)                                                     \
        for (int q = 0; q < s[o2]; q++) {                                                   \
            int ilo[3], isrc[3], ihi[3], isrc2[3];                                          \
            ilo[d] = a; isrc[d] = or_ghost_image(a, G, N);                                  \
            ihi[d] = G + N + a; isrc2[d] = or_ghost_image(G + N + a, G, N);                 \
            ilo[o1] = isrc[o1] = ihi[o1] = isrc2[o1] = p;                                   \
            ilo[o2] = isrc[o2] = ihi[o2] = isrc2[o2] = q;                                   \
            f[IDX3(s, ilo[0], ilo[1], ilo[2])] = f[IDX3(s, isrc[0], isrc[1], isrc[2])];     \
            f[IDX3(s, ihi[0], ihi[1], ihi[2])] = f[IDX3(s, isrc2[0], isrc2[1], isrc2[2])];  \
        }                                                                                   \
    }                                                                                       \
}                                                                                           \
void NAME_FOLD(const int s[3], int G, T* f)                                                 \
{                                                                                           \
    for (int d = 0; d < 3; d++) {                                                           \
        const int N = s[d] - 2 * G;                                                         \
        const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;                                       \
        for (int p = 0; p < s[o1]; p++)                                                     \
        for (int q = 0; q < s[o2]; q++) {                                                   \
            for (int a = 0; a < G; a++) {                                                   \
                int it[3], ig[3];                                                           \
                it[d] = or_ghost_image(a, G, N); ig[d] = a;                                 \
                it[o1] = ig[o1] = p; it[o2] = ig[o2] = q;                                   \
                f[IDX3(s, it[0], it[1], it[2])] += f[IDX3(s, ig[0], ig[1], ig[2])];         \
            }                                                                               \
            for (int a = 0; a < G; a++) {                                                   \
                int it[3], ig[3];                                                           \
                it[d] = or_ghost_image(G + N + a, G, N); ig[d] = G + N + a;                 \
                it[o1] = ig[o1] = p; it[o2] = ig[o2] = q;                                   \
                f[IDX3(s, it[0], it[1], it[2])] += f[IDX3(s, ig[0], ig[1], ig[2])];         \
            }                                                                               \
        }                                                                                   \
    }                                                                                       \
}

OR_DEFINE_GHOST(or_ghost_fill, or_ghost_fold, double)
OR_DEFINE_GHOST(or_ghost_fill_f32, or_ghost_fold_f32, float)

/* ------------------------------------------------------------------------------------ */
/* pica's plain (whole-grid) CIC interpolation and deposition, 3D.                       */
/* ------------------------------------------------------------------------------------ */
typedef struct { double inv[3], no[3], nos[3]; } or_norm;

/* getIndexCoeff, FieldInterpolatorYeeGrid.h:29-39 (identical in CurrentDepositorYeeGrid.h:30-40) */
static inline void or_index_coeff(const or_norm* nm, const double pos[3], int ic[3], double cc[3],
                                  int is[3], double cs[3])
{
    for (int d = 0; d < 3; d++) {
        const double np = pos[d] * nm->inv[d];
        double internal = np - nm->no[d];
        ic[d] = (int)internal;                       /* truncate, Vectors.h:673-679 */
        cc[d] = internal - (double)ic[d];
        internal = np - nm->nos[d];
        is[d] = (int)internal;
        cs[d] = internal - (double)is[d];
    }
}

/* 8-point CIC sum in pica's term order, FieldInterpolatorYeeGrid.h:171-182 / 269-281 */
static inline double or_cic8(const double* t, const int ts[3], const int b[3], const double c[3])
{
    const int i = b[0], j = b[1], k = b[2];
    return
        t[IDX3(ts, i, j, k)] * (1.0 - c[0]) * (1.0 - c[1]) * (1.0 - c[2]) +
        t[IDX3(ts, i, j, k + 1)] * (1.0 - c[0]) * (1.0 - c[1]) * c[2] +
        t[IDX3(ts, i, j + 1, k)] * (1.0 - c[0]) * c[1] * (1.0 - c[2]) +
        t[IDX3(ts, i, j + 1, k + 1)] * (1.0 - c[0]) * c[1] * c[2] +
        t[IDX3(ts, i + 1, j, k)] * c[0] * (1.0 - c[1]) * (1.0 - c[2]) +
        t[IDX3(ts, i + 1, j, k + 1)] * c[0] * (1.0 - c[1]) * c[2] +
        t[IDX3(ts, i + 1, j + 1, k)] * c[0] * c[1] * (1.0 - c[2]) +
        t[IDX3(ts, i + 1, j + 1, k + 1)] * c[0] * c[1] * c[2];
}

/* 8-point CIC scatter in pica's order, CurrentDepositorYeeGrid.h:153-163 / 251-262 */
static inline void or_cic8_add(double* t, const int ts[3], const int b[3], const double c[3], double v)
{
    const int i = b[0], j = b[1], k = b[2];
    t[IDX3(ts, i, j, k)] += (1.0 - c[0]) * (1.0 - c[1]) * (1.0 - c[2]) * v;
    t[IDX3(ts, i, j, k + 1)] += (1.0 - c[0]) * (1.0 - c[1]) * c[2] * v;
    t[IDX3(ts, i, j + 1, k)] += (1.0 - c[0]) * c[1] * (1.0 - c[2]) * v;
    t[IDX3(ts, i, j + 1, k + 1)] += (1.0 - c[0]) * c[1] * c[2] * v;
    t[IDX3(ts, i + 1, j, k)] += c[0] * (1.0 - c[1]) * (1.0 - c[2]) * v;
    t[IDX3(ts, i + 1, j, k + 1)] += c[0] * (1.0 - c[1]) * c[2] * v;
    t[IDX3(ts, i + 1, j + 1, k)] += c[0] * c[1] * (1.0 - c[2]) * v;
    t[IDX3(ts, i + 1, j + 1, k + 1)] += c[0] * c[1] * c[2] * v;
}

/* component -> (collocated? per dimension): E/J_x collocated in x, staggered in y,z, etc.
 * FieldInterpolatorYeeGrid.h:155-166, CurrentDepositorYeeGrid.h:143-148 */
static const int OR_COLLOC[6][3] = {
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1},   /* ex ey ez (and jx jy jz) */
    {0, 1, 1}, {1, 0, 1}, {1, 1, 0},   /* bx by bz */
};

static inline void or_pick(int comp, const int ic[3], const double cc[3], const int is[3],
                           const double cs[3], int b[3], double c[3])
{
    for (int d = 0; d < 3; d++) {
        b[d] = OR_COLLOC[comp][d] ? ic[d] : is[d];
        c[d] = OR_COLLOC[comp][d] ? cc[d] : cs[d];
    }
}

static void or_grid_norm(const or_grid* g, or_norm* nm)
{
    for (int d = 0; d < 3; d++) {
        nm->inv[d] = 1.0 / g->step[d];                          /* inverse(), Vectors.h:693-697 */
        nm->no[d] = g->origin[d] / g->step[d];                  /* FieldInterpolatorYeeGrid.h:26 */
        nm->nos[d] = nm->no[d] + 1.0 * 0.5;                     /* :27, ones * 0.5 */
    }
}

/* FieldInterpolatorCIC<YeeGrid<Three>>::get for n particles (FieldInterpolatorYeeGrid.h:150-167) */
void or_interp_cic(const or_grid* g, long n, const double* const x[3], double* e[3], double* b[3])
{
    or_norm nm;
    or_grid_norm(g, &nm);
    for (long p = 0; p < n; p++) {
        const double pos[3] = {x[0][p], x[1][p], x[2][p]};
        int ic[3], is[3], bi[3];
        double cc[3], cs[3], c[3];
        or_index_coeff(&nm, pos, ic, cc, is, cs);
        for (int comp = 0; comp < 6; comp++) {
            or_pick(comp, ic, cc, is, cs, bi, c);
            const double v = or_cic8(g->f[EX + comp], g->s, bi, c);
            if (comp < 3) e[comp][p] = v; else b[comp - 3][p] = v;
        }
    }
}

/* CurrentDepositorCIC<YeeGrid<Three>>::deposit for n particles (CurrentDepositorYeeGrid.h:138-163).
 * No 1/V normalisation, as in pica. */
void or_deposit_cic(or_grid* g, long n, const double* const x[3], const double* const cur[3])
{
    or_norm nm;
    or_grid_norm(g, &nm);
    for (long p = 0; p < n; p++) {
        const double pos[3] = {x[0][p], x[1][p], x[2][p]};
        int ic[3], is[3], bi[3];
        double cc[3], cs[3], c[3];
        or_index_coeff(&nm, pos, ic, cc, is, cs);
        for (int comp = 0; comp < 3; comp++) {
            or_pick(comp, ic, cc, is, cs, bi, c);
            or_cic8_add(g->f[JX + comp], g->s, bi, c, cur[comp][p]);
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* Boris pusher, BorisPusher.h:23-47 with ParticleArraySoA::ParticleRef (ParticleArray.h)   */
/* ------------------------------------------------------------------------------------ */
double or_boris_coeff(double q, double m, double dt)
{
    return q * dt / (2.0 * m * OR_C);                            /* BorisPusher.h:29 */
}

/* pushes one particle; u is p/(mc); returns velocity v (for the direct depositor's current) */
static inline void or_boris(double coeff, double dt, const double e[3], const double b[3],
                            double u[3], double x[3], double v[3])
{
    double eM[3], um[3], t[3], up[3], s[3], cr[3];
    for (int d = 0; d < 3; d++) eM[d] = e[d] * coeff;
    for (int d = 0; d < 3; d++) um[d] = u[d] + eM[d];
    const double um2 = um[0] * um[0] + um[1] * um[1] + um[2] * um[2];
    const double tk = coeff / sqrt(1.0 + um2);
    for (int d = 0; d < 3; d++) t[d] = b[d] * tk;
    cr[0] = um[1] * t[2] - um[2] * t[1];                         /* cross, Vectors.h:569-574 */
    cr[1] = um[2] * t[0] - um[0] * t[2];
    cr[2] = um[0] * t[1] - um[1] * t[0];
    for (int d = 0; d < 3; d++) up[d] = um[d] + cr[d];
    const double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    const double sk = 2.0 / (1.0 + t2);
    for (int d = 0; d < 3; d++) s[d] = t[d] * sk;
    cr[0] = up[1] * s[2] - up[2] * s[1];
    cr[1] = up[2] * s[0] - up[0] * s[2];
    cr[2] = up[0] * s[1] - up[1] * s[0];
    for (int d = 0; d < 3; d++) u[d] = um[d] + cr[d] + eM[d];
    const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double invGamma = 1.0 / sqrt(1.0 + u2);                /* setP, ParticleArray.h:262 */
    for (int d = 0; d < 3; d++) v[d] = u[d] * OR_C * invGamma;   /* getVelocity, ParticleArray.h:209 */
    for (int d = 0; d < 3; d++) x[d] = x[d] + v[d] * dt;         /* BorisPusher.h:44-46 */
}

void or_boris_push(long n, double coeff, double dt, double* x[3], double* u[3],
                   const double* const e[3], const double* const b[3])
{
    for (long p = 0; p < n; p++) {
        double xe[3] = {x[0][p], x[1][p], x[2][p]}, ue[3] = {u[0][p], u[1][p], u[2][p]};
        const double ee[3] = {e[0][p], e[1][p], e[2][p]}, be[3] = {b[0][p], b[1][p], b[2][p]};
        double v[3];
        or_boris(coeff, dt, ee, be, ue, xe, v);
        for (int d = 0; d < 3; d++) { x[d][p] = xe[d]; u[d][p] = ue[d]; }
    }
}

/* ------------------------------------------------------------------------------------ */
/* Index helpers                                                                        */
/* ------------------------------------------------------------------------------------ */
/* Grid::getCellIndex, Grid.h:36: truncate((x - origin) / step) */
void or_cell_index_grid(const or_grid* g, long n, const double* const x[3], int* out)
{
    for (long p = 0; p < n; p++)
        for (int d = 0; d < 3; d++)
            out[3 * p + d] = (int)((x[d][p] - g->origin[d]) / g->step[d]);
}

/* EnsembleSupercells::getSupercellIndex, EnsembleSupercells.h:31-33,53-57, with the
 * supercell geometry derived exactly as pica's constructor does.  Also used with ncps = 1
 * for the per-cell key (floor((x - min) * inverse(cellSize))). */
void or_supercell_index(const double mn[3], const double mx[3], const int ncell[3],
                        const int ncps[3], long n, const double* const x[3], int* out)
{
    double inv[3];
    for (int d = 0; d < 3; d++) {
        const double cellSize = (mx[d] - mn[d]) / (double)ncell[d];
        const double scSize = (double)ncps[d] * cellSize;
        inv[d] = 1.0 / scSize;
    }
    for (long p = 0; p < n; p++)
        for (int d = 0; d < 3; d++)
            out[3 * p + d] = (int)floor((x[d][p] - mn[d]) * inv[d]);
}

/* ------------------------------------------------------------------------------------ */
/* Supercell particle stage                                                             */
/* ------------------------------------------------------------------------------------ */
static inline void or_tsc_w(double d, double w[3])
{
    const double a = 0.5 - d, b = 0.5 + d;
    w[0] = 0.5 * a * a;
    w[1] = 0.75 - d * d;
    w[2] = 0.5 * b * b;
}

/* TSC 27-point gather: value = sum_a wx[a] * (sum_b wy[b] * (sum_c wz[c] * t)) */
static inline double or_tsc27(const double* t, const int ts[3], const int b[3], const double w[3][3])
{
    double sa = 0.0;
    for (int a = 0; a < 3; a++) {
        double sb = 0.0;
        for (int bb = 0; bb < 3; bb++) {
            double sc = 0.0;
            for (int c = 0; c < 3; c++)
                sc += w[2][c] * t[IDX3(ts, b[0] + a, b[1] + bb, b[2] + c)];
            sb += w[1][bb] * sc;
        }
        sa += w[0][a] * sb;
    }
    return sa;
}

static inline void or_tsc27_add(double* t, const int ts[3], const int b[3], const double w[3][3], double v)
{
    for (int a = 0; a < 3; a++)
        for (int bb = 0; bb < 3; bb++)
            for (int c = 0; c < 3; c++)
                t[IDX3(ts, b[0] + a, b[1] + bb, b[2] + c)] += w[0][a] * w[1][bb] * w[2][c] * v;
}

/* TSC node and weights for internal coordinate xi: nodes i0-1..i0+1, returns i0-1 */
static inline int or_tsc_node(double xi, double w[3])
{
    const int i0 = (int)floor(xi + 0.5);
    or_tsc_w(xi - (double)i0, w);
    return i0 - 1;
}

typedef struct {
    int shape, dep;
    double dt, coeff, qw_unused;
    /* interpolation tile: base (array idx) = truncate(scMin/step)+1, pica's baseIdx */
    int ibase[3], its[3];
    /* deposition tile */
    int dbase[3], dts[3];
    or_norm in;          /* pica supercell normalisation (origin scMin - step) */
    double nom[3];       /* origin scMin - 2 step, for the base-m deposition tiles */
} or_sc;

static int or_sc_setup(const or_grid* g, const or_ens* e, int sx, int sy, int sz, int shape, int dep,
                       or_sc* sc)
{
    const int sidx[3] = {sx, sy, sz};
    sc->shape = shape;
    sc->dep = dep;
    for (int d = 0; d < 3; d++) {
        const double scMin = (double)(sidx[d] * e->ncps[d]) * g->step[d];
        sc->in.inv[d] = 1.0 / g->step[d];
        sc->in.no[d] = (scMin - g->step[d]) / g->step[d];        /* FieldInterpolatorYeeGrid.h:201 */
        sc->in.nos[d] = sc->in.no[d] + 1.0 * 0.5;                /* :202 */
        sc->nom[d] = (scMin - 2.0 * g->step[d]) / g->step[d];
        const int base = (int)(scMin / g->step[d]) + 1;          /* :205 */
        sc->ibase[d] = base;
        sc->its[d] = e->ncps[d] + (shape == SHAPE_CIC ? 2 : 3);
        if (shape == SHAPE_CIC && dep == DEP_DIRECT) {
            sc->dbase[d] = base;                                 /* CurrentDepositorYeeGrid.h:185-186 */
            sc->dts[d] = e->ncps[d] + 3;
        } else {
            sc->dbase[d] = base - 1;
            sc->dts[d] = e->ncps[d] + 4;
        }
        if (sc->ibase[d] < 0 || sc->ibase[d] + sc->its[d] > g->s[d]) return -1;
        if (sc->dbase[d] < 0 || sc->dbase[d] + sc->dts[d] > g->s[d]) return -1;
    }
    return 0;
}

static inline int or_in_tile(const int b[3], int span, const int ts[3])
{
    for (int d = 0; d < 3; d++)
        if (b[d] < 0 || b[d] + span > ts[d]) return 0;
    return 1;
}

/* interpolate E, B at pos from the interpolation tile (6 components) */
static int or_sc_interp(const or_sc* sc, const double* tile, const double pos[3], double e[3], double b[3])
{
    const long tn = (long)sc->its[0] * sc->its[1] * sc->its[2];
    if (sc->shape == SHAPE_CIC) {
        int ic[3], is[3], bi[3];
        double cc[3], cs[3], c[3];
        or_index_coeff(&sc->in, pos, ic, cc, is, cs);            /* :257-267 */
        for (int comp = 0; comp < 6; comp++) {
            or_pick(comp, ic, cc, is, cs, bi, c);
            if (!or_in_tile(bi, 2, sc->its)) return -1;
            const double v = or_cic8(tile + comp * tn, sc->its, bi, c);
            if (comp < 3) e[comp] = v; else b[comp - 3] = v;
        }
    } else {
        int nc[3], ns[3], bi[3];
        double wc[3][3], ws[3][3], w[3][3];
        for (int d = 0; d < 3; d++) {
            const double np = pos[d] * sc->in.inv[d];
            nc[d] = or_tsc_node(np - sc->in.no[d], wc[d]);
            ns[d] = or_tsc_node(np - sc->in.nos[d], ws[d]);
        }
        for (int comp = 0; comp < 6; comp++) {
            for (int d = 0; d < 3; d++) {
                const int col = OR_COLLOC[comp][d];
                bi[d] = col ? nc[d] : ns[d];
                for (int k = 0; k < 3; k++) w[d][k] = col ? wc[d][k] : ws[d][k];
            }
            if (!or_in_tile(bi, 3, sc->its)) return -1;
            const double v = or_tsc27(tile + comp * tn, sc->its, bi, w);
            if (comp < 3) e[comp] = v; else b[comp - 3] = v;
        }
    }
    return 0;
}

/* direct deposition of current cur at pos into the J tile */
static int or_sc_deposit_direct(const or_sc* sc, double* jt, const double pos[3], const double cur[3])
{
    const long tn = (long)sc->dts[0] * sc->dts[1] * sc->dts[2];
    if (sc->shape == SHAPE_CIC) {
        int ic[3], is[3], bi[3];
        double cc[3], cs[3], c[3];
        or_index_coeff(&sc->in, pos, ic, cc, is, cs);            /* CurrentDepositorYeeGrid.h:239-249 */
        for (int comp = 0; comp < 3; comp++) {
            or_pick(comp, ic, cc, is, cs, bi, c);
            if (!or_in_tile(bi, 2, sc->dts)) return -1;
            or_cic8_add(jt + comp * tn, sc->dts, bi, c, cur[comp]);
        }
    } else {
        int nc[3], ns[3], bi[3];
        double wc[3][3], ws[3][3], w[3][3];
        for (int d = 0; d < 3; d++) {
            const double np = pos[d] * sc->in.inv[d];
            nc[d] = or_tsc_node(np - sc->nom[d], wc[d]);
            ns[d] = or_tsc_node(np - (sc->nom[d] + 0.5), ws[d]);
        }
        for (int comp = 0; comp < 3; comp++) {
            for (int d = 0; d < 3; d++) {
                const int col = OR_COLLOC[comp][d];
                bi[d] = col ? nc[d] : ns[d];
                for (int k = 0; k < 3; k++) w[d][k] = col ? wc[d][k] : ws[d][k];
            }
            if (!or_in_tile(bi, 3, sc->dts)) return -1;
            or_tsc27_add(jt + comp * tn, sc->dts, bi, w, cur[comp]);
        }
    }
    return 0;
}

/* 1D shape of the old and new positions on a common window of K nodes (rho nodes, i.e. cell
 * centres, relative to the base-m deposition tile).  Returns the window start. */
static inline int or_esk_1d(int shape, double xi0, double xi1, double S0[4], double DS[4], int* K)
{
    double S1[4];
    for (int k = 0; k < 4; k++) { S0[k] = 0.0; S1[k] = 0.0; }
    if (shape == SHAPE_CIC) {
        const int i0 = (int)floor(xi0), i1 = (int)floor(xi1);
        const double f0 = xi0 - (double)i0, f1 = xi1 - (double)i1;
        const int w = i0 < i1 ? i0 : i1;
        *K = 3;
        if (i0 - w > 1 || i1 - w > 1) return -1000000;
        S0[i0 - w] = 1.0 - f0; S0[i0 - w + 1] = f0;
        S1[i1 - w] = 1.0 - f1; S1[i1 - w + 1] = f1;
        for (int k = 0; k < 4; k++) DS[k] = S1[k] - S0[k];
        return w;
    } else {
        double w0[3], w1[3];
        const int a0 = or_tsc_node(xi0, w0), a1 = or_tsc_node(xi1, w1);
        const int w = a0 < a1 ? a0 : a1;
        *K = 4;
        if (a0 - w > 1 || a1 - w > 1) return -1000000;
        for (int k = 0; k < 3; k++) { S0[a0 - w + k] = w0[k]; S1[a1 - w + k] = w1[k]; }
        for (int k = 0; k < 4; k++) DS[k] = S1[k] - S0[k];
        return w;
    }
}

/* Esirkepov charge-conserving deposition of one particle moving pos0 -> pos1 with charge qw
 * into the J tile (values are J * V, i.e. before the 1/V flush normalisation). */
static int or_sc_deposit_esirkepov(const or_sc* sc, double* jt, const double pos0[3],
                                   const double pos1[3], double qw, double dt, const double step[3])
{
    const long tn = (long)sc->dts[0] * sc->dts[1] * sc->dts[2];
    double S0[3][4], DS[3][4];
    int w[3], K = 3;
    for (int d = 0; d < 3; d++) {
        const double off = sc->nom[d] + 0.5;
        const double xi0 = pos0[d] * sc->in.inv[d] - off;
        const double xi1 = pos1[d] * sc->in.inv[d] - off;
        w[d] = or_esk_1d(sc->shape, xi0, xi1, S0[d], DS[d], &K);
        if (w[d] < 0 || w[d] + K > sc->dts[d]) return -1;
    }
    const double third = 1.0 / 3.0;
    for (int comp = 0; comp < 3; comp++) {
        const int d1 = (comp + 1) % 3, d2 = (comp + 2) % 3;
        const double k = -qw * step[comp] / dt;
        double* t = jt + comp * tn;
        for (int p = 0; p < K; p++)
            for (int q = 0; q < K; q++) {
                double cum = 0.0;
                for (int a = 0; a < K - 1; a++) {
                    const double Wv = DS[comp][a] * (S0[d1][p] * S0[d2][q] + 0.5 * DS[d1][p] * S0[d2][q] +
                                                     0.5 * S0[d1][p] * DS[d2][q] + third * DS[d1][p] * DS[d2][q]);
                    cum += Wv;
                    int ix[3];
                    ix[comp] = w[comp] + a + 1;
                    ix[d1] = w[d1] + p;
                    ix[d2] = w[d2] + q;
                    t[IDX3(sc->dts, ix[0], ix[1], ix[2])] += k * cum;
                }
            }
    }
    return 0;
}

static inline int or_wrap_dir(const or_ens* e, int src[3], double x[3])
{
    const double L[3] = {(double)e->ncell[0] * e->step[0], (double)e->ncell[1] * e->step[1],
                         (double)e->ncell[2] * e->step[2]};
    const double mn[3] = {0.0, 0.0, 0.0};
    double inv[3];
    int code = 0;
    for (int d = 0; d < 3; d++) {
        /* supercell geometry as in EnsembleSupercells.h:31-33 */
        const double cellSize = (L[d] - mn[d]) / (double)e->ncell[d];
        inv[d] = 1.0 / ((double)e->ncps[d] * cellSize);
        if (x[d] < 0.0) x[d] += L[d];
        if (x[d] >= L[d]) x[d] -= L[d];
        int dst = (int)floor((x[d] - mn[d]) * inv[d]);
        if (dst < 0) dst = 0;
        if (dst > e->nsc[d] - 1) dst = e->nsc[d] - 1;
        int dd = dst - src[d];
        if (dd > 1) dd -= e->nsc[d];
        if (dd < -1) dd += e->nsc[d];
        if (dd < -1 || dd > 1) return -1;
        code = code * 3 + (dd + 1);
    }
    return code;
}

static int or_process_supercell(or_grid* g, or_ens* e, int sx, int sy, int sz, double dt, int shape,
                                int dep, double coeff, double* work)
{
    or_sc sc;
    if (or_sc_setup(g, e, sx, sy, sz, shape, dep, &sc)) return -1;
    const long itn = (long)sc.its[0] * sc.its[1] * sc.its[2];
    const long dtn = (long)sc.dts[0] * sc.dts[1] * sc.dts[2];
    double* itile = work;
    double* jtile = work + 6 * itn;
    /* gather, FieldInterpolatorYeeGrid.h:211-222 */
    for (int comp = 0; comp < 6; comp++)
        for (int i = 0; i < sc.its[0]; i++)
            for (int j = 0; j < sc.its[1]; j++)
                for (int k = 0; k < sc.its[2]; k++)
                    itile[comp * itn + IDX3(sc.its, i, j, k)] =
                        g->f[EX + comp][IDX3(g->s, sc.ibase[0] + i, sc.ibase[1] + j, sc.ibase[2] + k)];
    memset(jtile, 0, sizeof(double) * 3 * dtn);                 /* CurrentDepositorYeeGrid.h:192-198 */
    const long s = ((long)sx * e->nsc[1] + sy) * e->nsc[2] + sz;
    int src[3] = {sx, sy, sz};
    for (long p = e->off[s]; p < e->off[s + 1]; p++) {
        double x[3] = {e->x[0][p], e->x[1][p], e->x[2][p]};
        double u[3] = {e->u[0][p], e->u[1][p], e->u[2][p]};
        const double x0[3] = {x[0], x[1], x[2]};
        double ef[3], bf[3], v[3];
        if (or_sc_interp(&sc, itile, x, ef, bf)) return -2;
        or_boris(coeff, dt, ef, bf, u, x, v);
        const double qw = e->q * e->w[p];
        if (dep == DEP_DIRECT) {
            double cur[3];
            for (int d = 0; d < 3; d++) cur[d] = v[d] * qw;
            if (or_sc_deposit_direct(&sc, jtile, x, cur)) return -3;
        } else {
            if (or_sc_deposit_esirkepov(&sc, jtile, x0, x, qw, dt, g->step)) return -3;
        }
        const int code = or_wrap_dir(e, src, x);
        if (code < 0) return -4;
        for (int d = 0; d < 3; d++) { e->x[d][p] = x[d]; e->u[d][p] = u[d]; }
        e->cls[p] = (signed char)code;
    }
    /* flush, CurrentDepositorYeeGrid.h:201-213 */
    const double normalization = 1.0 / (g->step[0] * g->step[1] * g->step[2]);
    for (int comp = 0; comp < 3; comp++)
        for (int i = 0; i < sc.dts[0]; i++)
            for (int j = 0; j < sc.dts[1]; j++)
                for (int k = 0; k < sc.dts[2]; k++)
                    g->f[JX + comp][IDX3(g->s, sc.dbase[0] + i, sc.dbase[1] + j, sc.dbase[2] + k)] +=
                        jtile[comp * dtn + IDX3(sc.dts, i, j, k)] * normalization;
    return 0;
}

/* Standalone supercell interpolation (FieldInterpolatorCICSupercell for CIC; the TSC tile
 * for TSC) of n positions with supercell minimum scmin (grid origin must be -2 step). */
int or_interp_sc(const or_grid* g, const double scmin[3], const int ncps[3], int shape, long n,
                 const double* const x[3], double* e[3], double* b[3])
{
    or_sc sc;
    or_ens fake;
    memset(&fake, 0, sizeof(fake));
    int sidx[3];
    for (int d = 0; d < 3; d++) { fake.ncps[d] = ncps[d]; sidx[d] = 0; }
    if (or_sc_setup(g, &fake, sidx[0], sidx[1], sidx[2], shape, DEP_DIRECT, &sc)) return -1;
    for (int d = 0; d < 3; d++) {           /* re-derive with the given minimum position */
        sc.in.no[d] = (scmin[d] - g->step[d]) / g->step[d];
        sc.in.nos[d] = sc.in.no[d] + 1.0 * 0.5;
        sc.nom[d] = (scmin[d] - 2.0 * g->step[d]) / g->step[d];
        sc.ibase[d] = (int)(scmin[d] / g->step[d]) + 1;
        if (sc.ibase[d] < 0 || sc.ibase[d] + sc.its[d] > g->s[d]) return -1;
    }
    const long itn = (long)sc.its[0] * sc.its[1] * sc.its[2];
    double* tile = (double*)malloc(sizeof(double) * 6 * itn);
    for (int comp = 0; comp < 6; comp++)
        for (int i = 0; i < sc.its[0]; i++)
            for (int j = 0; j < sc.its[1]; j++)
                for (int k = 0; k < sc.its[2]; k++)
                    tile[comp * itn + IDX3(sc.its, i, j, k)] =
                        g->f[EX + comp][IDX3(g->s, sc.ibase[0] + i, sc.ibase[1] + j, sc.ibase[2] + k)];
    int rc = 0;
    for (long p = 0; p < n && !rc; p++) {
        const double pos[3] = {x[0][p], x[1][p], x[2][p]};
        double ee[3], bb[3];
        if (or_sc_interp(&sc, tile, pos, ee, bb)) { rc = -2; break; }
        for (int d = 0; d < 3; d++) { e[d][p] = ee[d]; b[d][p] = bb[d]; }
    }
    free(tile);
    return rc;
}

/* Standalone supercell direct deposition (CurrentDepositorCICSupercell for CIC: zero tile,
 * deposit, flush x 1/V) of n currents at positions x. */
int or_deposit_direct_sc(or_grid* g, const double scmin[3], const int ncps[3], int shape, long n,
                         const double* const x[3], const double* const cur[3])
{
    or_sc sc;
    or_ens fake;
    memset(&fake, 0, sizeof(fake));
    for (int d = 0; d < 3; d++) fake.ncps[d] = ncps[d];
    if (or_sc_setup(g, &fake, 0, 0, 0, shape, DEP_DIRECT, &sc)) return -1;
    for (int d = 0; d < 3; d++) {
        sc.in.no[d] = (scmin[d] - g->step[d]) / g->step[d];
        sc.in.nos[d] = sc.in.no[d] + 1.0 * 0.5;
        sc.nom[d] = (scmin[d] - 2.0 * g->step[d]) / g->step[d];
        const int base = (int)(scmin[d] / g->step[d]) + 1;
        sc.dbase[d] = (shape == SHAPE_CIC) ? base : base - 1;
        if (sc.dbase[d] < 0 || sc.dbase[d] + sc.dts[d] > g->s[d]) return -1;
    }
    const long dtn = (long)sc.dts[0] * sc.dts[1] * sc.dts[2];
    double* jt = (double*)calloc(3 * dtn, sizeof(double));
    int rc = 0;
    for (long p = 0; p < n; p++) {
        const double pos[3] = {x[0][p], x[1][p], x[2][p]};
        const double c[3] = {cur[0][p], cur[1][p], cur[2][p]};
        if (or_sc_deposit_direct(&sc, jt, pos, c)) { rc = -2; break; }
    }
    const double normalization = 1.0 / (g->step[0] * g->step[1] * g->step[2]);
    for (int comp = 0; comp < 3 && !rc; comp++)
        for (int i = 0; i < sc.dts[0]; i++)
            for (int j = 0; j < sc.dts[1]; j++)
                for (int k = 0; k < sc.dts[2]; k++)
                    g->f[JX + comp][IDX3(g->s, sc.dbase[0] + i, sc.dbase[1] + j, sc.dbase[2] + k)] +=
                        jt[comp * dtn + IDX3(sc.dts, i, j, k)] * normalization;
    free(jt);
    return rc;
}

/* Standalone Esirkepov deposition of n particles moving x0 -> x1 (charge q*w each) into the
 * base-m tile of the supercell with minimum scmin, flushed x 1/V. */
int or_deposit_esirkepov_sc(or_grid* g, const double scmin[3], const int ncps[3], int shape, long n,
                            const double* const x0[3], const double* const x1[3], const double* qw,
                            double dt)
{
    or_sc sc;
    or_ens fake;
    memset(&fake, 0, sizeof(fake));
    for (int d = 0; d < 3; d++) fake.ncps[d] = ncps[d];
    if (or_sc_setup(g, &fake, 0, 0, 0, shape, DEP_ESIRKEPOV, &sc)) return -1;
    for (int d = 0; d < 3; d++) {
        sc.in.no[d] = (scmin[d] - g->step[d]) / g->step[d];
        sc.in.nos[d] = sc.in.no[d] + 1.0 * 0.5;
        sc.nom[d] = (scmin[d] - 2.0 * g->step[d]) / g->step[d];
        sc.dbase[d] = (int)(scmin[d] / g->step[d]);
        if (sc.dbase[d] < 0 || sc.dbase[d] + sc.dts[d] > g->s[d]) return -1;
    }
    const long dtn = (long)sc.dts[0] * sc.dts[1] * sc.dts[2];
    double* jt = (double*)calloc(3 * dtn, sizeof(double));
    int rc = 0;
    for (long p = 0; p < n; p++) {
        const double a[3] = {x0[0][p], x0[1][p], x0[2][p]};
        const double b[3] = {x1[0][p], x1[1][p], x1[2][p]};
        if (or_sc_deposit_esirkepov(&sc, jt, a, b, qw[p], dt, g->step)) { rc = -2; break; }
    }
    const double normalization = 1.0 / (g->step[0] * g->step[1] * g->step[2]);
    for (int comp = 0; comp < 3 && !rc; comp++)
        for (int i = 0; i < sc.dts[0]; i++)
            for (int j = 0; j < sc.dts[1]; j++)
                for (int k = 0; k < sc.dts[2]; k++)
                    g->f[JX + comp][IDX3(g->s, sc.dbase[0] + i, sc.dbase[1] + j, sc.dbase[2] + k)] +=
                        jt[comp * dtn + IDX3(sc.dts, i, j, k)] * normalization;
    free(jt);
    return rc;
}

/* Particle stage over all supercells in colours: supercells p apart in a dimension, with
 * p = ceil((n + 4) / n) (2 for n >= 4, so 2x2x2 = 8 colours), have disjoint deposition tiles,
 * so one colour runs in parallel and the J sums do not depend on thread scheduling.
 * colour = ((sx % px) * py + sy % py) * pz + sz % pz, colours in increasing order. */
int or_particle_stage(or_grid* g, or_ens* e, double dt, int shape, int dep, int nthreads)
{
    const double coeff = or_boris_coeff(e->q, e->m, dt);
    const long nsct = (long)e->nsc[0] * e->nsc[1] * e->nsc[2];
    const long wn = 6L * (e->ncps[0] + 4) * (e->ncps[1] + 4) * (e->ncps[2] + 4) +
                    3L * (e->ncps[0] + 4) * (e->ncps[1] + 4) * (e->ncps[2] + 4);
    int err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    int pc[3];
    for (int d = 0; d < 3; d++) pc[d] = (e->ncps[d] + 4 + e->ncps[d] - 1) / e->ncps[d];
    const int ncolours = pc[0] * pc[1] * pc[2];
    for (int colour = 0; colour < ncolours; colour++) {
        #pragma omp parallel
        {
            double* work = (double*)malloc(sizeof(double) * wn);
            #pragma omp for schedule(dynamic, 4)
            for (long s = 0; s < nsct; s++) {
                const int sx = (int)(s / ((long)e->nsc[1] * e->nsc[2]));
                const int sy = (int)((s / e->nsc[2]) % e->nsc[1]);
                const int sz = (int)(s % e->nsc[2]);
                if (((sx % pc[0]) * pc[1] + sy % pc[1]) * pc[2] + sz % pc[2] != colour) continue;
                const int r = or_process_supercell(g, e, sx, sy, sz, dt, shape, dep, coeff, work);
                if (r) {
                    #pragma omp atomic write
                    err = r;
                }
            }
            free(work);
        }
        if (err) return err;
    }
    return 0;
}

/* Deterministic re-sort (the order the device kernels produce).  In supercell s with cnt
 * particles of which nmove left (class != 13), nfinal = cnt - nmove:
 *   - stayers keep their slot when it is < nfinal; the holes below nfinal left by movers are
 *     filled, in ascending order, by the stayers found at slots >= nfinal, in ascending order;
 *   - then, for each direction class d = 0..26 (d != 13), the movers of class d of the source
 *     supercell s - delta(d), in ascending slot order, are appended. */
static long or_sc_linear(const or_ens* e, int x, int y, int z)
{
    if (x < 0) x += e->nsc[0];
    if (x >= e->nsc[0]) x -= e->nsc[0];
    if (y < 0) y += e->nsc[1];
    if (y >= e->nsc[1]) y -= e->nsc[1];
    if (z < 0) z += e->nsc[2];
    if (z >= e->nsc[2]) z -= e->nsc[2];
    return ((long)x * e->nsc[1] + y) * e->nsc[2] + z;
}

int or_resort(or_ens* e)
{
    const long nsct = (long)e->nsc[0] * e->nsc[1] * e->nsc[2];
    const long np = e->np > 0 ? e->np : 1;
    long* order = (long*)malloc(sizeof(long) * np);     /* new position -> old index */
    long* noff = (long*)malloc(sizeof(long) * (nsct + 1));
    long* tail = (long*)malloc(sizeof(long) * np);
    if (!order || !noff || !tail) { free(order); free(noff); free(tail); return -1; }
    long k = 0;
    noff[0] = 0;
    for (long s = 0; s < nsct; s++) {
        const int sx = (int)(s / ((long)e->nsc[1] * e->nsc[2]));
        const int sy = (int)((s / e->nsc[2]) % e->nsc[1]);
        const int sz = (int)(s % e->nsc[2]);
        const long b = e->off[s], cnt = e->off[s + 1] - e->off[s];
        long nmove = 0;
        for (long j = 0; j < cnt; j++) nmove += (e->cls[b + j] != 13);
        const long nfinal = cnt - nmove;
        long nt = 0;
        for (long j = nfinal; j < cnt; j++)
            if (e->cls[b + j] == 13) tail[nt++] = b + j;
        long h = 0;
        for (long j = 0; j < nfinal; j++)
            order[k++] = (e->cls[b + j] == 13) ? b + j : tail[h++];
        for (int c = 0; c < 27; c++) {
            if (c == 13) continue;
            const long src = or_sc_linear(e, sx - (c / 9 - 1), sy - ((c / 3) % 3 - 1), sz - (c % 3 - 1));
            for (long j = e->off[src]; j < e->off[src + 1]; j++)
                if (e->cls[j] == c) order[k++] = j;
        }
        noff[s + 1] = k;
    }
    if (k != e->np) { free(order); free(noff); free(tail); return -2; }
    double* buf = (double*)malloc(sizeof(double) * np);
    double* arrs[7] = {e->x[0], e->x[1], e->x[2], e->u[0], e->u[1], e->u[2], e->w};
    for (int a = 0; a < 7; a++) {
        for (long p = 0; p < e->np; p++) buf[p] = arrs[a][order[p]];
        memcpy(arrs[a], buf, sizeof(double) * e->np);
    }
    for (long p = 0; p < e->np; p++) e->cls[p] = 13;
    memcpy(e->off, noff, sizeof(long) * (nsct + 1));
    free(buf); free(order); free(noff); free(tail);
    return 0;
}

/* Initial bucketing of unsorted particles: stable by supercell index (input order kept),
 * as repeated EnsembleSupercells::add calls do (EnsembleSupercells.h:41-45). */
int or_bucket(or_ens* e)
{
    const long nsct = (long)e->nsc[0] * e->nsc[1] * e->nsc[2];
    const long np = e->np > 0 ? e->np : 1;
    const double mn[3] = {0.0, 0.0, 0.0};
    const double mx[3] = {(double)e->ncell[0] * e->step[0], (double)e->ncell[1] * e->step[1],
                          (double)e->ncell[2] * e->step[2]};
    int* idx = (int*)malloc(sizeof(int) * 3 * np);
    long* keys = (long*)malloc(sizeof(long) * np);
    long* cnt = (long*)calloc(nsct + 1, sizeof(long));
    long* pos = (long*)malloc(sizeof(long) * np);
    double* buf = (double*)malloc(sizeof(double) * np);
    int rc = 0;
    if (!idx || !keys || !cnt || !pos || !buf) { rc = -2; goto done; }
    or_supercell_index(mn, mx, e->ncell, e->ncps, e->np, (const double* const*)e->x, idx);
    for (long p = 0; p < e->np; p++) {
        for (int d = 0; d < 3; d++)
            if (idx[3 * p + d] < 0 || idx[3 * p + d] >= e->nsc[d]) { rc = -1; goto done; }
        keys[p] = ((long)idx[3 * p] * e->nsc[1] + idx[3 * p + 1]) * e->nsc[2] + idx[3 * p + 2];
        cnt[keys[p] + 1]++;
    }
    for (long s = 0; s < nsct; s++) cnt[s + 1] += cnt[s];
    {
        long* next = (long*)malloc(sizeof(long) * (nsct + 1));
        memcpy(next, cnt, sizeof(long) * (nsct + 1));
        for (long p = 0; p < e->np; p++) pos[p] = next[keys[p]]++;
        free(next);
    }
    {
        double* arrs[7] = {e->x[0], e->x[1], e->x[2], e->u[0], e->u[1], e->u[2], e->w};
        for (int a = 0; a < 7; a++) {
            for (long p = 0; p < e->np; p++) buf[pos[p]] = arrs[a][p];
            memcpy(arrs[a], buf, sizeof(double) * e->np);
        }
    }
    for (long s = 0; s <= nsct; s++) e->off[s] = cnt[s];
    for (long p = 0; p < e->np; p++) e->cls[p] = 13;
done:
    free(idx); free(keys); free(cnt); free(pos); free(buf);
    return rc;
}

/* Field stage of the step (periodic): fold J, then B/2, E, B/2 with ghost fills. */
void or_field_stage(or_grid* g, int G, double dt)
{
    or_ghost_fold(g->s, G, g->f[JX]);
    or_ghost_fold(g->s, G, g->f[JY]);
    or_ghost_fold(g->s, G, g->f[JZ]);
    for (int c = JX; c <= JZ; c++) or_ghost_fill(g->s, G, g->f[c]);
    const double half = dt * 0.5;
    or_yee_update_b(g, half);
    for (int c = BX; c <= BZ; c++) or_ghost_fill(g->s, G, g->f[c]);
    or_yee_update_e(g, dt);
    for (int c = EX; c <= EZ; c++) or_ghost_fill(g->s, G, g->f[c]);
    or_yee_update_b(g, half);
    for (int c = BX; c <= BZ; c++) or_ghost_fill(g->s, G, g->f[c]);
}

/* One full periodic PIC step: J = 0, particle stage, field stage, re-sort. */
int or_step(or_grid* g, or_ens* e, double dt, int shape, int dep, int nthreads)
{
    const long n = (long)g->s[0] * g->s[1] * g->s[2];
    for (int c = JX; c <= JZ; c++) memset(g->f[c], 0, sizeof(double) * n);
    const int r = or_particle_stage(g, e, dt, shape, dep, nthreads);
    if (r) return r;
    or_field_stage(g, e->ghost, dt);
    return or_resort(e);
}

/* rho (charge density at cell centres) with the given shape, folded and filled periodically.
 * rho is a grid-sized array (incl. ghosts).  Sum order: particles in ensemble order. */
int or_deposit_rho(const or_grid* g, const or_ens* e, int shape, double* rho)
{
    const long n = (long)g->s[0] * g->s[1] * g->s[2];
    memset(rho, 0, sizeof(double) * n);
    const double normalization = 1.0 / (g->step[0] * g->step[1] * g->step[2]);
    for (long p = 0; p < e->np; p++) {
        const double qw = e->q * e->w[p];
        int base[3];
        double w[3][3];
        int K = shape == SHAPE_CIC ? 2 : 3;
        for (int d = 0; d < 3; d++) {
            /* cell-centre node a sits at origin + (a + 0.5) step */
            const double xi = e->x[d][p] * (1.0 / g->step[d]) - (g->origin[d] / g->step[d] + 0.5);
            if (shape == SHAPE_CIC) {
                const int i0 = (int)floor(xi);
                const double f = xi - (double)i0;
                base[d] = i0; w[d][0] = 1.0 - f; w[d][1] = f; w[d][2] = 0.0;
            } else {
                base[d] = or_tsc_node(xi, w[d]);
            }
            if (base[d] < 0 || base[d] + K > g->s[d]) return -1;
        }
        for (int a = 0; a < K; a++)
            for (int b = 0; b < K; b++)
                for (int c = 0; c < K; c++)
                    rho[IDX3(g->s, base[0] + a, base[1] + b, base[2] + c)] +=
                        w[0][a] * w[1][b] * w[2][c] * qw * normalization;
    }
    or_ghost_fold(g->s, e->ghost, rho);
    or_ghost_fill(g->s, e->ghost, rho);
    return 0;
}

/* Energies: field = sum over physical cells (E^2 + B^2) V / (8 pi); kinetic = sum w m c^2 (gamma-1)
 * with gamma - 1 = u^2 / (1 + gamma). */
void or_energy(const or_grid* g, const or_ens* e, double out[2])
{
    const int G = e->ghost;
    double fe = 0.0;
    for (int i = G; i < g->s[0] - G; i++)
        for (int j = G; j < g->s[1] - G; j++)
            for (int k = G; k < g->s[2] - G; k++) {
                const long c = IDX3(g->s, i, j, k);
                double s = 0.0;
                for (int comp = EX; comp <= BZ; comp++) s += g->f[comp][c] * g->f[comp][c];
                fe += s;
            }
    const double V = g->step[0] * g->step[1] * g->step[2];
    out[0] = fe * V / (8.0 * OR_PI);
    double ke = 0.0;
    for (long p = 0; p < e->np; p++) {
        const double u2 = e->u[0][p] * e->u[0][p] + e->u[1][p] * e->u[1][p] + e->u[2][p] * e->u[2][p];
        const double gam = sqrt(1.0 + u2);
        ke += e->w[p] * (u2 / (1.0 + gam));
    }
    out[1] = ke * e->m * OR_C * OR_C;
}

/* ------------------------------------------------------------------------------------ */
/* Seeded thermal loader: the CPU twin of the library's pk_ens_load_thermal.  pica has no  */
/* loader; this restates the library's counter-based one (SplitMix64 finaliser keyed by   */
/* seed and particle id, Box-Muller from ln and sincos written with IEEE +, -, *, / only), */
/* so that CPU and GPU runs start from bit-identical particles.                            */
/* ------------------------------------------------------------------------------------ */
static inline uint64_t or_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline double or_urand(uint64_t key, int j)
{
    return (double)(or_mix64(key + (uint64_t)j * 0xD1B54A32D192ED03ull) >> 11) * 0x1.0p-53;
}

/* ln(x), 0 < x <= 1: x = m 2^e, m in [sqrt(1/2), sqrt(2)], ln m = 2 atanh((m-1)/(m+1)) */
static double or_det_log(double x)
{
    uint64_t bits;
    memcpy(&bits, &x, 8);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    const uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    memcpy(&m, &mb, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    static const double inv_odd[12] = {1.0, 1.0 / 3.0, 1.0 / 5.0, 1.0 / 7.0, 1.0 / 9.0, 1.0 / 11.0, 1.0 / 13.0,
                                       1.0 / 15.0, 1.0 / 17.0, 1.0 / 19.0, 1.0 / 21.0, 1.0 / 23.0};
    double p = inv_odd[11];
    for (int k = 10; k >= 0; k--) p = inv_odd[k] + s2 * p;
    const double lnm = 2.0 * s * p;
    const double de = (double)e;
    return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + lnm);
}

/* sin and cos of 2 pi U, 0 <= U < 1, by octant reduction and Taylor series on [0, pi/4] */
static void or_det_sincos_2pi(double U, double* sn, double* cs)
{
    const double t = U * 8.0;
    const int q = (int)t;
    const double f = t - (double)q;
    const double r = ((q & 1) ? 1.0 - f : f) * 0.78539816339744830962;
    const double r2 = r * r;
    /* 1/(2k+1)! for k = 1..9 and 1/(2k)! for k = 1..10 */
    static const double is[9] = {1.0 / 6.0, 1.0 / 120.0, 1.0 / 5040.0, 1.0 / 362880.0, 1.0 / 39916800.0,
                                 1.0 / 6227020800.0, 1.0 / 1307674368000.0, 1.0 / 355687428096000.0,
                                 1.0 / 121645100408832000.0};
    static const double ic[10] = {0.5, 1.0 / 24.0, 1.0 / 720.0, 1.0 / 40320.0, 1.0 / 3628800.0, 1.0 / 479001600.0,
                                  1.0 / 87178291200.0, 1.0 / 20922789888000.0, 1.0 / 6402373705728000.0,
                                  1.0 / 2432902008176640000.0};
    double ps = is[8];
    for (int k = 7; k >= 0; k--) ps = is[k] - r2 * ps;
    const double sr = r - r * r2 * ps;
    double pc = ic[9];
    for (int k = 8; k >= 0; k--) pc = ic[k] - r2 * pc;
    const double cr = 1.0 - r2 * pc;
    switch (q & 7) {
    case 0: *sn = sr; *cs = cr; break;
    case 1: *sn = cr; *cs = sr; break;
    case 2: *sn = cr; *cs = -sr; break;
    case 3: *sn = sr; *cs = -cr; break;
    case 4: *sn = -sr; *cs = -cr; break;
    case 5: *sn = -cr; *cs = -sr; break;
    case 6: *sn = -cr; *cs = sr; break;
    default: *sn = -sr; *cs = cr; break;
    }
}

/* Particles of load_thermal in the order the library stores them: supercell by supercell
 * (z fastest), slot l of a supercell holding particle k of local cell lc with
 *   unpaired: lc = l mod cells, k = l / cells;
 *   paired:   pairs (2q, 2q+1) of a cell in adjacent slots, by (q, cell), then particle ppc-1
 *             of every cell when ppc is odd.
 * Particle id = cell id * ppc + k; key = mix(seed ^ mix(id)); U_j = urand(key, j):
 * position (cell + U_0..2) h, u = uth (r1 cos a1, r1 sin a1, r2 cos a2) + (+-drift, 0, 0) with
 * r = sqrt(-2 ln(1 - U_3 or 5)), a = 2 pi U_4 or 6, +drift for even ids.
 * Output: absolute positions (x), u, and the cell of each particle (cell, z fastest). */
void or_load_thermal(const int ncell[3], const int ncps[3], const double step[3], int ppc, double uth,
                     double drift, double weight, uint64_t seed, int paired, double* x[3], double* u[3],
                     double* w, int* cell)
{
    int nsc[3];
    for (int d = 0; d < 3; d++) nsc[d] = ncell[d] / ncps[d];
    const int ncsc = ncps[0] * ncps[1] * ncps[2];
    const int per_sc = ncsc * ppc;
    const long nsct = (long)nsc[0] * nsc[1] * nsc[2];
    for (long s = 0; s < nsct; s++) {
        const int sidx[3] = {(int)(s / ((long)nsc[1] * nsc[2])), (int)((s / nsc[2]) % nsc[1]), (int)(s % nsc[2])};
        for (int l = 0; l < per_sc; l++) {
            int lc = l % ncsc, k = l / ncsc;
            if (paired) {
                const int npair = (ppc & ~1) * ncsc;
                if (l < npair) {
                    lc = (l >> 1) % ncsc;
                    k = 2 * ((l >> 1) / ncsc) + (l & 1);
                } else {
                    lc = l - npair;
                    k = ppc - 1;
                }
            }
            const int a = lc / (ncps[1] * ncps[2]), b = (lc / ncps[2]) % ncps[1], c = lc % ncps[2];
            const int gc[3] = {sidx[0] * ncps[0] + a, sidx[1] * ncps[1] + b, sidx[2] * ncps[2] + c};
            const uint64_t id = ((((uint64_t)gc[0] * ncell[1] + gc[1]) * ncell[2] + gc[2]) * (uint64_t)ppc) + k;
            const uint64_t key = or_mix64(seed ^ or_mix64(id));
            const long p = s * per_sc + l;
            for (int d = 0; d < 3; d++) x[d][p] = ((double)gc[d] + or_urand(key, d)) * step[d];
            const double r1 = sqrt(-2.0 * or_det_log(1.0 - or_urand(key, 3)));
            const double r2 = sqrt(-2.0 * or_det_log(1.0 - or_urand(key, 5)));
            double s1, c1, s2, c2;
            or_det_sincos_2pi(or_urand(key, 4), &s1, &c1);
            or_det_sincos_2pi(or_urand(key, 6), &s2, &c2);
            u[0][p] = uth * r1 * c1 + ((id & 1ull) ? -drift : drift);
            u[1][p] = uth * r1 * s1;
            u[2][p] = uth * r2 * c2;
            w[p] = weight;
            cell[p] = (gc[0] * ncell[1] + gc[1]) * ncell[2] + gc[2];
        }
    }
}

/* ln and sincos of the loader, exposed for their tests */
double or_det_log_x(double x) { return or_det_log(x); }
void or_det_sincos_2pi_x(double U, double* sn, double* cs) { or_det_sincos_2pi(U, sn, cs); }

int or_version(void) { return 1; }
