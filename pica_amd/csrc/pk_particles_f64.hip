// pk_particles_f64.hip -- fused interpolation + Boris push + current deposition, FP64,
// absolute positions, one workgroup per supercell (gfx950).
//
// Per workgroup (supercell s, XCD-aware numbering):
//   1. gather the E/B stencil of the supercell into LDS (FieldInterpolatorCICSupercell ctor,
//      FieldInterpolatorYeeGrid.h:197-223) and zero an LDS J tile
//      (CurrentDepositorCICSupercell ctor, CurrentDepositorYeeGrid.h:178-199);
//   2. stream the supercell's SoA particles, each wavefront its own 64-particle groups with no
//      workgroup barrier, coalesced: interpolate (:225-242, 257-281), push
//      (BorisPusher.h:32-47), deposit into the LDS tile with LDS atomics
//      (CurrentDepositorYeeGrid.h:215-262, or Esirkepov), write x, u back in place;
//   3. particles that leave are copied to the outbox in slot order (with their direction
//      class), and the stayers found past the final count fill the holes movers left below it
//      (ascending order on both sides);
//   4. flush the J tile x 1/V with one pass of global FP64 atomics (the destructor, :201-213).
// This file is compiled with -ffp-contract=off: per particle, the arithmetic is pica's, in
// pica's order, so positions and momenta agree with pica to the last bit given the same
// fields.  The only FP difference left is the order of J sums (atomics).
#include "pk_internal.h"

namespace {

constexpr int WAVES = PK_BLOCK / 64;

struct ScGeom {
    double inv[3], no[3], nos[3], nom[3];
    int ibase[3], its[3], dbase[3], dts[3];
};

template <int SHAPE, int DEP>
__device__ __forceinline__ void sc_geom(const GridView<double>& g, const EnsGeom& eg, const int sidx[3], ScGeom& q)
{
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double step = g.step[d];
        const double scMin = (double)(sidx[d] * eg.ncps[d]) * step;
        q.inv[d] = 1.0 / step;
        q.no[d] = (scMin - step) / step;                 // FieldInterpolatorYeeGrid.h:201
        q.nos[d] = q.no[d] + 0.5;                        // :202
        q.nom[d] = (scMin - 2.0 * step) / step;
        const int base = (int)(scMin / step) + 1;       // :205
        q.ibase[d] = base;
        q.its[d] = eg.ncps[d] + (SHAPE == PK_CIC ? 2 : 3);
        if (SHAPE == PK_CIC && DEP == PK_DIRECT) {
            q.dbase[d] = base;                           // CurrentDepositorYeeGrid.h:185-186
            q.dts[d] = eg.ncps[d] + 3;
        } else {
            q.dbase[d] = base - 1;
            q.dts[d] = eg.ncps[d] + 4;
        }
    }
}

__device__ __forceinline__ int tidx(const int ts[3], int i, int j, int k) { return (i * ts[1] + j) * ts[2] + k; }

// E/J_x collocated in x; E/J_y in y; E/J_z in z; B_x staggered only in x, etc.
// (FieldInterpolatorYeeGrid.h:155-166)
__device__ __forceinline__ bool colloc(int comp, int d)
{
    // bit d of the mask of component comp
    const unsigned m[6] = {1u, 2u, 4u, 6u, 5u, 3u};
    return (m[comp] >> d) & 1u;
}

// pica's 8-point CIC sum, term order of FieldInterpolatorYeeGrid.h:273-280
__device__ __forceinline__ double cic8(const double* t, const int ts[3], int i, int j, int k, const double c[3])
{
    return t[tidx(ts, i, j, k)] * (1.0 - c[0]) * (1.0 - c[1]) * (1.0 - c[2]) +
           t[tidx(ts, i, j, k + 1)] * (1.0 - c[0]) * (1.0 - c[1]) * c[2] +
           t[tidx(ts, i, j + 1, k)] * (1.0 - c[0]) * c[1] * (1.0 - c[2]) +
           t[tidx(ts, i, j + 1, k + 1)] * (1.0 - c[0]) * c[1] * c[2] +
           t[tidx(ts, i + 1, j, k)] * c[0] * (1.0 - c[1]) * (1.0 - c[2]) +
           t[tidx(ts, i + 1, j, k + 1)] * c[0] * (1.0 - c[1]) * c[2] +
           t[tidx(ts, i + 1, j + 1, k)] * c[0] * c[1] * (1.0 - c[2]) +
           t[tidx(ts, i + 1, j + 1, k + 1)] * c[0] * c[1] * c[2];
}

// pica's 8-point CIC scatter, CurrentDepositorYeeGrid.h:254-261, into LDS with atomics
__device__ __forceinline__ void cic8_add(double* t, const int ts[3], int i, int j, int k, const double c[3], double v)
{
    atomicAdd(&t[tidx(ts, i, j, k)], (1.0 - c[0]) * (1.0 - c[1]) * (1.0 - c[2]) * v);
    atomicAdd(&t[tidx(ts, i, j, k + 1)], (1.0 - c[0]) * (1.0 - c[1]) * c[2] * v);
    atomicAdd(&t[tidx(ts, i, j + 1, k)], (1.0 - c[0]) * c[1] * (1.0 - c[2]) * v);
    atomicAdd(&t[tidx(ts, i, j + 1, k + 1)], (1.0 - c[0]) * c[1] * c[2] * v);
    atomicAdd(&t[tidx(ts, i + 1, j, k)], c[0] * (1.0 - c[1]) * (1.0 - c[2]) * v);
    atomicAdd(&t[tidx(ts, i + 1, j, k + 1)], c[0] * (1.0 - c[1]) * c[2] * v);
    atomicAdd(&t[tidx(ts, i + 1, j + 1, k)], c[0] * c[1] * (1.0 - c[2]) * v);
    atomicAdd(&t[tidx(ts, i + 1, j + 1, k + 1)], c[0] * c[1] * c[2] * v);
}

__device__ __forceinline__ void tsc_w(double d, double w[3])
{
    const double a = 0.5 - d, b = 0.5 + d;
    w[0] = 0.5 * a * a;
    w[1] = 0.75 - d * d;
    w[2] = 0.5 * b * b;
}

// nearest node i0 of xi; weights on i0-1..i0+1; returns i0-1
__device__ __forceinline__ int tsc_node(double xi, double w[3])
{
    const int i0 = (int)floor(xi + 0.5);
    tsc_w(xi - (double)i0, w);
    return i0 - 1;
}

__device__ __forceinline__ double tsc27(const double* t, const int ts[3], const int b[3], const double w[3][3])
{
    double sa = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double sb = 0.0;
#pragma unroll
        for (int bb = 0; bb < 3; bb++) {
            double sc = 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++) sc += w[2][c] * t[tidx(ts, b[0] + a, b[1] + bb, b[2] + c)];
            sb += w[1][bb] * sc;
        }
        sa += w[0][a] * sb;
    }
    return sa;
}

__device__ __forceinline__ void tsc27_add(double* t, const int ts[3], const int b[3], const double w[3][3], double v)
{
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int bb = 0; bb < 3; bb++)
#pragma unroll
            for (int c = 0; c < 3; c++)
                atomicAdd(&t[tidx(ts, b[0] + a, b[1] + bb, b[2] + c)], w[0][a] * w[1][bb] * w[2][c] * v);
}

__device__ __forceinline__ bool in_tile(const int b[3], int span, const int ts[3])
{
    return b[0] >= 0 && b[0] + span <= ts[0] && b[1] >= 0 && b[1] + span <= ts[1] && b[2] >= 0 &&
           b[2] + span <= ts[2];
}

template <int SHAPE>
__device__ __forceinline__ bool interp(const ScGeom& q, const double* tile, const double x[3], double e[3], double b[3])
{
    const int tn = q.its[0] * q.its[1] * q.its[2];
    if (SHAPE == PK_CIC) {
        int ic[3], is[3];
        double cc[3], cs[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {               // getIndexCoeff, :257-267
            const double np = x[d] * q.inv[d];
            double internal = np - q.no[d];
            ic[d] = (int)internal;
            cc[d] = internal - (double)ic[d];
            internal = np - q.nos[d];
            is[d] = (int)internal;
            cs[d] = internal - (double)is[d];
        }
        bool ok = true;
#pragma unroll
        for (int comp = 0; comp < 6; comp++) {
            int bi[3];
            double c[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                bi[d] = colloc(comp, d) ? ic[d] : is[d];
                c[d] = colloc(comp, d) ? cc[d] : cs[d];
            }
            ok = ok && in_tile(bi, 2, q.its);
            const double v = ok ? cic8(tile + comp * tn, q.its, bi[0], bi[1], bi[2], c) : 0.0;
            if (comp < 3) e[comp] = v; else b[comp - 3] = v;
        }
        return ok;
    } else {
        int nc[3], ns[3];
        double wc[3][3], ws[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double np = x[d] * q.inv[d];
            nc[d] = tsc_node(np - q.no[d], wc[d]);
            ns[d] = tsc_node(np - q.nos[d], ws[d]);
        }
        bool ok = true;
#pragma unroll
        for (int comp = 0; comp < 6; comp++) {
            int bi[3];
            double w[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const bool col = colloc(comp, d);
                bi[d] = col ? nc[d] : ns[d];
#pragma unroll
                for (int k = 0; k < 3; k++) w[d][k] = col ? wc[d][k] : ws[d][k];
            }
            ok = ok && in_tile(bi, 3, q.its);
            const double v = ok ? tsc27(tile + comp * tn, q.its, bi, w) : 0.0;
            if (comp < 3) e[comp] = v; else b[comp - 3] = v;
        }
        return ok;
    }
}

// BorisPusher::push (BorisPusher.h:32-47) with ParticleRef::setP / getVelocity
// (ParticleArray.h:209,257-263); returns v
__device__ __forceinline__ void boris(double coeff, double dt, const double e[3], const double b[3], double u[3],
                                      double x[3], double v[3])
{
    double eM[3], um[3], t[3], up[3], s[3], cr[3];
#pragma unroll
    for (int d = 0; d < 3; d++) eM[d] = e[d] * coeff;
#pragma unroll
    for (int d = 0; d < 3; d++) um[d] = u[d] + eM[d];
    const double um2 = um[0] * um[0] + um[1] * um[1] + um[2] * um[2];
    const double tk = coeff / sqrt(1.0 + um2);
#pragma unroll
    for (int d = 0; d < 3; d++) t[d] = b[d] * tk;
    cr[0] = um[1] * t[2] - um[2] * t[1];
    cr[1] = um[2] * t[0] - um[0] * t[2];
    cr[2] = um[0] * t[1] - um[1] * t[0];
#pragma unroll
    for (int d = 0; d < 3; d++) up[d] = um[d] + cr[d];
    const double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    const double sk = 2.0 / (1.0 + t2);
#pragma unroll
    for (int d = 0; d < 3; d++) s[d] = t[d] * sk;
    cr[0] = up[1] * s[2] - up[2] * s[1];
    cr[1] = up[2] * s[0] - up[0] * s[2];
    cr[2] = up[0] * s[1] - up[1] * s[0];
#pragma unroll
    for (int d = 0; d < 3; d++) u[d] = um[d] + cr[d] + eM[d];
    const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const double invGamma = 1.0 / sqrt(1.0 + u2);
#pragma unroll
    for (int d = 0; d < 3; d++) v[d] = u[d] * PK_C * invGamma;
#pragma unroll
    for (int d = 0; d < 3; d++) x[d] = x[d] + v[d] * dt;
}

template <int SHAPE>
__device__ __forceinline__ bool deposit_direct(const ScGeom& q, double* jt, const double x[3], const double cur[3])
{
    const int tn = q.dts[0] * q.dts[1] * q.dts[2];
    if (SHAPE == PK_CIC) {
        int ic[3], is[3];
        double cc[3], cs[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {               // CurrentDepositorYeeGrid.h:239-249
            const double np = x[d] * q.inv[d];
            double internal = np - q.no[d];
            ic[d] = (int)internal;
            cc[d] = internal - (double)ic[d];
            internal = np - q.nos[d];
            is[d] = (int)internal;
            cs[d] = internal - (double)is[d];
        }
        bool ok = true;
#pragma unroll
        for (int comp = 0; comp < 3; comp++) {
            int bi[3];
            double c[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                bi[d] = colloc(comp, d) ? ic[d] : is[d];
                c[d] = colloc(comp, d) ? cc[d] : cs[d];
            }
            ok = ok && in_tile(bi, 2, q.dts);
        }
        if (!ok) return false;
#pragma unroll
        for (int comp = 0; comp < 3; comp++) {
            int bi[3];
            double c[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                bi[d] = colloc(comp, d) ? ic[d] : is[d];
                c[d] = colloc(comp, d) ? cc[d] : cs[d];
            }
            cic8_add(jt + comp * tn, q.dts, bi[0], bi[1], bi[2], c, cur[comp]);
        }
        return true;
    } else {
        int nc[3], ns[3];
        double wc[3][3], ws[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const double np = x[d] * q.inv[d];
            nc[d] = tsc_node(np - q.nom[d], wc[d]);
            ns[d] = tsc_node(np - (q.nom[d] + 0.5), ws[d]);
        }
        bool ok = true;
#pragma unroll
        for (int comp = 0; comp < 3; comp++) {
            int bi[3];
#pragma unroll
            for (int d = 0; d < 3; d++) bi[d] = colloc(comp, d) ? nc[d] : ns[d];
            ok = ok && in_tile(bi, 3, q.dts);
        }
        if (!ok) return false;
#pragma unroll
        for (int comp = 0; comp < 3; comp++) {
            int bi[3];
            double w[3][3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const bool col = colloc(comp, d);
                bi[d] = col ? nc[d] : ns[d];
#pragma unroll
                for (int k = 0; k < 3; k++) w[d][k] = col ? wc[d][k] : ws[d][k];
            }
            tsc27_add(jt + comp * tn, q.dts, bi, w, cur[comp]);
        }
        return true;
    }
}

// 1D old/new shapes on a common window of K nodes (cell centres, base-m tile); -1 on failure
template <int SHAPE>
__device__ __forceinline__ int esk_1d(double xi0, double xi1, double S0[4], double DS[4])
{
    double S1[4];
#pragma unroll
    for (int k = 0; k < 4; k++) { S0[k] = 0.0; S1[k] = 0.0; }
    int w;
    if (SHAPE == PK_CIC) {
        const int i0 = (int)floor(xi0), i1 = (int)floor(xi1);
        const double f0 = xi0 - (double)i0, f1 = xi1 - (double)i1;
        w = i0 < i1 ? i0 : i1;
        if (i0 - w > 1 || i1 - w > 1) return -1;
        const int a = i0 - w, b = i1 - w;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            S0[k] = (k == a) ? 1.0 - f0 : ((k == a + 1) ? f0 : 0.0);
            S1[k] = (k == b) ? 1.0 - f1 : ((k == b + 1) ? f1 : 0.0);
        }
    } else {
        double w0[3], w1[3];
        const int a0 = tsc_node(xi0, w0), a1 = tsc_node(xi1, w1);
        w = a0 < a1 ? a0 : a1;
        if (a0 - w > 1 || a1 - w > 1) return -1;
        const int a = a0 - w, b = a1 - w;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            S0[k] = (k == a) ? w0[0] : (k == a + 1) ? w0[1] : (k == a + 2) ? w0[2] : 0.0;
            S1[k] = (k == b) ? w1[0] : (k == b + 1) ? w1[1] : (k == b + 2) ? w1[2] : 0.0;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) DS[k] = S1[k] - S0[k];
    return w;
}

// Esirkepov (CPC 135 (2001) 144) deposition of a particle moving x0 -> x1 with charge qw
template <int SHAPE>
__device__ __forceinline__ bool deposit_esirkepov(const ScGeom& q, double* jt, const double x0[3], const double x1[3],
                                                  double qw, double dt, const double step[3])
{
    constexpr int K = SHAPE == PK_CIC ? 3 : 4;
    const int tn = q.dts[0] * q.dts[1] * q.dts[2];
    double S0[3][4], DS[3][4];
    int w[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double off = q.nom[d] + 0.5;
        const double xi0 = x0[d] * q.inv[d] - off;
        const double xi1 = x1[d] * q.inv[d] - off;
        w[d] = esk_1d<SHAPE>(xi0, xi1, S0[d], DS[d]);
        if (w[d] < 0 || w[d] + K > q.dts[d]) return false;
    }
    const double third = 1.0 / 3.0;
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
        const int d1 = (comp + 1) % 3, d2 = (comp + 2) % 3;
        const double kq = -qw * step[comp] / dt;
        double* t = jt + comp * tn;
#pragma unroll
        for (int p = 0; p < K; p++)
#pragma unroll
            for (int r = 0; r < K; r++) {
                double cum = 0.0;
#pragma unroll
                for (int a = 0; a < K - 1; a++) {
                    const double Wv = DS[comp][a] * (S0[d1][p] * S0[d2][r] + 0.5 * DS[d1][p] * S0[d2][r] +
                                                     0.5 * S0[d1][p] * DS[d2][r] + third * DS[d1][p] * DS[d2][r]);
                    cum += Wv;
                    int ix[3];
                    ix[comp] = w[comp] + a + 1;
                    ix[d1] = w[d1] + p;
                    ix[d2] = w[d2] + r;
                    const double val = kq * cum;
                    if (val != 0.0) atomicAdd(&t[tidx(q.dts, ix[0], ix[1], ix[2])], val);
                }
            }
    }
    return true;
}

struct StepParams {
    double dt, coeff, q;
    double L[3], scinv[3];
    int mask_words;
};

// all arguments of one ensemble's particle step (one kernel argument, or one entry of a batch)
struct F64Args {
    GridView<double> g;
    EnsGeom eg;
    PartF64 pm, po;
    uint8_t* odir;
    const int64_t* off;
    const int32_t* cnt;
    int32_t *nstay, *nmove, *mvc, *flags;
    StepParams sp;
};


template <int SHAPE, int DEP>
__global__ void __launch_bounds__(PK_BLOCK) k_particles(GridView<double> g, EnsGeom eg, PartF64 pm, PartF64 po,
                                                        uint8_t* __restrict__ odir, const int64_t* __restrict__ off,
                                                        const int32_t* __restrict__ cnt, int32_t* __restrict__ nstay,
                                                        int32_t* __restrict__ nmove, int32_t* __restrict__ mvc,
                                                        int32_t* __restrict__ flags, StepParams sp)
{
    const long s = pk_xcd_supercell(blockIdx.x, eg.xcd_chunk);
    if (s >= eg.nsct) return;
#include "pk_particles_f64_body.inc"
}

// many ensembles: a list of (ensemble, supercell) items; gridDim.x < nitems makes it persistent
template <int SHAPE, int DEP>
__global__ void __launch_bounds__(PK_BLOCK) k_particles_batch(const F64Args* __restrict__ jobs,
                                                              const int2* __restrict__ items, long nitems,
                                                              const int* __restrict__ halt)
{
    if (halt[PK_HALT_STOP]) return;
    for (long w = blockIdx.x; w < nitems; w += gridDim.x) {
        __syncthreads();
        const int2 it = items[w];
        const long s = it.y;
        const F64Args& A = jobs[it.x];      // arguments read in place (a copy would live in scratch)
        const GridView<double>& g = A.g;
        const EnsGeom& eg = A.eg;
        const PartF64& pm = A.pm;
        const PartF64& po = A.po;
        uint8_t* __restrict__ odir = A.odir;
        const int64_t* __restrict__ off = A.off;
        const int32_t* __restrict__ cnt = A.cnt;
        int32_t* __restrict__ nstay = A.nstay;
        int32_t* __restrict__ nmove = A.nmove;
        int32_t* __restrict__ mvc = A.mvc;
        int32_t* __restrict__ flags = A.flags;
        const StepParams& sp = A.sp;
#include "pk_particles_f64_body.inc"
    }
}

StepParams step_params(const pk_ens* e, double dt)
{
    StepParams sp;
    sp.dt = dt;
    const double q = e->ctx->charge[e->type], m = e->ctx->mass[e->type];
    sp.coeff = q * dt / (2.0 * m * PK_C);        // BorisPusher.h:29
    sp.q = q;
    for (int d = 0; d < 3; d++) {
        sp.L[d] = (double)e->ncell[d] * e->step[d];
        const double cellSize = (sp.L[d] - 0.0) / (double)e->ncell[d];      // EnsembleSupercells.h:31-33
        sp.scinv[d] = 1.0 / ((double)e->ncps[d] * cellSize);
    }
    sp.mask_words = (e->maxcap + PK_BLOCK + 63) / 64 + 1;
    return sp;
}

// dynamic LDS of k_particles: field and J tiles, mover mask, counters, mask prefix
size_t f64_lds(const pk_ens* e, int shape, int dep, int mask_words)
{
    int its[3], dts[3];
    for (int d = 0; d < 3; d++) {
        its[d] = e->ncps[d] + (shape == PK_CIC ? 2 : 3);
        dts[d] = e->ncps[d] + ((shape == PK_CIC && dep == PK_DIRECT) ? 3 : 4);
    }
    return sizeof(double) * (6 * its[0] * its[1] * its[2] + 3 * dts[0] * dts[1] * dts[2]) +
           sizeof(uint64_t) * mask_words + sizeof(int) * 32 + sizeof(long) * (mask_words + 1) + 16;
}

F64Args f64_args(pk_ens* e, pk_grid* g, const EnsGeom& eg, const StepParams& sp)
{
    F64Args A;
    A.g = grid_view<double>(g);
    A.eg = eg;
    A.pm = part_f64(e, e->main);
    A.po = part_f64(e, e->outbox);
    A.odir = e->odir;
    A.off = e->off;
    A.cnt = e->cnt;
    A.nstay = e->nstay;
    A.nmove = e->nmove;
    A.mvc = e->mvc;
    A.flags = e->flags;
    A.sp = sp;
    return A;
}

template <int SHAPE, int DEP>
int launch(pk_ens* e, pk_grid* g, double dt)
{
    EnsGeom eg = ens_geom(e);
    const StepParams sp = step_params(e, dt);
    const size_t lds = f64_lds(e, SHAPE, DEP, sp.mask_words);
    if (lds > 160 * 1024)
        return pk_fail(e->ctx, PK_E_CAPACITY, "particle step: supercell too large for LDS");
    const unsigned nblocks = (unsigned)(PK_XCD * (long)eg.xcd_chunk);
    const F64Args A = f64_args(e, g, eg, sp);
    pk_ktimer tm(e->ctx, "particles");
    hipLaunchKernelGGL((k_particles<SHAPE, DEP>), dim3(nblocks), dim3(PK_BLOCK), lds, e->ctx->stream, A.g, A.eg, A.pm,
                       A.po, A.odir, A.off, A.cnt, A.nstay, A.nmove, A.mvc, A.flags, A.sp);
    PK_HIP(e->ctx, hipGetLastError());
    return PK_OK;
}

// the items of one form factor (SHAPE) in one launch
template <int SHAPE, int DEP>
int batch_go(pk_batch* b, bool prepare)
{
    const void* fn = (const void*)k_particles_batch<SHAPE, DEP>;
    const size_t lds = b->cplds[SHAPE];
    const long nitems = b->cnitems[SHAPE];
    if (nitems == 0) return PK_OK;
    if (prepare) {
        PK_HIP(b->ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        return PK_OK;
    }
    const long wg = b->knob[PK_BATCH_WG_PER_CU];
    const long nb = wg > 0 && wg * 256 < nitems ? wg * 256 : nitems;
    hipLaunchKernelGGL((k_particles_batch<SHAPE, DEP>), dim3((unsigned)nb), dim3(PK_BLOCK), lds, b->ctx->stream,
                       (const F64Args*)b->pjobs, b->citems[SHAPE], nitems, b->halt);
    PK_HIP(b->ctx, hipGetLastError());
    return PK_OK;
}

// every form factor present, one launch each
int batch_dispatch(pk_batch* b, int dep, bool prepare)
{
    int r;
    if (dep == PK_DIRECT) {
        if ((r = batch_go<PK_CIC, PK_DIRECT>(b, prepare))) return r;
        return batch_go<PK_TSC, PK_DIRECT>(b, prepare);
    }
    if (dep == PK_ESIRKEPOV) {
        if ((r = batch_go<PK_CIC, PK_ESIRKEPOV>(b, prepare))) return r;
        return batch_go<PK_TSC, PK_ESIRKEPOV>(b, prepare);
    }
    return pk_fail(b->ctx, PK_E_ARG, "particle step: unknown depositor");
}

}  // namespace

int pk_particles_f64(pk_ens* e, pk_grid* g, double dt, int shape, int dep)
{
    if (shape == PK_CIC && dep == PK_DIRECT) return launch<PK_CIC, PK_DIRECT>(e, g, dt);
    if (shape == PK_CIC && dep == PK_ESIRKEPOV) return launch<PK_CIC, PK_ESIRKEPOV>(e, g, dt);
    if (shape == PK_TSC && dep == PK_DIRECT) return launch<PK_TSC, PK_DIRECT>(e, g, dt);
    if (shape == PK_TSC && dep == PK_ESIRKEPOV) return launch<PK_TSC, PK_ESIRKEPOV>(e, g, dt);
    return pk_fail(e->ctx, PK_E_ARG, "particle step: unknown shape/depositor");
}

int pk_batch_f64_prepare(pk_batch* b, double dt, int shape, int dep)
{
    const int n = b->n();
    std::vector<F64Args> jobs(n);
    b->cplds[0] = b->cplds[1] = 0;
    for (int i = 0; i < n; i++) {
        pk_ens* e = b->ens[i];
        const StepParams sp = step_params(e, dt);
        jobs[i] = f64_args(e, b->grids[i], ens_geom(e), sp);
        const int sh = pk_batch_shape_of(b, i, shape);
        const size_t l = f64_lds(e, sh, dep, sp.mask_words);
        if (l > b->cplds[sh]) b->cplds[sh] = l;
    }
    if (b->cplds[0] > 160 * 1024 || b->cplds[1] > 160 * 1024)
        return pk_fail(b->ctx, PK_E_CAPACITY, "particle step: supercell too large for LDS");
    if (int r = pk_batch_upload(b, &b->pjobs, jobs.data(), sizeof(F64Args) * n)) return r;
    return batch_dispatch(b, dep, true);
}

int pk_batch_f64_launch(pk_batch* b, int dep) { return batch_dispatch(b, dep, false); }
