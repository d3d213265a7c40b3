// pk_ensemble.hip -- supercell-ordered particle storage on the device (gfx950).
//
// Layout: supercell s owns slots [off[s], off[s+1]) of every SoA field array; cnt[s] of them
// are live.  The slack (capacity - count) absorbs the particles that arrive each step, so the
// per-step re-sort moves only the few particles that changed supercell:
//   - the particle kernel leaves stayers compacted in [0, nstay[s]) and writes leavers to the
//     outbox (same slot layout) with their direction class (0..26, 13 = none);
//   - k_insert pulls, for each direction class d in a fixed order, the class-d leavers of the
//     neighbour s - delta(d) and appends them.  No atomics decide the order: the result is
//     deterministic and equals the oracle's re-sort rule.
//   - if a supercell's slack is exhausted, a rebuild (stable rocPRIM radix sort on
//     dest * 28 + class rank) recomputes all capacities; it yields the very same order.
// pica's EnsembleSupercells (EnsembleSupercells.h:15-69) never moves particles between
// supercells; this is the device counterpart that the periodic PIC step needs.
#include <string.h>

#include "pk_internal.h"

#include <rocprim/rocprim.hpp>

namespace {

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

template <class W>
struct Fields {
    W* f[8];
};

template <class W>
Fields<W> fields_of(const pk_ens* e, void* buf)
{
    Fields<W> r;
    for (int k = 0; k < 8; k++) r.f[k] = k < e->nfield() ? (W*)e->field(buf, k) : nullptr;
    return r;
}

__device__ __forceinline__ long sc_neighbour(const EnsGeom& eg, long s, int code, int sign)
{
    int c[3] = {(int)(s / ((long)eg.nsc[1] * eg.nsc[2])), (int)((s / eg.nsc[2]) % eg.nsc[1]), (int)(s % eg.nsc[2])};
    const int dd[3] = {code / 9 - 1, (code / 3) % 3 - 1, code % 3 - 1};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        c[d] += sign * dd[d];
        if (c[d] < 0) c[d] += eg.nsc[d];
        if (c[d] >= eg.nsc[d]) c[d] -= eg.nsc[d];
    }
    return ((long)c[0] * eg.nsc[1] + c[1]) * eg.nsc[2] + c[2];
}

template <class W>
struct InsertArgs {
    EnsGeom eg;
    Fields<W> pm, po;
    const uint8_t* odir;
    const int64_t* off;
    int32_t* cnt;
    const int32_t *nstay, *nmove, *mvc;
    int32_t* flags;
};

// INSERT_SC supercells per workgroup, one per wavefront (XCD-aware over groups of them);
// INSERT_LIST sources of incoming particles per wavefront in LDS
constexpr int INSERT_SC = 4, INSERT_LIST = 256;

// one wavefront appends the incoming movers of supercell s in class order (the class-d leavers
// of the class-d neighbour, in that neighbour's outbox order); an overflow of the slack leaves
// the supercell as it is, raises the ensemble's overflow flag (and *pending).
// Three memory round trips: the 26 class counts and source outboxes; the first 64 direction
// bytes of every source outbox; the copies, one lane per arrival.  Between them the arrivals'
// sources are listed in LDS.  More than INSERT_LIST arrivals, or a source outbox above 64
// entries, take the class-by-class path.
template <class W, int NF>
__device__ __forceinline__ void insert_sc(const InsertArgs<W>& A, long s, int* pending)
{
    __shared__ long list_all[INSERT_SC][INSERT_LIST];
    const EnsGeom& eg = A.eg;
    const int lane = threadIdx.x & 63;
    long* list = list_all[(threadIdx.x >> 6) % INSERT_SC];
    const int nfinal = A.nstay[s];
    const long cap = A.off[s + 1] - A.off[s];
    // lane d < 27 (d != 13): the neighbour the class-d leavers come from, their count, and that
    // neighbour's outbox
    int m_d = 0, nm_d = 0;
    long ob_d = 0;
    if (lane < 27 && lane != 13) {
        const long src = sc_neighbour(eg, s, lane, -1);
        m_d = A.mvc[src * 27 + lane];
        if (m_d > 0) {
            nm_d = A.nmove[src];
            ob_d = A.off[src];
        }
    }
    int nin = m_d;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) nin += __shfl_xor(nin, o);
    if ((long)nfinal + nin > cap) {
        if (lane == 0) {
            atomicOr(&A.flags[PK_FLAG_OVERFLOW], 1);
            if (pending) atomicOr(pending, 1);
        }
        return;
    }
    if (lane == 0) A.cnt[s] = nfinal + nin;
    if (nin == 0) return;
    const uint64_t todo0 = __ballot(m_d > 0);
    const long base = A.off[s] + nfinal;
    const bool small = nin <= INSERT_LIST && __ballot(nm_d > 64) == 0ull;
    if (small) {
        // every source outbox has at most 64 entries: lane k loads entry k of each at once
        uint8_t od[27];
#pragma unroll
        for (int d = 0; d < 27; d++) {
            const int nm = __shfl(nm_d, d);
            const long ob = __shfl(ob_d, d);
            od[d] = ((todo0 >> d) & 1ull) && lane < nm ? A.odir[ob + lane] : (uint8_t)255;
        }
        int p = 0, found = 0;
#pragma unroll
        for (int d = 0; d < 27; d++) {
            if (!((todo0 >> d) & 1ull)) continue;
            const bool f = od[d] == d;
            const uint64_t bal = __ballot(f);
            const int t = p + __popcll(bal & ((1ull << lane) - 1ull));
            const long ob = __shfl(ob_d, d);       // by every lane: lane d may not take the branch
            if (f && t < nin) list[t] = ob + lane;
            p += __shfl(m_d, d);
            found += __popcll(bal);
        }
        if (found != nin) {            // outbox and class counts disagree: report, copy nothing
            if (lane == 0) atomicOr(&A.flags[PK_FLAG_ERR], 1);
            return;
        }
        __builtin_amdgcn_wave_barrier();
        for (int t = lane; t < nin; t += 64) {
            const long src = list[t];
#pragma unroll
            for (int q = 0; q < NF; q++) A.pm.f[q][base + t] = A.po.f[q][src];
        }
        return;
    }
    long p = base;
    for (uint64_t todo = todo0; todo; todo &= todo - 1ull) {
        const int d = __ffsll((long long)todo) - 1;
        const int m = __shfl(m_d, d), nm = __shfl(nm_d, d);
        const long ob = __shfl(ob_d, d);
        int found = 0;
        for (int k0 = 0; k0 < nm && found < m; k0 += 64) {
            const int k = k0 + lane;
            const bool f = k < nm && A.odir[ob + k] == d;
            const uint64_t bal = __ballot(f);
            if (f) {
                const long dst = p + found + __popcll(bal & ((1ull << lane) - 1ull));
#pragma unroll
                for (int q = 0; q < NF; q++) A.pm.f[q][dst] = A.po.f[q][ob + k];
            }
            found += __popcll(bal);
        }
        p += m;
    }
}



template <class W, int NF>
__global__ void __launch_bounds__(64 * INSERT_SC) k_insert(InsertArgs<W> A, int gchunk)
{
    const long s = pk_xcd_supercell(blockIdx.x, gchunk) * INSERT_SC + (threadIdx.x >> 6);
    if (s >= A.eg.nsct) return;
    insert_sc<W, NF>(A, s, nullptr);
}

template <class W, int NF>
__global__ void __launch_bounds__(64) k_insert_batch(const InsertArgs<W>* __restrict__ jobs,
                                                     const int2* __restrict__ items, long nitems, int* halt)
{
    if (halt[PK_HALT_STOP]) return;
    for (long w = blockIdx.x; w < nitems; w += gridDim.x) {
        const int2 it = items[w];
        insert_sc<W, NF>(jobs[it.x], it.y, halt + PK_HALT_PENDING);
    }
}

template <class W>
InsertArgs<W> insert_args(const pk_ens* e)
{
    InsertArgs<W> A;
    A.eg = ens_geom(e);
    A.pm = fields_of<W>(e, e->main);
    A.po = fields_of<W>(e, e->outbox);
    A.odir = e->odir;
    A.off = e->off;
    A.cnt = e->cnt;
    A.nstay = e->nstay;
    A.nmove = e->nmove;
    A.mvc = e->mvc;
    A.flags = e->flags;
    return A;
}

// ---- rebuild ----
constexpr uint64_t SRC_MAIN = 0, SRC_OUT = 1ull << 62, SRC_STAGE = 2ull << 62;
constexpr uint64_t SRC_MASK = (1ull << 62) - 1;

__global__ void k_entries_per_sc(long nsct, const int32_t* nstay, const int32_t* nmove, int64_t* n)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsct) n[s] = (int64_t)nstay[s] + nmove[s];
}

__global__ void __launch_bounds__(64) k_keys_main(EnsGeom eg, const int64_t* off, const int32_t* nstay,
                                                  const int32_t* nmove, const uint8_t* odir, const int64_t* ebase,
                                                  uint32_t* keys, uint64_t* srcs)
{
    const long s = blockIdx.x;
    if (s >= eg.nsct) return;
    const long b = ebase[s];
    const int ns = nstay[s], nm = nmove[s];
    for (int j = threadIdx.x; j < ns; j += 64) {
        keys[b + j] = (uint32_t)(s * 28);
        srcs[b + j] = SRC_MAIN | (uint64_t)(off[s] + j);
    }
    for (int k = threadIdx.x; k < nm; k += 64) {
        const int code = odir[off[s] + k];
        const long dst = sc_neighbour(eg, s, code, +1);
        keys[b + ns + k] = (uint32_t)(dst * 28 + code + 1);
        srcs[b + ns + k] = SRC_OUT | (uint64_t)(off[s] + k);
    }
}

struct StageGeom {
    double cellinv[3], scinv[3], L[3];
};

// supercell key of a host-supplied particle: pica's getSupercellIndex for absolute positions
// (EnsembleSupercells.h:56); for the split form the supercell of the particle's cell.
__global__ void k_keys_stage(EnsGeom eg, StageGeom sg, int split, const double* x, const double* y, const double* z,
                             long n, uint32_t* keys, uint64_t* srcs, int32_t* flags)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double p[3] = {x[i], y[i], z[i]};
    int sc[3];
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        if (!(p[d] >= 0.0 && p[d] < sg.L[d])) ok = false;
        if (split) {
            int c = (int)floor(p[d] * sg.cellinv[d]);
            c = c < 0 ? 0 : (c >= eg.ncell[d] ? eg.ncell[d] - 1 : c);
            sc[d] = c / eg.ncps[d];
        } else {
            sc[d] = (int)floor((p[d] - 0.0) * sg.scinv[d]);
            if (sc[d] < 0 || sc[d] >= eg.nsc[d]) ok = false;
            sc[d] = sc[d] < 0 ? 0 : (sc[d] >= eg.nsc[d] ? eg.nsc[d] - 1 : sc[d]);
        }
    }
    if (!ok) atomicOr(&flags[PK_FLAG_ERR], 1);
    keys[i] = (uint32_t)((((long)sc[0] * eg.nsc[1] + sc[1]) * eg.nsc[2] + sc[2]) * 28);
    srcs[i] = SRC_STAGE | (uint64_t)i;
}

// per-destination counts of the sorted keys: the keys of one destination are consecutive, so the
// first lane of each run within a wavefront adds the run's length (one atomic per run, not per key)
__global__ void k_count_keys(long n, const uint32_t* keys, int32_t* cnt)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool act = i < n;
    const uint32_t d = act ? keys[i] / 28 : 0xffffffffu;
    const uint32_t dprev = __shfl_up(d, 1);
    const bool head = act && (lane == 0 || d != dprev);
    const uint64_t heads = __ballot(head);
    const int nact = __popcll(__ballot(act));          // the active lanes are a prefix of the wavefront
    if (head) {
        const uint64_t later = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
        const int next = later ? __ffsll((long long)later) - 1 : nact;
        atomicAdd(&cnt[d], next - lane);
    }
}

__global__ void k_caps(long nsct, const int32_t* cnt, double frac, int add, int64_t* cap64, int64_t* cnt64,
                       int32_t* maxcap)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int capi = 0;
    if (s < nsct) {
        const int c = cnt[s];
        const long cap = (long)c + (long)(c * frac) + add;
        cap64[s] = cap;
        cnt64[s] = c;
        capi = (int)cap;
    }
    // the wavefront's largest capacity, then one atomic per wavefront
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int t = __shfl_xor(capi, o);
        capi = t > capi ? t : capi;
    }
    if ((threadIdx.x & 63) == 0 && capi > 0) atomicMax(maxcap, capi);
}

template <class W, int NF>
__global__ void k_gather(long n, const uint32_t* keys, const uint64_t* srcs, const int64_t* newoff,
                         const int64_t* runstart, Fields<W> dst, Fields<W> from_main, Fields<W> from_out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long d = keys[i] / 28;
    const long slot = newoff[d] + (i - runstart[d]);
    const uint64_t sv = srcs[i];
    const long si = (long)(sv & SRC_MASK);
    const Fields<W>& from = (sv >> 62) == 0 ? from_main : from_out;
#pragma unroll
    for (int q = 0; q < NF; q++) dst.f[q][slot] = from.f[q][si];
}

// staged FP64 particles -> FP64 absolute storage
__global__ void k_gather_stage_f64(long n, const uint32_t* keys, const uint64_t* srcs, const int64_t* newoff,
                                   const int64_t* runstart, PartF64 dst, Fields<double> st)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long d = keys[i] / 28;
    const long slot = newoff[d] + (i - runstart[d]);
    const long si = (long)(srcs[i] & SRC_MASK);
    dst.x[slot] = st.f[0][si]; dst.y[slot] = st.f[1][si]; dst.z[slot] = st.f[2][si];
    dst.ux[slot] = st.f[3][si]; dst.uy[slot] = st.f[4][si]; dst.uz[slot] = st.f[5][si];
    dst.w[slot] = st.f[6][si];
}

// cell index and in-cell offset of an absolute position for the split form
__device__ __forceinline__ void to_split(double x, double cellinv, int ncell, int& c, float& o)
{
    const double t = x * cellinv;
    c = (int)floor(t);
    c = c < 0 ? 0 : (c >= ncell ? ncell - 1 : c);
    float f = (float)(t - (double)c);
    if (!(f >= 0.0f)) f = 0.0f;
    if (f >= 1.0f) f = 0x1.fffffep-1f;   // largest float below 1
    o = f;
}

__global__ void k_gather_stage_split(EnsGeom eg, StageGeom sg, long n, const uint32_t* keys, const uint64_t* srcs,
                                     const int64_t* newoff, const int64_t* runstart, PartSplit dst, Fields<double> st)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long d = keys[i] / 28;
    const long slot = newoff[d] + (i - runstart[d]);
    const long si = (long)(srcs[i] & SRC_MASK);
    int c[3];
    float o[3];
#pragma unroll
    for (int k = 0; k < 3; k++) to_split(st.f[k][si], sg.cellinv[k], eg.ncell[k], c[k], o[k]);
    dst.cell[slot] = (c[0] * eg.ncell[1] + c[1]) * eg.ncell[2] + c[2];
    dst.ox[slot] = o[0]; dst.oy[slot] = o[1]; dst.oz[slot] = o[2];
    dst.ux[slot] = (float)st.f[3][si]; dst.uy[slot] = (float)st.f[4][si]; dst.uz[slot] = (float)st.f[5][si];
    dst.w[slot] = (float)st.f[6][si];
}

__global__ void k_reset_after_rebuild(long nsct, int32_t* nstay, int32_t* nmove, const int32_t* cnt)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsct) {
        nstay[s] = cnt[s];
        nmove[s] = 0;
    }
}

// ---- loader ----
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0, 1) from the counter (seed, particle id, stream j)
__device__ __forceinline__ double urand(uint64_t key, int j)
{
    return (double)(mix64(key + (uint64_t)j * 0xD1B54A32D192ED03ull) >> 11) * 0x1.0p-53;
}

// ln(x) for 0 < x <= 1 and (sin, cos)(2 pi U) for 0 <= U < 1 from IEEE +, -, *, / only (this file
// is built without FP contraction), so that a CPU restatement of the loader with the same
// operations reproduces every bit.  ln: x = m 2^e with m in [sqrt(1/2), sqrt(2)], ln m = 2 atanh(s),
// s = (m - 1)/(m + 1), |s| <= 0.172, odd series to s^23 (truncation below 1e-19).  sincos: U in
// octant q = floor(8U) at fraction f, reduced to r = f pi/4 or (1 - f) pi/4 in [0, pi/4], Taylor
// series to r^19 / r^20 (truncation below 1e-21).
__device__ __forceinline__ double det_log(double x)
{
    uint64_t bits = (uint64_t)__double_as_longlong(x);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull));
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double p = 1.0 / 23.0;
    p = 1.0 / 21.0 + s2 * p;
    p = 1.0 / 19.0 + s2 * p;
    p = 1.0 / 17.0 + s2 * p;
    p = 1.0 / 15.0 + s2 * p;
    p = 1.0 / 13.0 + s2 * p;
    p = 1.0 / 11.0 + s2 * p;
    p = 1.0 / 9.0 + s2 * p;
    p = 1.0 / 7.0 + s2 * p;
    p = 1.0 / 5.0 + s2 * p;
    p = 1.0 / 3.0 + s2 * p;
    p = 1.0 + s2 * p;
    const double lnm = 2.0 * s * p;
    const double de = (double)e;
    return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + lnm);
}

__device__ __forceinline__ void det_sincos_2pi(double U, double* sn, double* cs)
{
    const double t = U * 8.0;
    const int q = (int)t;
    const double f = t - (double)q;
    const double r = ((q & 1) ? 1.0 - f : f) * 0.78539816339744830962;
    const double r2 = r * r;
    double ps = 1.0 / 121645100408832000.0;                // 1/19!
    ps = 1.0 / 355687428096000.0 - r2 * ps;                  // 1/17!
    ps = 1.0 / 1307674368000.0 - r2 * ps;                    // 1/15!
    ps = 1.0 / 6227020800.0 - r2 * ps;                       // 1/13!
    ps = 1.0 / 39916800.0 - r2 * ps;                         // 1/11!
    ps = 1.0 / 362880.0 - r2 * ps;                           // 1/9!
    ps = 1.0 / 5040.0 - r2 * ps;                             // 1/7!
    ps = 1.0 / 120.0 - r2 * ps;                              // 1/5!
    ps = 1.0 / 6.0 - r2 * ps;                                // 1/3!
    const double sr = r - r * r2 * ps;
    double pc = 1.0 / 2432902008176640000.0;               // 1/20!
    pc = 1.0 / 6402373705728000.0 - r2 * pc;                 // 1/18!
    pc = 1.0 / 20922789888000.0 - r2 * pc;                   // 1/16!
    pc = 1.0 / 87178291200.0 - r2 * pc;                      // 1/14!
    pc = 1.0 / 479001600.0 - r2 * pc;                        // 1/12!
    pc = 1.0 / 3628800.0 - r2 * pc;                          // 1/10!
    pc = 1.0 / 40320.0 - r2 * pc;                            // 1/8!
    pc = 1.0 / 720.0 - r2 * pc;                              // 1/6!
    pc = 1.0 / 24.0 - r2 * pc;                               // 1/4!
    pc = 0.5 - r2 * pc;                                      // 1/2!
    const double cr = 1.0 - r2 * pc;
    // octant q: angle q pi/4 + f pi/4; odd octants measured back from the next multiple of pi/2
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

struct LoadParams {
    int ppc;
    double uth, drift, weight;
    uint64_t seed;
    double step[3];
    int nmodes;              // velocity modes: u += a sin(2 pi k . x / L + phase)
    const double* modes;     // nmodes x (kx, ky, kz, ax, ay, az, phase), device
};

// Particle k of every cell has an id from its cell and k; the slot l it takes in its supercell
// interleaves the cells, so that the lanes of a wavefront sit in different cells and their LDS
// deposits do not collide on one address.  Unpaired: slot = k * cells + cell.  Paired (the FP32
// split form, whose lanes take two adjacent slots): particles 2q and 2q + 1 of a cell in
// adjacent slots, pairs by (q, cell), then particle ppc - 1 of every cell when ppc is odd --
// the order k_interleave_split gives.
__device__ __forceinline__ void thermal_particle(const EnsGeom& eg, const LoadParams& lp, long s, int l, bool paired,
                                                 int gc[3], double U[3], double u[3])
{
    const int sidx[3] = {(int)(s / ((long)eg.nsc[1] * eg.nsc[2])), (int)((s / eg.nsc[2]) % eg.nsc[1]),
                         (int)(s % eg.nsc[2])};
    const int ncsc = eg.ncps[0] * eg.ncps[1] * eg.ncps[2];
    int lc = l % ncsc, k = l / ncsc;
    if (paired) {
        const int npair = (lp.ppc & ~1) * ncsc;
        if (l < npair) {
            lc = (l >> 1) % ncsc;
            k = 2 * ((l >> 1) / ncsc) + (l & 1);
        } else {
            lc = l - npair;
            k = lp.ppc - 1;
        }
    }
    const int a = lc / (eg.ncps[1] * eg.ncps[2]), b = (lc / eg.ncps[2]) % eg.ncps[1], c = lc % eg.ncps[2];
    gc[0] = sidx[0] * eg.ncps[0] + a;
    gc[1] = sidx[1] * eg.ncps[1] + b;
    gc[2] = sidx[2] * eg.ncps[2] + c;
    const uint64_t id = ((((uint64_t)gc[0] * eg.ncell[1] + gc[1]) * eg.ncell[2] + gc[2]) * lp.ppc) + k;
    const uint64_t key = mix64(lp.seed ^ mix64(id));
    U[0] = urand(key, 0); U[1] = urand(key, 1); U[2] = urand(key, 2);
    // Box-Muller normals from streams 3..6 (1 - U is exact and > 0)
    const double r1 = sqrt(-2.0 * det_log(1.0 - urand(key, 3)));
    const double r2 = sqrt(-2.0 * det_log(1.0 - urand(key, 5)));
    double s1, c1, s2, c2;
    det_sincos_2pi(urand(key, 4), &s1, &c1);
    det_sincos_2pi(urand(key, 6), &s2, &c2);
    u[0] = lp.uth * r1 * c1 + ((id & 1ull) ? -lp.drift : lp.drift);
    u[1] = lp.uth * r1 * s1;
    u[2] = lp.uth * r2 * c2;
    for (int k = 0; k < lp.nmodes; k++) {
        const double* md = lp.modes + 7 * k;
        double ph = md[6];
        for (int d = 0; d < 3; d++) ph += 2.0 * PK_PI * md[d] * ((double)gc[d] + U[d]) / (double)eg.ncell[d];
        const double sn = sin(ph);
        for (int d = 0; d < 3; d++) u[d] += md[3 + d] * sn;
    }
}

__global__ void __launch_bounds__(256) k_load_f64(EnsGeom eg, LoadParams lp, const int64_t* off, int per_sc, PartF64 pm)
{
    const long s = blockIdx.x;
    for (int l = threadIdx.x; l < per_sc; l += blockDim.x) {
        int gc[3];
        double U[3], u[3];
        thermal_particle(eg, lp, s, l, false, gc, U, u);
        const long p = off[s] + l;
        pm.x[p] = ((double)gc[0] + U[0]) * lp.step[0];
        pm.y[p] = ((double)gc[1] + U[1]) * lp.step[1];
        pm.z[p] = ((double)gc[2] + U[2]) * lp.step[2];
        pm.ux[p] = u[0]; pm.uy[p] = u[1]; pm.uz[p] = u[2];
        pm.w[p] = lp.weight;
    }
}

__global__ void __launch_bounds__(256) k_load_split(EnsGeom eg, LoadParams lp, const int64_t* off, int per_sc,
                                                    PartSplit pm)
{
    const long s = blockIdx.x;
    for (int l = threadIdx.x; l < per_sc; l += blockDim.x) {
        int gc[3];
        double U[3], u[3];
        thermal_particle(eg, lp, s, l, true, gc, U, u);
        const long p = off[s] + l;
        float o[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            // the same rounding as the absolute form's (gc + U) * h, taken relative to the cell
            const double xa = ((double)gc[d] + U[d]) * lp.step[d];
            int c;
            to_split(xa, 1.0 / lp.step[d], eg.ncell[d], c, o[d]);
            if (c != gc[d]) o[d] = c < gc[d] ? 0.0f : 0x1.fffffep-1f;
        }
        pm.cell[p] = (gc[0] * eg.ncell[1] + gc[1]) * eg.ncell[2] + gc[2];
        pm.ox[p] = o[0]; pm.oy[p] = o[1]; pm.oz[p] = o[2];
        pm.ux[p] = (float)u[0]; pm.uy[p] = (float)u[1]; pm.uz[p] = (float)u[2];
        pm.w[p] = (float)lp.weight;
    }
}

__global__ void k_fill_i32(int32_t* p, long n, int32_t v)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_uniform_off(long nsct, int64_t cap, int64_t* off)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s <= nsct) off[s] = s * cap;
}

// ---- download / cells ----
__global__ void __launch_bounds__(256) k_compact_f64(EnsGeom eg, PartF64 pm, const int64_t* off, const int32_t* cnt,
                                                     const int64_t* runstart, Fields<double> out)
{
    const long s = blockIdx.x;
    const int n = cnt[s];
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
        const long p = off[s] + l, o = runstart[s] + l;
        out.f[0][o] = pm.x[p]; out.f[1][o] = pm.y[p]; out.f[2][o] = pm.z[p];
        out.f[3][o] = pm.ux[p]; out.f[4][o] = pm.uy[p]; out.f[5][o] = pm.uz[p];
        out.f[6][o] = pm.w[p];
    }
}

__global__ void __launch_bounds__(256) k_compact_split(EnsGeom eg, PartSplit pm, const int64_t* off, const int32_t* cnt,
                                                       const int64_t* runstart, Fields<double> out)
{
    const long s = blockIdx.x;
    const int n = cnt[s];
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
        const long p = off[s] + l, o = runstart[s] + l;
        const int c = pm.cell[p];
        int cc[3];
        cell_coords(c, eg, cc);
        out.f[0][o] = ((double)cc[0] + (double)pm.ox[p]) * eg.step[0];
        out.f[1][o] = ((double)cc[1] + (double)pm.oy[p]) * eg.step[1];
        out.f[2][o] = ((double)cc[2] + (double)pm.oz[p]) * eg.step[2];
        out.f[3][o] = pm.ux[p]; out.f[4][o] = pm.uy[p]; out.f[5][o] = pm.uz[p];
        out.f[6][o] = pm.w[p];
    }
}

// linear physical cell of every particle (supercell order) and per-cell counts.  For the
// absolute form the key is floor((x - min) * inverse(cellSize)) per dimension, pica's
// getSupercellIndex with one cell per supercell (EnsembleSupercells.h:31-33,56).
__global__ void __launch_bounds__(256) k_cells(EnsGeom eg, StageGeom sg, int split, Fields<double> pf,
                                               const int32_t* scell, const int64_t* off, const int32_t* cnt,
                                               const int64_t* runstart, int32_t* counts, int32_t* cells)
{
    const long s = blockIdx.x;
    const int n = cnt[s];
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
        const long p = off[s] + l;
        int c;
        if (split) {
            c = scell[p];
        } else {
            int k[3];
#pragma unroll
            for (int d = 0; d < 3; d++) k[d] = (int)floor((pf.f[d][p] - 0.0) * sg.cellinv[d]);
            c = (k[0] * eg.ncell[1] + k[1]) * eg.ncell[2] + k[2];
        }
        if (cells) cells[runstart[s] + l] = c;
        if (counts) atomicAdd(&counts[c], 1);
    }
}

StageGeom stage_geom(const pk_ens* e)
{
    StageGeom sg;
    for (int d = 0; d < 3; d++) {
        sg.L[d] = (double)e->ncell[d] * e->step[d];
        const double cellSize = (sg.L[d] - 0.0) / (double)e->ncell[d];   // EnsembleSupercells.h:31
        sg.cellinv[d] = 1.0 / cellSize;
        sg.scinv[d] = 1.0 / ((double)e->ncps[d] * cellSize);            // :32-33
    }
    return sg;
}

// RAII device buffer
struct DBuf {
    void* p = nullptr;
    ~DBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes)
    {
        return hipMalloc(&p, bytes ? bytes : 16);
    }
    template <class T>
    T* as() const
    {
        return (T*)p;
    }
};

template <class T>
int exclusive_scan(pk_ctx* c, const T* in, T* out, long n)
{
    size_t bytes = 0;
    PK_HIP(c, rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), c->stream));
    DBuf tmp;
    PK_HIP(c, tmp.alloc(bytes));
    PK_HIP(c, rocprim::exclusive_scan(tmp.p, bytes, in, out, T(0), (size_t)n, rocprim::plus<T>(), c->stream));
    return PK_OK;
}

}  // namespace

EnsGeom ens_geom(const pk_ens* e)
{
    EnsGeom g;
    for (int d = 0; d < 3; d++) {
        g.nsc[d] = e->nsc[d];
        g.ncps[d] = e->ncps[d];
        g.ncell[d] = e->ncell[d];
        g.step[d] = e->step[d];
    }
    g.nsct = e->nsct;
    g.xcd_chunk = (int)((e->nsct + PK_XCD - 1) / PK_XCD);
    g.div12 = make_fastdiv((uint32_t)(e->ncell[1] * e->ncell[2]));
    g.div2 = make_fastdiv((uint32_t)e->ncell[2]);
    return g;
}

int pk_ens_alloc(pk_ens* e, size_t capacity)
{
    pk_ctx* c = e->ctx;
    void *m = nullptr, *o = nullptr;
    uint8_t* od = nullptr;
    const size_t bytes = capacity * e->esize() * e->nfield();
    if (hipMalloc(&m, bytes ? bytes : 16) != hipSuccess || hipMalloc(&o, bytes ? bytes : 16) != hipSuccess ||
        hipMalloc((void**)&od, capacity ? capacity : 16) != hipSuccess) {
        if (m) (void)hipFree(m);
        if (o) (void)hipFree(o);
        if (od) (void)hipFree(od);
        (void)hipGetLastError();
        return pk_fail(c, PK_E_ALLOC, "particle buffers: hipMalloc failed");
    }
    if (e->main) (void)hipFree(e->main);
    if (e->outbox) (void)hipFree(e->outbox);
    if (e->odir) (void)hipFree(e->odir);
    e->main = m;
    e->outbox = o;
    e->odir = od;
    e->capacity = capacity;
    return PK_OK;
}

// ---- interleave (FP32 split ensembles) ----
constexpr int IL_RMAX = 128;     // most particles of one cell that are interleaved
// particles held per thread: IL_P = 8 for supercells of up to 2048 particles, 16 for 2049 to 4096
// (launched only when the largest supercell capacity exceeds 2048)
inline bool il_big(int maxcap) { return maxcap > 256 * 8; }

// One workgroup reorders supercell s in place.  With r the rank of a particle among those of its
// cell (in storage order) and h the cell's count, the order is: first the pairs (ranks 2q and
// 2q + 1 of one cell) by (q, cell), the two of a pair adjacent; then the last particle of every
// cell with an odd count, by cell.  So a particle and its neighbour slot mostly share a cell, and
// consecutive pairs sit in different cells.  Supercells of more than 64 cells or 256 IL_P
// particles, or with a cell of more than IL_RMAX particles, are left as they are.
//
// Ranks come 256 particles at a time: each particle sets its lane bit in cm[cell][wavefront]
// (one LDS atomic OR), and its rank within the chunk is the popcount of the lanes below it in
// the same cell.  The per-pair-row cell masks M and their prefix counts come from wavefront 0
// with ballots and shuffles.  The eight 32-bit fields then pass one at a time through an LDS
// stage (scattered LDS writes, coalesced global loads and stores), two fields ahead in flight.
template <int IL_P>
__device__ __forceinline__ void interleave_sc(const EnsGeom& eg, const PartSplit& pm, const int64_t* __restrict__ off,
                                              const int32_t* __restrict__ cnt, long s)
{
    const int n = cnt[s];
    const int ncl = eg.ncps[0] * eg.ncps[1] * eg.ncps[2];
    // each size class of supercells has its own form: 16 per thread only above 256 * 8 particles
    if (n < 2 || n > 256 * IL_P || (IL_P > 8 && n <= 256 * 8) || ncl > 64) return;
    __shared__ unsigned long long cm[64][4];
    __shared__ int hist[64];
    __shared__ unsigned long long M[IL_RMAX / 2];
    __shared__ int start[IL_RMAX / 2];
    __shared__ unsigned long long odd;
    __shared__ int maxc, bad, npair;
    __shared__ uint32_t stage[2][256 * IL_P];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const long base = off[s];
    const int m0[3] = {(int)(s / ((long)eg.nsc[1] * eg.nsc[2])) * eg.ncps[0],
                       (int)((s / eg.nsc[2]) % eg.nsc[1]) * eg.ncps[1], (int)(s % eg.nsc[2]) * eg.ncps[2]};
    uint32_t* const fld[8] = {(uint32_t*)pm.cell, (uint32_t*)pm.ox, (uint32_t*)pm.oy, (uint32_t*)pm.oz,
                              (uint32_t*)pm.ux,   (uint32_t*)pm.uy, (uint32_t*)pm.uz, (uint32_t*)pm.w};
    uint32_t v[8][IL_P];
    int key[IL_P], pos[IL_P];
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < IL_P; k++)
            if (tid + 256 * k < n) v[f][k] = fld[f][base + tid + 256 * k];
#pragma unroll
    for (int k = 0; k < IL_P; k++) {
        key[k] = -1;
        pos[k] = 0;
        if (tid + 256 * k < n) {
            int cc[3];
            cell_coords((int32_t)v[0][k], eg, cc);
            int kk = 0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                int l = cc[d] - m0[d];
                l = l < 0 ? 0 : (l >= eg.ncps[d] ? eg.ncps[d] - 1 : l);   // a stray particle still gets a slot
                kk = kk * eg.ncps[d] + l;
            }
            key[k] = kk;
        }
    }
    cm[tid >> 2][tid & 3] = 0ull;
    if (tid < 64) hist[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < IL_P; k++) {
        if (256 * k >= n) break;
        const int kk = key[k];
        if (kk >= 0) atomicOr(&cm[kk][wave], 1ull << lane);
        __syncthreads();
        if (kk >= 0) {
            int r = hist[kk] + __popcll(cm[kk][wave] & below);
            for (int w = 0; w < wave; w++) r += __popcll(cm[kk][w]);
            pos[k] = r;                                   // the rank for now
        }
        __syncthreads();
        // the thread that counts cell tid also clears its masks: no other thread touches them here
        if (tid < 64) {
            hist[tid] += __popcll(cm[tid][0]) + __popcll(cm[tid][1]) + __popcll(cm[tid][2]) + __popcll(cm[tid][3]);
            cm[tid][0] = 0ull; cm[tid][1] = 0ull; cm[tid][2] = 0ull; cm[tid][3] = 0ull;
        }
        __syncthreads();
    }
    // pair row q: M[q] bit c = cell c holds more than 2q + 1 particles, start[q] = slots of the
    // rows below; odd bit c = cell c has an odd count; npair = slots of all pairs
    if (wave == 0) {
        const int h = lane < ncl ? hist[lane] : 0;
        int mx = h;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const int t = __shfl_xor(mx, o);
            mx = t > mx ? t : mx;
        }
        uint64_t mine = 0ull;
        const int qmax = mx > IL_RMAX ? IL_RMAX / 2 : mx / 2;
        for (int q = 0; q < qmax; q++) {
            const uint64_t b = __ballot(h > 2 * q + 1);
            if (lane == q) mine = b;
        }
        const uint64_t ob = __ballot(h & 1);
        const int c = lane < qmax ? 2 * __popcll(mine) : 0;
        int incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o);
            if (lane >= o) incl += t;
        }
        if (lane < qmax) {
            M[lane] = mine;
            start[lane] = incl - c;
        }
        if (lane == 63) {
            maxc = mx;
            odd = ob;
            npair = incl;
            bad = mx > IL_RMAX || incl + __popcll(ob) != n;
        }
    }
    __syncthreads();
    const int mx = maxc;
    // every particle must name a set bit of M or of odd: otherwise the supercell is left as it is
    bool ok = true;
#pragma unroll
    for (int k = 0; k < IL_P; k++)
        if (key[k] >= 0) {
            const int r = pos[k], h = hist[key[k]];
            const uint64_t below_c = (1ull << key[k]) - 1ull;
            if (r < 0 || r >= h || h > mx || mx > IL_RMAX)
                ok = false;
            else if ((h & 1) && r == h - 1)
                pos[k] = npair + __popcll(odd & below_c);
            else if ((M[r >> 1] >> key[k]) & 1ull)
                pos[k] = start[r >> 1] + 2 * __popcll(M[r >> 1] & below_c) + (r & 1);
            else
                ok = false;
        }
    if (!ok) bad = 1;
    __syncthreads();
    if (bad) return;
#pragma unroll
    for (int f = 0; f < 8; f++) {
        uint32_t* const st = stage[f & 1];
#pragma unroll
        for (int k = 0; k < IL_P; k++)
            if (key[k] >= 0) st[pos[k]] = v[f][k];
        __syncthreads();
        if (f + 3 < 8)
#pragma unroll
            for (int k = 0; k < IL_P; k++)
                if (tid + 256 * k < n) v[f + 3][k] = fld[f + 3][base + tid + 256 * k];
#pragma unroll
        for (int k = 0; k < IL_P; k++)
            if (tid + 256 * k < n) fld[f][base + tid + 256 * k] = st[tid + 256 * k];
    }
}

// FP64 ensembles (one particle per lane in the FP64 kernel): the order is rank-major, the particles
// of rank r (r-th of their cell in storage order) by cell, so the lanes of a wavefront sit in
// different cells, as the loader leaves them (slot = r * cells + cell when all counts are equal).
// A particle's cell is floor(x / cellSize) per dimension (k_cells), relative to the supercell.
// Ranks as in interleave_sc; the seven 64-bit fields then pass one at a time through a 64-bit
// LDS stage, two fields ahead in flight.  Supercells of more than 64 cells or 256 IL_P particles,
// or with a cell of more than IL_RMAX particles, are left as they are.
template <int IL_P>
__device__ __forceinline__ void interleave_f64_sc(const EnsGeom& eg, const StageGeom& sg, const PartF64& pm,
                                                  const int64_t* __restrict__ off, const int32_t* __restrict__ cnt, long s)
{
    const int n = cnt[s];
    const int ncl = eg.ncps[0] * eg.ncps[1] * eg.ncps[2];
    if (n < 2 || n > 256 * IL_P || (IL_P > 8 && n <= 256 * 8) || ncl > 64) return;
    __shared__ unsigned long long cm[64][4];
    __shared__ int hist[64];
    __shared__ unsigned long long M[IL_RMAX];
    __shared__ int start[IL_RMAX];
    __shared__ int maxc, bad;
    __shared__ __align__(16) uint64_t stage[2][256 * IL_P];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const long base = off[s];
    const int m0[3] = {(int)(s / ((long)eg.nsc[1] * eg.nsc[2])) * eg.ncps[0],
                       (int)((s / eg.nsc[2]) % eg.nsc[1]) * eg.ncps[1], (int)(s % eg.nsc[2]) * eg.ncps[2]};
    uint64_t* const fld[7] = {(uint64_t*)pm.x,  (uint64_t*)pm.y,  (uint64_t*)pm.z, (uint64_t*)pm.ux,
                              (uint64_t*)pm.uy, (uint64_t*)pm.uz, (uint64_t*)pm.w};
    constexpr int RING = 3;            // fields held in registers: the three positions, then a ring
    uint64_t v[RING][IL_P];
    int key[IL_P], pos[IL_P];
#pragma unroll
    for (int f = 0; f < RING; f++)
#pragma unroll
        for (int k = 0; k < IL_P; k++)
            if (tid + 256 * k < n) v[f][k] = fld[f][base + tid + 256 * k];
#pragma unroll
    for (int k = 0; k < IL_P; k++) {
        key[k] = -1;
        pos[k] = 0;
        if (tid + 256 * k < n) {
            int kk = 0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
                int l = (int)floor((__longlong_as_double((long long)v[d][k]) - 0.0) * sg.cellinv[d]) - m0[d];
                l = l < 0 ? 0 : (l >= eg.ncps[d] ? eg.ncps[d] - 1 : l);   // a stray particle still gets a slot
                kk = kk * eg.ncps[d] + l;
            }
            key[k] = kk;
        }
    }
    cm[tid >> 2][tid & 3] = 0ull;
    if (tid < 64) hist[tid] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < IL_P; k++) {
        if (256 * k >= n) break;
        const int kk = key[k];
        if (kk >= 0) atomicOr(&cm[kk][wave], 1ull << lane);
        __syncthreads();
        if (kk >= 0) {
            int r = hist[kk] + __popcll(cm[kk][wave] & below);
            for (int w = 0; w < wave; w++) r += __popcll(cm[kk][w]);
            pos[k] = r;                                   // the rank for now
        }
        __syncthreads();
        if (tid < 64) {
            hist[tid] += __popcll(cm[tid][0]) + __popcll(cm[tid][1]) + __popcll(cm[tid][2]) + __popcll(cm[tid][3]);
            cm[tid][0] = 0ull; cm[tid][1] = 0ull; cm[tid][2] = 0ull; cm[tid][3] = 0ull;
        }
        __syncthreads();
    }
    // rank row r: M[r] bit c = cell c holds more than r particles, start[r] = slots of the rows
    // below; lane l builds rows l and l + 64
    if (wave == 0) {
        const int h = lane < ncl ? hist[lane] : 0;
        int mx = h;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const int t = __shfl_xor(mx, o);
            mx = t > mx ? t : mx;
        }
        const int rmax = mx > IL_RMAX ? IL_RMAX : mx;
        uint64_t mine0 = 0ull, mine1 = 0ull;
        for (int r = 0; r < rmax; r++) {
            const uint64_t b = __ballot(h > r);
            if (r == lane) mine0 = b;
            if (r == lane + 64) mine1 = b;
        }
        const int c0 = lane < rmax ? __popcll(mine0) : 0, c1 = lane + 64 < rmax ? __popcll(mine1) : 0;
        int i0 = c0, i1 = c1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t0 = __shfl_up(i0, o), t1 = __shfl_up(i1, o);
            if (lane >= o) {
                i0 += t0;
                i1 += t1;
            }
        }
        const int tot0 = __shfl(i0, 63), tot1 = __shfl(i1, 63);
        M[lane] = mine0;
        start[lane] = i0 - c0;
        if (lane + 64 < IL_RMAX) {
            M[lane + 64] = mine1;
            start[lane + 64] = tot0 + i1 - c1;
        }
        if (lane == 63) {
            maxc = mx;
            bad = mx > IL_RMAX || tot0 + tot1 != n;
        }
    }
    __syncthreads();
    const int mx = maxc;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < IL_P; k++)
        if (key[k] >= 0) {
            const int r = pos[k], h = hist[key[k]];
            const uint64_t below_c = (1ull << key[k]) - 1ull;
            if (r < 0 || r >= h || h > mx || mx > IL_RMAX || !((M[r] >> key[k]) & 1ull))
                ok = false;
            else
                pos[k] = start[r] + __popcll(M[r] & below_c);
        }
    if (!ok) bad = 1;
    __syncthreads();
    if (bad) return;
#pragma unroll
    for (int f = 0; f < 7; f++) {
        uint64_t* const st = stage[f & 1];
#pragma unroll
        for (int k = 0; k < IL_P; k++)
            if (key[k] >= 0) st[pos[k]] = v[f % RING][k];
        __syncthreads();
        if (f + RING < 7)
#pragma unroll
            for (int k = 0; k < IL_P; k++)
                if (tid + 256 * k < n) v[f % RING][k] = fld[f + RING][base + tid + 256 * k];
#pragma unroll
        for (int k = 0; k < IL_P; k++)
            if (tid + 256 * k < n) fld[f][base + tid + 256 * k] = st[tid + 256 * k];
    }
}

template <int IL_P>
__global__ void __launch_bounds__(256) k_interleave_f64(EnsGeom eg, StageGeom sg, PartF64 pm,
                                                        const int64_t* __restrict__ off, const int32_t* __restrict__ cnt)
{
    const long s = pk_xcd_supercell(blockIdx.x, eg.xcd_chunk);
    if (s >= eg.nsct) return;
    interleave_f64_sc<IL_P>(eg, sg, pm, off, cnt, s);
}

template <int IL_P>
__global__ void __launch_bounds__(256) k_interleave_split(EnsGeom eg, PartSplit pm, const int64_t* __restrict__ off,
                                                          const int32_t* __restrict__ cnt)
{
    const long s = pk_xcd_supercell(blockIdx.x, eg.xcd_chunk);
    if (s >= eg.nsct) return;
    interleave_sc<IL_P>(eg, pm, off, cnt, s);
}

struct InterleaveArgs {
    EnsGeom eg;
    PartSplit pm;
    const int64_t* off;
    const int32_t* cnt;
};

// batched form, last kernel of a batch step: ensemble j is reordered when its re-sort count
// base[j] + (steps done in this pass) is a multiple of period; nothing after a halt
template <int IL_P>
__global__ void __launch_bounds__(256) k_interleave_split_batch(const InterleaveArgs* __restrict__ jobs,
                                                                const int2* __restrict__ items, long nitems,
                                                                const int* __restrict__ halt,
                                                                const int64_t* __restrict__ base, int period)
{
    if (halt[PK_HALT_STOP]) return;
    for (long w = blockIdx.x; w < nitems; w += gridDim.x) {
        __syncthreads();
        const int2 it = items[w];
        if ((base[it.x] + halt[PK_HALT_DONE]) % period != 0) continue;
        const InterleaveArgs& a = jobs[it.x];
        interleave_sc<IL_P>(a.eg, a.pm, a.off, a.cnt, it.y);
    }
}

int pk_ens_interleave_launch(pk_ens* e)
{
    if (e->posmode != PK_POS_SPLIT || e->np == 0) return PK_OK;
    EnsGeom eg = ens_geom(e);
    pk_ktimer tm(e->ctx, "interleave");
    hipLaunchKernelGGL(k_interleave_split<8>, dim3((unsigned)(PK_XCD * (long)eg.xcd_chunk)), dim3(256), 0,
                       e->ctx->stream, eg, part_split(e, e->main), e->off, e->cnt);
    if (il_big(e->maxcap))
        hipLaunchKernelGGL(k_interleave_split<16>, dim3((unsigned)(PK_XCD * (long)eg.xcd_chunk)), dim3(256), 0,
                           e->ctx->stream, eg, part_split(e, e->main), e->off, e->cnt);
    PK_HIP(e->ctx, hipGetLastError());
    return PK_OK;
}

int pk_ens_interleave_f64_launch(pk_ens* e)
{
    if (e->posmode != PK_POS_ABS || e->np == 0) return PK_OK;
    EnsGeom eg = ens_geom(e);
    StageGeom sg = stage_geom(e);
    pk_ktimer tm(e->ctx, "interleave");
    hipLaunchKernelGGL(k_interleave_f64<8>, dim3((unsigned)(PK_XCD * (long)eg.xcd_chunk)), dim3(256), 0,
                       e->ctx->stream, eg, sg, part_f64(e, e->main), e->off, e->cnt);
    if (il_big(e->maxcap))
        hipLaunchKernelGGL(k_interleave_f64<16>, dim3((unsigned)(PK_XCD * (long)eg.xcd_chunk)), dim3(256), 0,
                           e->ctx->stream, eg, sg, part_f64(e, e->main), e->off, e->cnt);
    PK_HIP(e->ctx, hipGetLastError());
    return PK_OK;
}

extern "C" int pk_ens_interleave(pk_ens* e)
{
    if (!e) return PK_E_ARG;
    if (int r = pk_ens_resort(e)) return r;
    return pk_ens_interleave_launch(e);
}

int pk_batch_interleave_prepare(pk_batch* b)
{
    const int n = b->n();
    const int period = b->knob[PK_BATCH_INTERLEAVE] >= 0 ? b->knob[PK_BATCH_INTERLEAVE] : b->ctx->opt[PK_OPT_INTERLEAVE];
    b->ilperiod = b->ens[0]->posmode == PK_POS_SPLIT ? period : 0;
    if (b->ilperiod <= 0) return PK_OK;
    std::vector<InterleaveArgs> jobs(n);
    for (int i = 0; i < n; i++) {
        pk_ens* e = b->ens[i];
        jobs[i] = InterleaveArgs{ens_geom(e), part_split(e, e->main), e->off, e->cnt};
    }
    b->ilbase_host.assign(n, 0);
    int r = pk_batch_upload(b, &b->iljobs, jobs.data(), sizeof(jobs[0]) * n);
    if (r) return r;
    return pk_batch_upload(b, (void**)&b->ilbase, b->ilbase_host.data(), sizeof(int64_t) * n);
}

int pk_batch_interleave_launch(pk_batch* b)
{
    if (b->ilperiod <= 0 || b->nitems == 0) return PK_OK;
    int maxcap = 0;
    for (pk_ens* e : b->ens) maxcap = e->maxcap > maxcap ? e->maxcap : maxcap;
    hipLaunchKernelGGL(k_interleave_split_batch<8>, dim3((unsigned)b->nitems), dim3(256), 0, b->ctx->stream,
                       (const InterleaveArgs*)b->iljobs, b->items, b->nitems, b->halt, b->ilbase, b->ilperiod);
    if (il_big(maxcap))
        hipLaunchKernelGGL(k_interleave_split_batch<16>, dim3((unsigned)b->nitems), dim3(256), 0, b->ctx->stream,
                           (const InterleaveArgs*)b->iljobs, b->items, b->nitems, b->halt, b->ilbase, b->ilperiod);
    PK_HIP(b->ctx, hipGetLastError());
    return PK_OK;
}

int pk_ens_insert_launch(pk_ens* e)
{
    EnsGeom eg = ens_geom(e);
    const long ngroups = (eg.nsct + INSERT_SC - 1) / INSERT_SC;
    const int gchunk = (int)((ngroups + PK_XCD - 1) / PK_XCD);
    const unsigned nb = (unsigned)(PK_XCD * (long)gchunk);
    pk_ktimer tm(e->ctx, "insert");
    if (e->prec == PK_F64)
        hipLaunchKernelGGL((k_insert<uint64_t, 7>), dim3(nb), dim3(64 * INSERT_SC), 0, e->ctx->stream, insert_args<uint64_t>(e),
                           gchunk);
    else
        hipLaunchKernelGGL((k_insert<uint32_t, 8>), dim3(nb), dim3(64 * INSERT_SC), 0, e->ctx->stream, insert_args<uint32_t>(e),
                           gchunk);
    PK_HIP(e->ctx, hipGetLastError());
    return PK_OK;
}

int pk_batch_insert_prepare(pk_batch* b)
{
    const int n = b->n();
    if (b->prec == PK_F64) {
        std::vector<InsertArgs<uint64_t>> jobs(n);
        for (int i = 0; i < n; i++) jobs[i] = insert_args<uint64_t>(b->ens[i]);
        return pk_batch_upload(b, &b->ijobs, jobs.data(), sizeof(jobs[0]) * n);
    }
    std::vector<InsertArgs<uint32_t>> jobs(n);
    for (int i = 0; i < n; i++) jobs[i] = insert_args<uint32_t>(b->ens[i]);
    return pk_batch_upload(b, &b->ijobs, jobs.data(), sizeof(jobs[0]) * n);
}

int pk_batch_insert_launch(pk_batch* b)
{
    if (b->nitems == 0) return PK_OK;
    const unsigned nb = (unsigned)b->nitems;
    if (b->prec == PK_F64)
        hipLaunchKernelGGL((k_insert_batch<uint64_t, 7>), dim3(nb), dim3(64), 0, b->ctx->stream,
                           (const InsertArgs<uint64_t>*)b->ijobs, b->items, b->nitems, b->halt);
    else
        hipLaunchKernelGGL((k_insert_batch<uint32_t, 8>), dim3(nb), dim3(64), 0, b->ctx->stream,
                           (const InsertArgs<uint32_t>*)b->ijobs, b->items, b->nitems, b->halt);
    PK_HIP(b->ctx, hipGetLastError());
    return PK_OK;
}

// Stable counting rebuild.  source 0: current main (stayers) + outbox (movers);
// source 1: staged host upload (stage pointer in e->outbox, n entries in np).
static int rebuild_from(pk_ens* e, long nent, DBuf& keys_in, DBuf& srcs_in, Fields<double>* stage)
{
    pk_ctx* c = e->ctx;
    const long nsct = e->nsct;
    DBuf keys_out, srcs_out, tmp, cnt64, cap64, newoff, runstart, newcnt, maxcap;
    PK_HIP(c, keys_out.alloc(sizeof(uint32_t) * nent));
    PK_HIP(c, srcs_out.alloc(sizeof(uint64_t) * nent));
    unsigned bits = 1;
    while (((uint64_t)1 << bits) < (uint64_t)nsct * 28) bits++;
    size_t bytes = 0;
    PK_HIP(c, rocprim::radix_sort_pairs(nullptr, bytes, keys_in.as<uint32_t>(), keys_out.as<uint32_t>(),
                                        srcs_in.as<uint64_t>(), srcs_out.as<uint64_t>(), (size_t)nent, 0u, bits,
                                        c->stream));
    PK_HIP(c, tmp.alloc(bytes));
    PK_HIP(c, rocprim::radix_sort_pairs(tmp.p, bytes, keys_in.as<uint32_t>(), keys_out.as<uint32_t>(),
                                        srcs_in.as<uint64_t>(), srcs_out.as<uint64_t>(), (size_t)nent, 0u, bits,
                                        c->stream));
    PK_HIP(c, newcnt.alloc(sizeof(int32_t) * nsct));
    PK_HIP(c, cnt64.alloc(sizeof(int64_t) * (nsct + 1)));
    PK_HIP(c, cap64.alloc(sizeof(int64_t) * (nsct + 1)));
    PK_HIP(c, newoff.alloc(sizeof(int64_t) * (nsct + 1)));
    PK_HIP(c, runstart.alloc(sizeof(int64_t) * (nsct + 1)));
    PK_HIP(c, maxcap.alloc(sizeof(int32_t)));
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(nsct, 256)), dim3(256), 0, c->stream, newcnt.as<int32_t>(), nsct, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, c->stream, maxcap.as<int32_t>(), 1L, 0);
    if (nent > 0)
        hipLaunchKernelGGL(k_count_keys, dim3(nblk(nent, 256)), dim3(256), 0, c->stream, nent, keys_out.as<uint32_t>(),
                           newcnt.as<int32_t>());
    hipLaunchKernelGGL(k_caps, dim3(nblk(nsct, 256)), dim3(256), 0, c->stream, nsct, newcnt.as<int32_t>(),
                       e->slack_frac, e->slack_add, cap64.as<int64_t>(), cnt64.as<int64_t>(), maxcap.as<int32_t>());
    PK_HIP(c, hipGetLastError());
    // one extra zero element so that the scans also give the totals at index nsct
    PK_HIP(c, hipMemsetAsync(cap64.as<int64_t>() + nsct, 0, sizeof(int64_t), c->stream));
    PK_HIP(c, hipMemsetAsync(cnt64.as<int64_t>() + nsct, 0, sizeof(int64_t), c->stream));
    int r = exclusive_scan<int64_t>(c, cap64.as<int64_t>(), newoff.as<int64_t>(), nsct + 1);
    if (r) return r;
    r = exclusive_scan<int64_t>(c, cnt64.as<int64_t>(), runstart.as<int64_t>(), nsct + 1);
    if (r) return r;
    int64_t total_cap = 0;
    int32_t mc = 0;
    PK_HIP(c, hipMemcpyAsync(&total_cap, newoff.as<int64_t>() + nsct, sizeof(int64_t), hipMemcpyDeviceToHost,
                             c->stream));
    PK_HIP(c, hipMemcpyAsync(&mc, maxcap.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    PK_HIP(c, hipStreamSynchronize(c->stream));
    if (mc > PK_MAX_SC_PARTICLES)
        return pk_fail(c, PK_E_CAPACITY, "supercell holds more than 65536 particles; use smaller supercells");
    // new storage
    void *nmain = nullptr, *nout = nullptr;
    uint8_t* nodir = nullptr;
    const size_t fbytes = (size_t)total_cap * e->esize();
    const size_t bbytes = fbytes * e->nfield();
    if (hipMalloc(&nmain, bbytes ? bbytes : 16) != hipSuccess || hipMalloc(&nout, bbytes ? bbytes : 16) != hipSuccess ||
        hipMalloc((void**)&nodir, total_cap ? total_cap : 16) != hipSuccess) {
        if (nmain) (void)hipFree(nmain);
        if (nout) (void)hipFree(nout);
        if (nodir) (void)hipFree(nodir);
        (void)hipGetLastError();
        return pk_fail(c, PK_E_ALLOC, "rebuild: hipMalloc of particle buffers failed");
    }
    pk_ens tmpe = *e;
    tmpe.capacity = (size_t)total_cap;
    EnsGeom eg = ens_geom(e);
    if (nent > 0) {
        if (stage) {
            if (e->prec == PK_F64)
                hipLaunchKernelGGL(k_gather_stage_f64, dim3(nblk(nent, 256)), dim3(256), 0, c->stream, nent,
                                   keys_out.as<uint32_t>(), srcs_out.as<uint64_t>(), newoff.as<int64_t>(),
                                   runstart.as<int64_t>(), part_f64(&tmpe, nmain), *stage);
            else
                hipLaunchKernelGGL(k_gather_stage_split, dim3(nblk(nent, 256)), dim3(256), 0, c->stream, eg,
                                   stage_geom(e), nent, keys_out.as<uint32_t>(), srcs_out.as<uint64_t>(),
                                   newoff.as<int64_t>(), runstart.as<int64_t>(), part_split(&tmpe, nmain), *stage);
        } else {
            if (e->prec == PK_F64)
                hipLaunchKernelGGL((k_gather<uint64_t, 7>), dim3(nblk(nent, 256)), dim3(256), 0, c->stream, nent,
                                   keys_out.as<uint32_t>(), srcs_out.as<uint64_t>(), newoff.as<int64_t>(),
                                   runstart.as<int64_t>(), fields_of<uint64_t>(&tmpe, nmain),
                                   fields_of<uint64_t>(e, e->main), fields_of<uint64_t>(e, e->outbox));
            else
                hipLaunchKernelGGL((k_gather<uint32_t, 8>), dim3(nblk(nent, 256)), dim3(256), 0, c->stream, nent,
                                   keys_out.as<uint32_t>(), srcs_out.as<uint64_t>(), newoff.as<int64_t>(),
                                   runstart.as<int64_t>(), fields_of<uint32_t>(&tmpe, nmain),
                                   fields_of<uint32_t>(e, e->main), fields_of<uint32_t>(e, e->outbox));
        }
        PK_HIP(c, hipGetLastError());
    }
    PK_HIP(c, hipMemcpyAsync(e->off, newoff.p, sizeof(int64_t) * (nsct + 1), hipMemcpyDeviceToDevice, c->stream));
    PK_HIP(c, hipMemcpyAsync(e->cnt, newcnt.p, sizeof(int32_t) * nsct, hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_reset_after_rebuild, dim3(nblk(nsct, 256)), dim3(256), 0, c->stream, nsct, e->nstay, e->nmove,
                       e->cnt);
    PK_HIP(c, hipGetLastError());
    PK_HIP(c, hipStreamSynchronize(c->stream));
    if (e->main) (void)hipFree(e->main);
    if (e->outbox) (void)hipFree(e->outbox);
    if (e->odir) (void)hipFree(e->odir);
    e->main = nmain;
    e->outbox = nout;
    e->odir = nodir;
    e->capacity = (size_t)total_cap;
    e->maxcap = mc;
    e->np = nent;
    e->pending_movers = false;
    e->failed = false;
    return PK_OK;
}

// keys of every live particle under another supercell geometry (re-bucketing): particle j of old
// supercell s is entry ebase[s] + j; its key is its new supercell x 28 (class 0), its source the
// old main slot.  The split form keys by its cell, the absolute form by pica's getSupercellIndex
// with the new supercell size (EnsembleSupercells.h:31-33,56).
__global__ void __launch_bounds__(64) k_keys_rebucket(EnsGeom old_eg, EnsGeom new_eg, StageGeom new_sg, int split,
                                                      PartF64 pf, PartSplit ps, const int64_t* off,
                                                      const int32_t* cnt, const int64_t* ebase, uint32_t* keys,
                                                      uint64_t* srcs)
{
    const long s = blockIdx.x;
    if (s >= old_eg.nsct) return;
    const int n = cnt[s];
    for (int j = threadIdx.x; j < n; j += 64) {
        const long p = off[s] + j;
        int sc[3];
        if (split) {
            int c[3];
            cell_coords(ps.cell[p], old_eg, c);
#pragma unroll
            for (int d = 0; d < 3; d++) sc[d] = c[d] / new_eg.ncps[d];
        } else {
            const double x[3] = {pf.x[p], pf.y[p], pf.z[p]};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const int k = (int)floor((x[d] - 0.0) * new_sg.scinv[d]);
                sc[d] = k < 0 ? 0 : (k >= new_eg.nsc[d] ? new_eg.nsc[d] - 1 : k);
            }
        }
        keys[ebase[s] + j] = (uint32_t)((((long)sc[0] * new_eg.nsc[1] + sc[1]) * new_eg.nsc[2] + sc[2]) * 28);
        srcs[ebase[s] + j] = SRC_MAIN | (uint64_t)p;
    }
}

__global__ void k_count64(long nsct, const int32_t* cnt, int64_t* out)
{
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsct) out[s] = cnt[s];
    if (s == nsct) out[s] = 0;
}

// Supercell arrays (off, cnt, nstay, nmove, mvc) of a new supercell shape replace the old ones; the
// particle buffers are left as they are (the caller rebuilds or overwrites them).
int pk_ens_reshape(pk_ens* e, const int ncps[3])
{
    pk_ctx* c = e->ctx;
    long nsct = 1;
    int nsc[3];
    for (int d = 0; d < 3; d++) {
        nsc[d] = e->ncell[d] / ncps[d];
        nsct *= nsc[d];
    }
    int64_t* noff = nullptr;
    int32_t *ncnt = nullptr, *nstay = nullptr, *nmove = nullptr, *nmvc = nullptr;
    const bool ok = hipMalloc((void**)&noff, sizeof(int64_t) * (nsct + 1)) == hipSuccess &&
                    hipMalloc((void**)&ncnt, sizeof(int32_t) * nsct) == hipSuccess &&
                    hipMalloc((void**)&nstay, sizeof(int32_t) * nsct) == hipSuccess &&
                    hipMalloc((void**)&nmove, sizeof(int32_t) * nsct) == hipSuccess &&
                    hipMalloc((void**)&nmvc, sizeof(int32_t) * nsct * 27) == hipSuccess;
    if (!ok) {
        for (void* p : {(void*)noff, (void*)ncnt, (void*)nstay, (void*)nmove, (void*)nmvc})
            if (p) (void)hipFree(p);
        (void)hipGetLastError();
        return pk_fail(c, PK_E_ALLOC, "supercell re-shape: hipMalloc failed");
    }
    PK_HIP(c, hipStreamSynchronize(c->stream));
    for (void* p : {(void*)e->off, (void*)e->cnt, (void*)e->nstay, (void*)e->nmove, (void*)e->mvc})
        if (p) (void)hipFree(p);
    e->off = noff;
    e->cnt = ncnt;
    e->nstay = nstay;
    e->nmove = nmove;
    e->mvc = nmvc;
    for (int d = 0; d < 3; d++) {
        e->ncps[d] = ncps[d];
        e->nsc[d] = nsc[d];
    }
    e->nsct = nsct;
    return PK_OK;
}

// Re-bucket every particle into supercells of ncps cells (the particle data move, the storage is
// rebuilt).  On failure after the geometry changed the ensemble is marked unusable.
extern "C" int pk_ens_set_supercells(pk_ens* e, const int ncps[3])
{
    if (!e || !ncps) return PK_E_ARG;
    pk_ctx* c = e->ctx;
    long nsct = 1;
    int nsc[3];
    for (int d = 0; d < 3; d++) {
        if (ncps[d] <= 0 || e->ncell[d] % ncps[d] != 0)
            return pk_fail(c, PK_E_ARG, "pk_ens_set_supercells: ncell must be a positive multiple of ncps");
        if (ncps[d] > 12) return pk_fail(c, PK_E_ARG, "pk_ens_set_supercells: supercells are limited to 12 cells per side");
        nsc[d] = e->ncell[d] / ncps[d];
        nsct *= nsc[d];
    }
    if (nsct * 28 >= (1L << 31)) return pk_fail(c, PK_E_ARG, "pk_ens_set_supercells: too many supercells");
    int r = pk_ens_resort(e);
    if (r) return r;
    if (ncps[0] == e->ncps[0] && ncps[1] == e->ncps[1] && ncps[2] == e->ncps[2]) return PK_OK;
    const long onsct = e->nsct;
    const long nent = (long)e->np;
    pk_ens ne = *e;      // the new geometry, for its EnsGeom / StageGeom
    for (int d = 0; d < 3; d++) {
        ne.ncps[d] = ncps[d];
        ne.nsc[d] = nsc[d];
    }
    ne.nsct = nsct;
    DBuf ecount, ebase, keys, srcs;
    PK_HIP(c, ecount.alloc(sizeof(int64_t) * (onsct + 1)));
    PK_HIP(c, ebase.alloc(sizeof(int64_t) * (onsct + 1)));
    PK_HIP(c, keys.alloc(sizeof(uint32_t) * (nent > 0 ? nent : 1)));
    PK_HIP(c, srcs.alloc(sizeof(uint64_t) * (nent > 0 ? nent : 1)));
    hipLaunchKernelGGL(k_count64, dim3(nblk(onsct + 1, 256)), dim3(256), 0, c->stream, onsct, e->cnt,
                       ecount.as<int64_t>());
    if ((r = exclusive_scan<int64_t>(c, ecount.as<int64_t>(), ebase.as<int64_t>(), onsct + 1))) return r;
    PartF64 pf{};
    PartSplit ps{};
    if (e->prec == PK_F64) pf = part_f64(e, e->main);
    else ps = part_split(e, e->main);
    hipLaunchKernelGGL(k_keys_rebucket, dim3((unsigned)onsct), dim3(64), 0, c->stream, ens_geom(e), ens_geom(&ne),
                       stage_geom(&ne), e->posmode == PK_POS_SPLIT ? 1 : 0, pf, ps, e->off, e->cnt,
                       ebase.as<int64_t>(), keys.as<uint32_t>(), srcs.as<uint64_t>());
    PK_HIP(c, hipGetLastError());
    PK_HIP(c, hipStreamSynchronize(c->stream));      // the key kernel has read the old arrays
    if ((r = pk_ens_reshape(e, ncps))) return r;
    pk_ktimer tm(c, "rebuild");
    r = rebuild_from(e, nent, keys, srcs, nullptr);
    if (r) {
        e->failed = true;
        c->err += " (the ensemble is unusable until it is uploaded or loaded again)";
        return r;
    }
    // the rebuild keeps the old slot order inside each new supercell; pair same-cell particles again
    if (e->posmode == PK_POS_SPLIT && c->opt[PK_OPT_INTERLEAVE] > 0) return pk_ens_interleave_launch(e);
    return PK_OK;
}

int pk_ens_rebuild(pk_ens* e, int source)
{
    (void)source;
    pk_ctx* c = e->ctx;
    const long nsct = e->nsct;
    pk_ktimer tm(c, "rebuild");
    DBuf ecount, ebase;
    PK_HIP(c, ecount.alloc(sizeof(int64_t) * (nsct + 1)));
    PK_HIP(c, ebase.alloc(sizeof(int64_t) * (nsct + 1)));
    hipLaunchKernelGGL(k_entries_per_sc, dim3(nblk(nsct, 256)), dim3(256), 0, c->stream, nsct, e->nstay, e->nmove,
                       ecount.as<int64_t>());
    PK_HIP(c, hipMemsetAsync(ecount.as<int64_t>() + nsct, 0, sizeof(int64_t), c->stream));
    int r = exclusive_scan<int64_t>(c, ecount.as<int64_t>(), ebase.as<int64_t>(), nsct + 1);
    if (r) return r;
    int64_t nent = 0;
    PK_HIP(c, hipMemcpyAsync(&nent, ebase.as<int64_t>() + nsct, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    PK_HIP(c, hipStreamSynchronize(c->stream));
    DBuf keys, srcs;
    PK_HIP(c, keys.alloc(sizeof(uint32_t) * nent));
    PK_HIP(c, srcs.alloc(sizeof(uint64_t) * nent));
    EnsGeom eg = ens_geom(e);
    hipLaunchKernelGGL(k_keys_main, dim3((unsigned)nsct), dim3(64), 0, c->stream, eg, e->off, e->nstay, e->nmove,
                       e->odir, ebase.as<int64_t>(), keys.as<uint32_t>(), srcs.as<uint64_t>());
    PK_HIP(c, hipGetLastError());
    return rebuild_from(e, nent, keys, srcs, nullptr);
}

int pk_ens_upload_impl(pk_ens* e, size_t n, const double* const* host)
{
    pk_ctx* c = e->ctx;
    DBuf stage, keys, srcs;
    PK_HIP(c, stage.alloc(sizeof(double) * 7 * n));
    Fields<double> st;
    for (int k = 0; k < 8; k++) st.f[k] = k < 7 ? stage.as<double>() + (size_t)k * n : nullptr;
    for (int k = 0; k < 7; k++)
        PK_HIP(c, hipMemcpyAsync(st.f[k], host[k], sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    PK_HIP(c, keys.alloc(sizeof(uint32_t) * n));
    PK_HIP(c, srcs.alloc(sizeof(uint64_t) * n));
    PK_HIP(c, hipMemsetAsync(e->flags, 0, sizeof(int32_t) * PK_NFLAGS, c->stream));
    if (n > 0)
        hipLaunchKernelGGL(k_keys_stage, dim3(nblk((long)n, 256)), dim3(256), 0, c->stream, ens_geom(e), stage_geom(e),
                           e->posmode == PK_POS_SPLIT ? 1 : 0, st.f[0], st.f[1], st.f[2], (long)n, keys.as<uint32_t>(),
                           srcs.as<uint64_t>(), e->flags);
    PK_HIP(c, hipGetLastError());
    int32_t fl[PK_NFLAGS];
    PK_HIP(c, hipMemcpyAsync(fl, e->flags, sizeof(fl), hipMemcpyDeviceToHost, c->stream));
    PK_HIP(c, hipStreamSynchronize(c->stream));
    if (fl[PK_FLAG_ERR]) {
        PK_HIP(c, hipMemsetAsync(e->flags, 0, sizeof(int32_t) * PK_NFLAGS, c->stream));
        return pk_fail(c, PK_E_ARG, "pk_ens_upload: particle outside the domain [0, ncell*step) or not finite");
    }
    pk_ktimer tm(c, "rebuild");
    return rebuild_from(e, (long)n, keys, srcs, &st);
}

int pk_ens_load_impl(pk_ens* e, int ppc, double uth, double drift, double weight, uint64_t seed, int nmodes,
                     const double* modes)
{
    pk_ctx* c = e->ctx;
    DBuf dmodes;
    if (nmodes > 0) {
        PK_HIP(c, dmodes.alloc(sizeof(double) * 7 * nmodes));
        PK_HIP(c, hipMemcpy(dmodes.p, modes, sizeof(double) * 7 * nmodes, hipMemcpyHostToDevice));
    }
    const int per_sc = e->ncps[0] * e->ncps[1] * e->ncps[2] * ppc;
    const long cap = (long)per_sc + (long)(per_sc * e->slack_frac) + e->slack_add;
    if (cap > PK_MAX_SC_PARTICLES)
        return pk_fail(c, PK_E_CAPACITY, "pk_ens_load_thermal: more than 65536 particles per supercell");
    const size_t capacity = (size_t)cap * (size_t)e->nsct;
    int r = pk_ens_alloc(e, capacity);
    if (r) return r;
    hipLaunchKernelGGL(k_uniform_off, dim3(nblk(e->nsct + 1, 256)), dim3(256), 0, c->stream, e->nsct, (int64_t)cap,
                       e->off);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(e->nsct, 256)), dim3(256), 0, c->stream, e->cnt, e->nsct, per_sc);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(e->nsct, 256)), dim3(256), 0, c->stream, e->nstay, e->nsct, per_sc);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(e->nsct, 256)), dim3(256), 0, c->stream, e->nmove, e->nsct, 0);
    LoadParams lp;
    lp.ppc = ppc;
    lp.uth = uth;
    lp.drift = drift;
    lp.weight = weight;
    lp.seed = seed;
    lp.nmodes = nmodes;
    lp.modes = nmodes > 0 ? dmodes.as<double>() : nullptr;
    for (int d = 0; d < 3; d++) lp.step[d] = e->step[d];
    EnsGeom eg = ens_geom(e);
    pk_ktimer tm(c, "load");
    if (e->prec == PK_F64)
        hipLaunchKernelGGL(k_load_f64, dim3((unsigned)e->nsct), dim3(256), 0, c->stream, eg, lp, e->off, per_sc,
                           part_f64(e, e->main));
    else
        hipLaunchKernelGGL(k_load_split, dim3((unsigned)e->nsct), dim3(256), 0, c->stream, eg, lp, e->off, per_sc,
                           part_split(e, e->main));
    PK_HIP(c, hipGetLastError());
    if (nmodes > 0) PK_HIP(c, hipStreamSynchronize(c->stream));   // before the mode table is freed
    e->maxcap = (int)cap;
    e->np = (int64_t)per_sc * e->nsct;
    e->pending_movers = false;
    e->failed = false;
    return PK_OK;
}

static int runstart_of(pk_ens* e, DBuf& runstart)
{
    pk_ctx* c = e->ctx;
    const long nsct = e->nsct;
    DBuf c64;
    PK_HIP(c, c64.alloc(sizeof(int64_t) * (nsct + 1)));
    PK_HIP(c, runstart.alloc(sizeof(int64_t) * (nsct + 1)));
    DBuf dummy_cap, dummy_max;
    PK_HIP(c, dummy_cap.alloc(sizeof(int64_t) * (nsct + 1)));
    PK_HIP(c, dummy_max.alloc(sizeof(int32_t)));
    hipLaunchKernelGGL(k_caps, dim3(nblk(nsct, 256)), dim3(256), 0, c->stream, nsct, e->cnt, 0.0, 0,
                       dummy_cap.as<int64_t>(), c64.as<int64_t>(), dummy_max.as<int32_t>());
    PK_HIP(c, hipMemsetAsync(c64.as<int64_t>() + nsct, 0, sizeof(int64_t), c->stream));
    return exclusive_scan<int64_t>(c, c64.as<int64_t>(), runstart.as<int64_t>(), nsct + 1);
}

int pk_ens_download_impl(pk_ens* e, double* const* host, int64_t* off)
{
    pk_ctx* c = e->ctx;
    DBuf runstart, out;
    int r = runstart_of(e, runstart);
    if (r) return r;
    const long n = (long)e->np;
    PK_HIP(c, out.alloc(sizeof(double) * 7 * (n > 0 ? n : 1)));
    Fields<double> of;
    for (int k = 0; k < 8; k++) of.f[k] = k < 7 ? out.as<double>() + (size_t)k * n : nullptr;
    EnsGeom eg = ens_geom(e);
    if (e->prec == PK_F64)
        hipLaunchKernelGGL(k_compact_f64, dim3((unsigned)e->nsct), dim3(256), 0, c->stream, eg, part_f64(e, e->main),
                           e->off, e->cnt, runstart.as<int64_t>(), of);
    else
        hipLaunchKernelGGL(k_compact_split, dim3((unsigned)e->nsct), dim3(256), 0, c->stream, eg,
                           part_split(e, e->main), e->off, e->cnt, runstart.as<int64_t>(), of);
    PK_HIP(c, hipGetLastError());
    for (int k = 0; k < 7; k++)
        if (host[k]) PK_HIP(c, hipMemcpyAsync(host[k], of.f[k], sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (off)
        PK_HIP(c, hipMemcpyAsync(off, runstart.p, sizeof(int64_t) * (e->nsct + 1), hipMemcpyDeviceToHost, c->stream));
    PK_HIP(c, hipStreamSynchronize(c->stream));
    return PK_OK;
}

int pk_ens_cells_impl(pk_ens* e, int32_t* counts, int32_t* cells)
{
    pk_ctx* c = e->ctx;
    DBuf runstart, dcounts, dcells;
    int r = runstart_of(e, runstart);
    if (r) return r;
    const long ncells = (long)e->ncell[0] * e->ncell[1] * e->ncell[2];
    const long n = (long)e->np;
    PK_HIP(c, dcounts.alloc(sizeof(int32_t) * ncells));
    PK_HIP(c, dcells.alloc(sizeof(int32_t) * (n > 0 ? n : 1)));
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(ncells, 256)), dim3(256), 0, c->stream, dcounts.as<int32_t>(), ncells, 0);
    Fields<double> pf;
    for (int k = 0; k < 8; k++) pf.f[k] = (e->prec == PK_F64 && k < 7) ? (double*)e->field(e->main, k) : nullptr;
    const int32_t* scell = e->prec == PK_F64 ? nullptr : (const int32_t*)e->field(e->main, 0);
    hipLaunchKernelGGL(k_cells, dim3((unsigned)e->nsct), dim3(256), 0, c->stream, ens_geom(e), stage_geom(e),
                       e->prec == PK_F64 ? 0 : 1, pf, scell, e->off, e->cnt, runstart.as<int64_t>(),
                       dcounts.as<int32_t>(), dcells.as<int32_t>());
    PK_HIP(c, hipGetLastError());
    if (counts)
        PK_HIP(c, hipMemcpyAsync(counts, dcounts.p, sizeof(int32_t) * ncells, hipMemcpyDeviceToHost, c->stream));
    if (cells) PK_HIP(c, hipMemcpyAsync(cells, dcells.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    PK_HIP(c, hipStreamSynchronize(c->stream));
    return PK_OK;
}
