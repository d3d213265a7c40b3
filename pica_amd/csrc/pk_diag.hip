// pk_diag.hip -- FP64 diagnostics on the device: rho deposition, field and kinetic energy,
// total charge and the Gauss-law residual (no pica counterpart; definitions follow the
// oracle's or_deposit_rho / or_energy and the Yee-grid divergence at cell centres).
// All sums are two-pass (per-block partials, then one fixed-order pass): deterministic.
#include "pk_internal.h"

namespace {

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }
constexpr int RB = 256;             // reduction block
constexpr int RGRID = 1024;         // reduction blocks for grid sums
constexpr int NR = 7;               // grid reduction slots

// rho of one supercell's particles into an LDS tile (cell-centre nodes, base m+1, n+2 wide),
// flushed x 1/V with FP64 atomics.  Shape weights as the oracle's or_deposit_rho.
template <int SHAPE, bool SPLIT>
__global__ void __launch_bounds__(256) k_rho(EnsGeom eg, PartF64 pf, PartSplit ps, const int64_t* off,
                                             const int32_t* cnt, double q, double* rho, int s1, int s2)
{
    extern __shared__ double tile[];
    const long s = pk_xcd_supercell(blockIdx.x, eg.xcd_chunk);
    if (s >= eg.nsct) return;
    const int sidx[3] = {(int)(s / ((long)eg.nsc[1] * eg.nsc[2])), (int)((s / eg.nsc[2]) % eg.nsc[1]),
                         (int)(s % eg.nsc[2])};
    int ts[3], base[3];
    for (int d = 0; d < 3; d++) {
        ts[d] = eg.ncps[d] + 2;
        base[d] = sidx[d] * eg.ncps[d] + 1;      // array index of the tile origin (ghost width 2)
    }
    const int tn = ts[0] * ts[1] * ts[2];
    for (int t = threadIdx.x; t < tn; t += blockDim.x) tile[t] = 0.0;
    __syncthreads();
    const int n = cnt[s];
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
        const long p = off[s] + l;
        double xi[3], w;
        if (SPLIT) {
            const int c = ps.cell[p];
            int cc[3];
            cell_coords(c, eg, cc);
            const float o[3] = {ps.ox[p], ps.oy[p], ps.oz[p]};
            for (int d = 0; d < 3; d++) xi[d] = (double)cc[d] + (double)o[d] + 1.5;   // node a at (a - 1.5) h
            w = ps.w[p];
        } else {
            const double x[3] = {pf.x[p], pf.y[p], pf.z[p]};
            for (int d = 0; d < 3; d++) xi[d] = x[d] * (1.0 / eg.step[d]) - (-2.0 + 0.5);
            w = pf.w[p];
        }
        int b[3];
        double wt[3][3];
        const int K = SHAPE == PK_CIC ? 2 : 3;
        bool ok = true;
        for (int d = 0; d < 3; d++) {
            if (SHAPE == PK_CIC) {
                const int i0 = (int)floor(xi[d]);
                const double f = xi[d] - (double)i0;
                b[d] = i0;
                wt[d][0] = 1.0 - f; wt[d][1] = f; wt[d][2] = 0.0;
            } else {
                const int i0 = (int)floor(xi[d] + 0.5);
                const double dd = xi[d] - (double)i0;
                const double a0 = 0.5 - dd, a1 = 0.5 + dd;
                b[d] = i0 - 1;
                wt[d][0] = 0.5 * a0 * a0; wt[d][1] = 0.75 - dd * dd; wt[d][2] = 0.5 * a1 * a1;
            }
            b[d] -= base[d];
            if (b[d] < 0 || b[d] + K > ts[d]) ok = false;
        }
        if (!ok) continue;
        const double qw = q * w;
        for (int a = 0; a < K; a++)
            for (int bb = 0; bb < K; bb++)
                for (int c = 0; c < K; c++)
                    atomicAdd(&tile[((b[0] + a) * ts[1] + (b[1] + bb)) * ts[2] + (b[2] + c)],
                              wt[0][a] * wt[1][bb] * wt[2][c] * qw);
    }
    __syncthreads();
    const double norm = 1.0 / (eg.step[0] * eg.step[1] * eg.step[2]);
    for (int t = threadIdx.x; t < tn; t += blockDim.x) {
        const double v = tile[t];
        if (v == 0.0) continue;
        const int i = t / (ts[1] * ts[2]), j = (t / ts[2]) % ts[1], k = t % ts[2];
        atomicAdd(&rho[((long)(base[0] + i) * s1 + (base[1] + j)) * s2 + (base[2] + k)], v * norm);
    }
}

__device__ __forceinline__ void block_sum(double* sh, double v, double* out)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
    __syncthreads();
}

__device__ __forceinline__ void block_max(double* sh, double v, double* out)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
    __syncthreads();
}

// grid reductions over physical cells: [sum E^2+B^2, sum rho, max|r|, max|r - ref|, max|4 pi rho|,
// max|c|, max|div J|] with r = div E - 4 pi rho and c = (rho - rho_prev) * inv_dt + div J at the cell
// centre of array cell (i, j, k) (c only when rho_prev is given).
template <class T>
__global__ void __launch_bounds__(RB) k_grid_reduce(GridView<T> g, const double* rho, double* ref, int set_ref,
                                                    int has_ref, const double* rho_prev, double inv_dt,
                                                    double* partial)
{
    __shared__ double sh[RB];
    const int G = g.G;
    const int N0 = g.s0 - 2 * G, N1 = g.s1 - 2 * G, N2 = g.s2 - 2 * G;
    const long n = (long)N0 * N1 * N2;
    double e2 = 0.0, sr = 0.0, mr = 0.0, mdr = 0.0, mrho = 0.0, mc = 0.0, mdj = 0.0;
    const long s12 = (long)g.s1 * g.s2;
    for (long t = (long)blockIdx.x * RB + threadIdx.x; t < n; t += (long)gridDim.x * RB) {
        const int i = G + (int)(t / ((long)N1 * N2)), j = G + (int)((t / N2) % N1), k = G + (int)(t % N2);
        const long c = (long)i * s12 + (long)j * g.s2 + k;
        double s = 0.0;
#pragma unroll
        for (int comp = PK_EX; comp <= PK_BZ; comp++) {
            const double v = (double)g.f[comp][c];
            s += v * v;
        }
        e2 += s;
        if (!rho) continue;    // field energy only
        const double rh = rho[c];
        sr += rh;
        const double divE = ((double)g.f[PK_EX][c + s12] - (double)g.f[PK_EX][c]) / g.step[0] +
                            ((double)g.f[PK_EY][c + g.s2] - (double)g.f[PK_EY][c]) / g.step[1] +
                            ((double)g.f[PK_EZ][c + 1] - (double)g.f[PK_EZ][c]) / g.step[2];
        const double fr = 4.0 * PK_PI * rh;
        const double r = divE - fr;
        mr = fmax(mr, fabs(r));
        mrho = fmax(mrho, fabs(fr));
        if (set_ref) ref[t] = r;
        if (has_ref) mdr = fmax(mdr, fabs(r - ref[t]));
        if (rho_prev) {
            // div J with the same forward differences under which Esirkepov deposition conserves charge
            const double divJ = ((double)g.f[PK_JX][c + s12] - (double)g.f[PK_JX][c]) / g.step[0] +
                                ((double)g.f[PK_JY][c + g.s2] - (double)g.f[PK_JY][c]) / g.step[1] +
                                ((double)g.f[PK_JZ][c + 1] - (double)g.f[PK_JZ][c]) / g.step[2];
            mc = fmax(mc, fabs((rh - rho_prev[c]) * inv_dt + divJ));
            mdj = fmax(mdj, fabs(divJ));
        }
    }
    block_sum(sh, e2, &partial[blockIdx.x * NR + 0]);
    block_sum(sh, sr, &partial[blockIdx.x * NR + 1]);
    block_max(sh, mr, &partial[blockIdx.x * NR + 2]);
    block_max(sh, mdr, &partial[blockIdx.x * NR + 3]);
    block_max(sh, mrho, &partial[blockIdx.x * NR + 4]);
    block_max(sh, mc, &partial[blockIdx.x * NR + 5]);
    block_max(sh, mdj, &partial[blockIdx.x * NR + 6]);
}

// sum_p w * u^2 / (1 + gamma), per supercell partial
template <bool SPLIT>
__global__ void __launch_bounds__(RB) k_kinetic(EnsGeom eg, PartF64 pf, PartSplit ps, const int64_t* off,
                                                const int32_t* cnt, double* partial)
{
    __shared__ double sh[RB];
    const long s = blockIdx.x;
    const int n = cnt[s];
    double acc = 0.0;
    for (int l = threadIdx.x; l < n; l += RB) {
        const long p = off[s] + l;
        double u[3], w;
        if (SPLIT) {
            u[0] = ps.ux[p]; u[1] = ps.uy[p]; u[2] = ps.uz[p];
            w = ps.w[p];
        } else {
            u[0] = pf.ux[p]; u[1] = pf.uy[p]; u[2] = pf.uz[p];
            w = pf.w[p];
        }
        const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        acc += w * (u2 / (1.0 + sqrt(1.0 + u2)));
    }
    block_sum(sh, acc, &partial[s]);
}

// fixed-order final pass over the partials
__global__ void k_final(const double* gp, int ngp, const double* kp, long nkp, double* out)
{
    __shared__ double sh[RB];
    double a[NR] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < ngp; b += RB) {
        a[0] += gp[b * NR + 0];
        a[1] += gp[b * NR + 1];
        for (int k = 2; k < NR; k++) a[k] = fmax(a[k], gp[b * NR + k]);
    }
    double kin = 0.0;
    for (long b = threadIdx.x; b < nkp; b += RB) kin += kp[b];
    block_sum(sh, a[0], &out[0]);
    block_sum(sh, a[1], &out[1]);
    block_max(sh, a[2], &out[2]);
    block_max(sh, a[3], &out[3]);
    block_max(sh, a[4], &out[4]);
    block_sum(sh, kin, &out[5]);
    block_max(sh, a[5], &out[6]);
    block_max(sh, a[6], &out[7]);
}

__global__ void k_zero_d(double* p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

template <class T>
__global__ void k_copy_to_grid(const double* src, T* dst, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (T)src[i];
}

// Scratch of pk_diag_energies_static: a file-scope device array each GPU holds once for the
// life of the process (no allocation at call time).  Layout: RGRID * NR grid partials, then
// PK_STATIC_MAX_SUPERCELLS kinetic partials, then 16 results.
constexpr long PK_STATIC_MAX_SUPERCELLS = 1L << 20;
constexpr long STATIC_SCRATCH_N = (long)RGRID * NR + PK_STATIC_MAX_SUPERCELLS + 16;
__device__ __align__(256) double scratch[STATIC_SCRATCH_N];

// the device address of `scratch` per device, taken once with that device current
constexpr int MAX_DEVICES = 64;
double* g_static_scratch[MAX_DEVICES] = {};

int ensure_scratch(pk_grid* g, pk_ens* e)
{
    const size_t need = g->n + (size_t)RGRID * NR + (size_t)e->nsct + 16;   // 16: final results (8 used)
    if (g->scratch && g->scratch_n >= need) return PK_OK;
    if (g->scratch) (void)hipFree(g->scratch);
    g->scratch = nullptr;
    if (hipMalloc((void**)&g->scratch, need * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        g->scratch_n = 0;
        return pk_fail(g->ctx, PK_E_ALLOC, "diagnostics: scratch allocation failed");
    }
    g->scratch_n = need;
    return PK_OK;
}

template <int SHAPE>
void launch_rho(pk_ens* e, pk_grid* g, double* rho)
{
    EnsGeom eg = ens_geom(e);
    const double q = e->ctx->charge[e->type];
    const size_t lds = sizeof(double) * (e->ncps[0] + 2) * (e->ncps[1] + 2) * (e->ncps[2] + 2);
    const unsigned nb = (unsigned)(PK_XCD * (long)eg.xcd_chunk);
    PartF64 pf{};
    PartSplit ps{};
    if (e->prec == PK_F64) {
        pf = part_f64(e, e->main);
        hipLaunchKernelGGL((k_rho<SHAPE, false>), dim3(nb), dim3(256), lds, e->ctx->stream, eg, pf, ps, e->off, e->cnt,
                           q, rho, g->s[1], g->s[2]);
    } else {
        ps = part_split(e, e->main);
        hipLaunchKernelGGL((k_rho<SHAPE, true>), dim3(nb), dim3(256), lds, e->ctx->stream, eg, pf, ps, e->off, e->cnt,
                           q, rho, g->s[1], g->s[2]);
    }
}

}  // namespace

// Called by pk_ctx_create with `device` current: resolves this device's copy of `scratch`.
int pk_diag_static_init(pk_ctx* c, int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return pk_fail(c, PK_E_UNSUPPORTED, "pk_ctx_create: device index beyond the static scratch table");
    if (g_static_scratch[device]) return PK_OK;
    void* ptr = nullptr;
    const hipError_t er = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (er != hipSuccess || !ptr)
        return pk_fail(c, PK_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(er));
    g_static_scratch[device] = (double*)ptr;
    return PK_OK;
}

// FP64 rho (folded, filled) into g->scratch[0 .. n)
int pk_rho_fp64(pk_ens* e, pk_grid* g, int shape)
{
    pk_ctx* c = e->ctx;
    int r = ensure_scratch(g, e);
    if (r) return r;
    double* rho = g->scratch;
    hipLaunchKernelGGL(k_zero_d, dim3(nblk((long)g->n, 256)), dim3(256), 0, c->stream, rho, (long)g->n);
    pk_ktimer tm(c, "rho");
    if (shape == PK_CIC) launch_rho<PK_CIC>(e, g, rho);
    else launch_rho<PK_TSC>(e, g, rho);
    PK_HIP(c, hipGetLastError());
    // periodic fold + fill of the FP64 rho through a grid of the same geometry
    pk_grid tmp = *g;
    tmp.prec = PK_F64;
    tmp.pitch = 0;
    tmp.data = rho;       // comp(c) = data + c * 0: every component aliases rho; only PK_RHO is touched
    tmp.scratch = nullptr;
    tmp.gauss_ref = nullptr;
    r = pk_field_ghost(&tmp, PK_FOLD_RHO);
    tmp.data = nullptr;
    return r;
}

int pk_rho_impl(pk_ens* e, pk_grid* g, int shape)
{
    int r = pk_rho_fp64(e, g, shape);
    if (r) return r;
    const long n = (long)g->n;
    if (g->prec == PK_F64)
        hipLaunchKernelGGL(k_copy_to_grid<double>, dim3(nblk(n, 256)), dim3(256), 0, e->ctx->stream, g->scratch,
                           (double*)g->comp(PK_RHO), n);
    else
        hipLaunchKernelGGL(k_copy_to_grid<float>, dim3(nblk(n, 256)), dim3(256), 0, e->ctx->stream, g->scratch,
                           (float*)g->comp(PK_RHO), n);
    PK_HIP(e->ctx, hipGetLastError());
    PK_HIP(e->ctx, hipStreamSynchronize(e->ctx->stream));
    return PK_OK;
}

// Field and kinetic energy with pk_diag's kernels and reduction order (bit-identical to its
// PK_DIAG_FIELD_ENERGY / PK_DIAG_KINETIC_ENERGY), on the device's static scratch.  Calls on one
// GPU share that scratch, so calls that overlap on one GPU race; every value a call reads was
// written by that call (the partial and result slots are written in full), so nothing needs
// clearing between calls.
int pk_diag_energies_static_impl(pk_ens* e, pk_grid* g, double out[2])
{
    pk_ctx* c = e->ctx;
    if (e->nsct > PK_STATIC_MAX_SUPERCELLS)
        return pk_fail(c, PK_E_CAPACITY,
                       "pk_diag_energies_static: more than 1,048,576 supercells (the static scratch holds one "
                       "kinetic partial per supercell); use pk_diag");
    double* base = (c->device >= 0 && c->device < MAX_DEVICES) ? g_static_scratch[c->device] : nullptr;
    if (!base) return pk_fail(c, PK_E_STATE, "pk_diag_energies_static: static scratch not resolved for this device");
    double* gp = base;
    double* kp = gp + (size_t)RGRID * NR;
    double* res = kp + PK_STATIC_MAX_SUPERCELLS;
    if (g->prec == PK_F64)
        hipLaunchKernelGGL(k_grid_reduce<double>, dim3(RGRID), dim3(RB), 0, c->stream, grid_view<double>(g), nullptr,
                           nullptr, 0, 0, nullptr, 0.0, gp);
    else
        hipLaunchKernelGGL(k_grid_reduce<float>, dim3(RGRID), dim3(RB), 0, c->stream, grid_view<float>(g), nullptr,
                           nullptr, 0, 0, nullptr, 0.0, gp);
    EnsGeom eg = ens_geom(e);
    PartF64 pf{};
    PartSplit ps{};
    if (e->nsct > 0) {
        if (e->prec == PK_F64) {
            pf = part_f64(e, e->main);
            hipLaunchKernelGGL(k_kinetic<false>, dim3((unsigned)e->nsct), dim3(RB), 0, c->stream, eg, pf, ps, e->off,
                               e->cnt, kp);
        } else {
            ps = part_split(e, e->main);
            hipLaunchKernelGGL(k_kinetic<true>, dim3((unsigned)e->nsct), dim3(RB), 0, c->stream, eg, pf, ps, e->off,
                               e->cnt, kp);
        }
    }
    hipLaunchKernelGGL(k_final, dim3(1), dim3(RB), 0, c->stream, gp, RGRID, kp, e->nsct, res);
    PK_HIP(c, hipGetLastError());
    double h[8];
    PK_HIP(c, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    PK_HIP(c, hipStreamSynchronize(c->stream));
    const double V = g->step[0] * g->step[1] * g->step[2];
    out[0] = h[0] * V / (8.0 * PK_PI);
    out[1] = h[5] * c->mass[e->type] * PK_C * PK_C;
    return PK_OK;
}

int pk_diag_impl(pk_ens* e, pk_grid* g, int shape, double* out, bool set_ref)
{
    pk_ctx* c = e->ctx;
    int r = pk_rho_fp64(e, g, shape);
    if (r) return r;
    const long nphys = (long)(g->s[0] - 2 * g->ghost) * (g->s[1] - 2 * g->ghost) * (g->s[2] - 2 * g->ghost);
    if (set_ref && !g->gauss_ref) {
        if (hipMalloc((void**)&g->gauss_ref, sizeof(double) * nphys) != hipSuccess) {
            (void)hipGetLastError();
            return pk_fail(c, PK_E_ALLOC, "pk_diag_set_reference: allocation failed");
        }
    }
    double* gp = g->scratch + g->n;
    double* kp = gp + (size_t)RGRID * NR;
    double* res = kp + e->nsct;
    pk_ktimer tm(c, "diag");
    const double* rp = g->last_dt > 0.0 ? g->rho_prev : nullptr;
    const double inv_dt = g->last_dt > 0.0 ? 1.0 / g->last_dt : 0.0;
    if (g->prec == PK_F64)
        hipLaunchKernelGGL(k_grid_reduce<double>, dim3(RGRID), dim3(RB), 0, c->stream, grid_view<double>(g), g->scratch,
                           g->gauss_ref, set_ref ? 1 : 0, (!set_ref && g->gauss_ref) ? 1 : 0, rp, inv_dt, gp);
    else
        hipLaunchKernelGGL(k_grid_reduce<float>, dim3(RGRID), dim3(RB), 0, c->stream, grid_view<float>(g), g->scratch,
                           g->gauss_ref, set_ref ? 1 : 0, (!set_ref && g->gauss_ref) ? 1 : 0, rp, inv_dt, gp);
    EnsGeom eg = ens_geom(e);
    PartF64 pf{};
    PartSplit ps{};
    if (e->prec == PK_F64) {
        pf = part_f64(e, e->main);
        hipLaunchKernelGGL(k_kinetic<false>, dim3((unsigned)e->nsct), dim3(RB), 0, c->stream, eg, pf, ps, e->off,
                           e->cnt, kp);
    } else {
        ps = part_split(e, e->main);
        hipLaunchKernelGGL(k_kinetic<true>, dim3((unsigned)e->nsct), dim3(RB), 0, c->stream, eg, pf, ps, e->off,
                           e->cnt, kp);
    }
    hipLaunchKernelGGL(k_final, dim3(1), dim3(RB), 0, c->stream, gp, RGRID, kp, e->nsct, res);
    PK_HIP(c, hipGetLastError());
    double h[8];
    PK_HIP(c, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    PK_HIP(c, hipStreamSynchronize(c->stream));
    const double V = g->step[0] * g->step[1] * g->step[2];
    const double m = c->mass[e->type];
    out[PK_DIAG_FIELD_ENERGY] = h[0] * V / (8.0 * PK_PI);
    out[PK_DIAG_KINETIC_ENERGY] = h[5] * m * PK_C * PK_C;
    out[PK_DIAG_CHARGE] = h[1] * V;
    out[PK_DIAG_GAUSS_MAX] = h[2];
    out[PK_DIAG_GAUSS_DRIFT] = h[3];
    out[PK_DIAG_RHO_MAX] = h[4];
    out[PK_DIAG_CONTINUITY_MAX] = h[6];
    out[PK_DIAG_DIVJ_MAX] = h[7];
    return PK_OK;
}
