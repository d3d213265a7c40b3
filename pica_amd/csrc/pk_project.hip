// pk_project.hip -- charge-continuity projection of J, and the PIC step with the learned
// denoiser (no counterpart in pica).
//
// After J has been changed by anything that ignores charge conservation (the denoiser), the
// projection subtracts a discrete gradient: J <- J - G phi with  L phi = D J + (rho1 - rho0)/dt
// (or, when the current before the change was Esirkepov's, L phi = D J - S D J_esirkepov with S
// the denoiser's charge filter, a fixed radial filter with gain g(|k|), g(0) = 1, applied in
// Fourier space: then D J = -S (rho1 - rho0)/dt, so continuity and Gauss's law hold exactly for
// the filtered charge S rho; without a filter S is the identity, the particles' own charge),
// where D is the forward-difference divergence at the cell-centre nodes (the stencil under which
// Esirkepov deposition conserves charge exactly), G = -D^T the backward-difference gradient onto
// the J faces, and L = D G the 7-point periodic Laplacian.  Then D J = -(rho1 - rho0)/dt.  L is
// diagonal in the discrete Fourier basis, eigenvalues -sum_d 4 sin^2(pi k_d / N_d) / h_d^2, so
// phi comes from three passes of batched radix-2 FFTs (16 pencils per workgroup, in LDS), in the
// grid's precision.
#include <cmath>

#include "pk_internal.h"

int pk_dn_apply(pk_dn* d, pk_grid* g);
void pk_dn_charge_gain(const pk_dn* d, const double** gain, int* n, double* kmax);

namespace {

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

struct Phys {
    int n[3];
    long nv;
};

// complex value of the projection's work volume: FP64 for FP64 grids, FP32 for FP32 grids (the
// correction is stored in the grid's precision either way)
template <class R>
struct alignas(2 * sizeof(R)) Cx {
    R x, y;
};

__device__ __forceinline__ long grid_index(const int s1, const int s2, int G, int i, int j, int k)
{
    return ((long)(i + G) * s1 + (j + G)) * s2 + (k + G);
}

// w[v] = (D J + (rho1 - rho0) / dt, 0) at the physical nodes; J ghosts must be filled
template <class T>
__global__ void k_residual(GridView<T> g, Phys P, const double* __restrict__ rho1, const double* __restrict__ rho0,
                           double inv_dt, Cx<T>* __restrict__ w)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.nv) return;
    const int i = (int)(v / ((long)P.n[1] * P.n[2])), j = (int)((v / P.n[2]) % P.n[1]), k = (int)(v % P.n[2]);
    const long c = grid_index(g.s1, g.s2, g.G, i, j, k);
    const long s12 = (long)g.s1 * g.s2;
    double r = ((double)g.f[PK_JX][c + s12] - (double)g.f[PK_JX][c]) / g.step[0];
    r += ((double)g.f[PK_JY][c + g.s2] - (double)g.f[PK_JY][c]) / g.step[1];
    r += ((double)g.f[PK_JZ][c + 1] - (double)g.f[PK_JZ][c]) / g.step[2];
    r += (rho1[c] - rho0[c]) * inv_dt;
    w[v] = Cx<T>{(T)r, (T)0};
}

// physical node v -> (i, j, k); 32-bit division when the volume allows it
__device__ __forceinline__ void phys_ijk(const Phys& P, long v, int& i, int& j, int& k)
{
    if (P.nv <= 0xffffffffl) {
        const unsigned n12 = (unsigned)P.n[1] * (unsigned)P.n[2], u = (unsigned)v;
        i = (int)(u / n12);
        const unsigned r = u - (unsigned)i * n12;
        j = (int)(r / (unsigned)P.n[2]);
        k = (int)(r - (unsigned)j * (unsigned)P.n[2]);
    } else {
        i = (int)(v / ((long)P.n[1] * P.n[2]));
        j = (int)((v / P.n[2]) % P.n[1]);
        k = (int)(v % P.n[2]);
    }
}

// D J at physical node v (pk_div_j)
template <class T>
__device__ __forceinline__ double div_j(const GridView<T>& g, const Phys& P, long v)
{
    int i, j, k;
    phys_ijk(P, v, i, j, k);
    return pk_div_j(g, P.n, i, j, k);
}

// w[v] = (D J - d0[v], 0) = (D (J - Jref), 0): the residual when the current Jref, whose
// divergence d0 was saved, satisfies continuity exactly, as the Esirkepov current does
// pack_ref: the imaginary part carries d0 itself, so that one complex transform gives the
// transforms of both real fields (the charge filter needs both)
template <class T>
__global__ void k_residual_ref(GridView<T> g, Phys P, const double* __restrict__ d0, Cx<T>* __restrict__ w,
                               int pack_ref)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.nv) return;
    w[v] = Cx<T>{(T)(div_j(g, P, v) - d0[v]), pack_ref ? (T)d0[v] : (T)0};
}

// in-place radix-2 FFTs of pb pencils per workgroup: pencil p holds w[base(p) + j * stride],
// j < n, base(p) = (p / inner) * ostride + p % inner (n, pb, inner powers of two, passed as
// logarithms); tw[k] = exp(-2 pi i k / n), k < n / 2;
// inverse = conjugate twiddles, no 1/n.  `along` = 1 when a pencil is contiguous (z), else the
// pencils of a workgroup are adjacent in memory and are read across pencils (coalesced).
// solve != null: each value is first divided by lambda(k) nv, the Laplacian's eigenvalue times
// the volume (lambda = solve[i] + solve[n0 + j] + solve[n0 + n1 + k]; the mean mode is zeroed) --
// the Poisson solve folded into the first inverse pass.
// res.d0 != null: the pass reads no volume but forms the residual (D J - d0, pack ? d0 : 0), as
// k_residual_ref writes it, from the grid (the first forward pass)
template <class R>
struct ResidualSrc {
    GridView<R> g;
    const double* d0 = nullptr;
    int pack = 0;
};
template <class R>
__global__ void __launch_bounds__(256) k_fft(Cx<R>* __restrict__ w, int logn, long npencil, int logpb, int loginner,
                                             long ostride, long stride, int along, const double2* __restrict__ tw,
                                             int inverse, const double* __restrict__ solve, Phys P, ResidualSrc<R> res)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = 1 << logn, pb = 1 << logpb;
    Cx<R>* twl = (Cx<R>*)smem;        // n / 2 twiddles, conjugated for the inverse
    // pb rows of n + n / 8 + 1 values: element i of a row sits at i + i / 8 (one pad value after every
    // eight, so that the bit-reversed writes and the butterflies' strided accesses spread over the
    // banks), and the last pad spreads the rows' starts
    Cx<R>* a = twl + (n >> 1);
    const int row = n + (n >> 3) + 1;
    const auto at = [row](int pl, int i) { return pl * row + i + (i >> 3); };
    for (int t = threadIdx.x; t < (n >> 1); t += blockDim.x) {
        const double2 wd = tw[t];
        twl[t] = Cx<R>{(R)wd.x, inverse ? (R)-wd.y : (R)wd.y};
    }
    const long inner_mask = (1l << loginner) - 1;
    const long p0 = (long)blockIdx.x << logpb;
    const int tot = pb << logn;
    constexpr int CH = 8;             // loads in flight per thread before their LDS writes
    for (int i0 = threadIdx.x; i0 < tot; i0 += CH * 256) {
        Cx<R> v[CH];
#pragma unroll
        for (int k = 0; k < CH; k++) {
            const int idx = i0 + 256 * k;
            const int pl = along ? idx >> logn : idx & (pb - 1), j = along ? idx & (n - 1) : idx >> logpb;
            const long p = p0 + pl;
            if (idx < tot && p < npencil) {
                const long gv = ((p >> loginner) * ostride + (p & inner_mask)) + (long)j * stride;
                if (res.d0) {
                    const double d0 = res.d0[gv];
                    v[k] = Cx<R>{(R)(div_j(res.g, P, gv) - d0), res.pack ? (R)d0 : (R)0};
                } else {
                    v[k] = w[gv];
                }
                if (solve) {
                    int vi, vj, vk;
                    phys_ijk(P, gv, vi, vj, vk);
                    const double lam = solve[vi] + solve[P.n[0] + vj] + solve[P.n[0] + P.n[1] + vk];
                    const R f = gv == 0 ? (R)0 : (R)(1.0 / (lam * (double)P.nv));
                    v[k] = Cx<R>{v[k].x * f, v[k].y * f};
                }
            }
        }
#pragma unroll
        for (int k = 0; k < CH; k++) {
            const int idx = i0 + 256 * k;
            const int pl = along ? idx >> logn : idx & (pb - 1), j = along ? idx & (n - 1) : idx >> logpb;
            if (idx < tot && p0 + pl < npencil) a[at(pl, (int)(__brev((unsigned)j) >> (32 - logn)))] = v[k];
        }
    }
    __syncthreads();
    // radix-2 stages two at a time: a thread takes the four elements i, i + h, i + 2h, i + 3h of a
    // block of 4h and applies stage lh (pairs at distance h) and stage lh + 1 (pairs at distance 2h)
    // in registers, with the radix-2 stages' own operations and twiddles (so the same results), one
    // LDS round trip and one barrier per two stages; an odd last stage runs alone
    const auto mul = [](Cx<R> x, Cx<R> w) { return Cx<R>{x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x}; };
    int lh = 0;
    for (; lh + 1 < logn; lh += 2) {
        const int h = 1 << lh;
        const int nq = pb << (logn - 2);
        for (int t = threadIdx.x; t < nq; t += blockDim.x) {
            const int pl = t >> (logn - 2), tt = t & ((n >> 2) - 1);
            const int grp = tt >> lh, pos = tt & (h - 1);
            const int b0 = (grp << (lh + 2)) + pos;
            const int i0 = at(pl, b0), i1 = at(pl, b0 + h), i2 = at(pl, b0 + 2 * h), i3 = at(pl, b0 + 3 * h);
            const Cx<R> x0 = a[i0], x1 = a[i1], x2 = a[i2], x3 = a[i3];
            const Cx<R> w1 = twl[pos << (logn - 1 - lh)];
            const Cx<R> v1 = mul(x1, w1), v3 = mul(x3, w1);
            const Cx<R> y0{x0.x + v1.x, x0.y + v1.y}, y1{x0.x - v1.x, x0.y - v1.y};
            const Cx<R> y2{x2.x + v3.x, x2.y + v3.y}, y3{x2.x - v3.x, x2.y - v3.y};
            const Cx<R> w2 = twl[pos << (logn - 2 - lh)], w3 = twl[(pos + h) << (logn - 2 - lh)];
            const Cx<R> v2 = mul(y2, w2), u3 = mul(y3, w3);
            a[i0] = Cx<R>{y0.x + v2.x, y0.y + v2.y};
            a[i2] = Cx<R>{y0.x - v2.x, y0.y - v2.y};
            a[i1] = Cx<R>{y1.x + u3.x, y1.y + u3.y};
            a[i3] = Cx<R>{y1.x - u3.x, y1.y - u3.y};
        }
        __syncthreads();
    }
    if (lh < logn) {                                   // half = 2^lh, len = 2 half
        const int half = 1 << lh;
        const int nb = pb << (logn - 1);
        for (int t = threadIdx.x; t < nb; t += blockDim.x) {
            const int pl = t >> (logn - 1), tt = t & ((n >> 1) - 1);
            const int grp = tt >> lh, pos = tt & (half - 1);
            const int b0 = (grp << (lh + 1)) + pos;
            const int i0 = at(pl, b0), i1 = at(pl, b0 + half);
            const Cx<R> u = a[i0], v = mul(a[i1], twl[pos << (logn - 1 - lh)]);
            a[i0] = Cx<R>{u.x + v.x, u.y + v.y};
            a[i1] = Cx<R>{u.x - v.x, u.y - v.y};
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
        const int pl = along ? idx >> logn : idx & (pb - 1), j = along ? idx & (n - 1) : idx >> logpb;
        const long p = p0 + pl;
        if (p >= npencil) continue;
        const long base = (p >> loginner) * ostride + (p & inner_mask);
        w[base + (long)j * stride] = a[at(pl, j)];
    }
}

// J_d[c] -= (phi[c] - phi[c - e_d]) / h_d on the physical faces
template <class T>
__global__ void k_correct(GridView<T> g, Phys P, const Cx<T>* __restrict__ w)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.nv) return;
    int i, j, k;
    phys_ijk(P, v, i, j, k);
    const long c = grid_index(g.s1, g.s2, g.G, i, j, k);
    const double ph = (double)w[v].x;
    const int im = i == 0 ? P.n[0] - 1 : i - 1, jm = j == 0 ? P.n[1] - 1 : j - 1, km = k == 0 ? P.n[2] - 1 : k - 1;
    const double px = (double)w[((long)im * P.n[1] + j) * P.n[2] + k].x;
    const double py = (double)w[((long)i * P.n[1] + jm) * P.n[2] + k].x;
    const double pz = (double)w[((long)i * P.n[1] + j) * P.n[2] + km].x;
    g.f[PK_JX][c] = (T)((double)g.f[PK_JX][c] - (ph - px) / g.step[0]);
    g.f[PK_JY][c] = (T)((double)g.f[PK_JY][c] - (ph - py) / g.step[1]);
    g.f[PK_JZ][c] = (T)((double)g.f[PK_JZ][c] - (ph - pz) / g.step[2]);
}

// Z = transform of a + i b for real volumes a (the residual D J - d0) and b (d0), whose
// transforms are A(k) = (Z(k) + conj Z(-k)) / 2 and B(k) = (Z(k) - conj Z(-k)) / 2i.  Writes
// out = A + (1 - g(|k|)) B, the transform of D J - S d0 with S the radial filter of gain g (linear
// interpolation of nb values over [0, kmax], the last value beyond); |k| in radians per cell,
// k_d = 2 pi f_d / n_d with f_d the signed frequency
template <class R>
__global__ void k_gain_combine(const Cx<R>* __restrict__ z, Cx<R>* __restrict__ out, Phys P,
                               const double* __restrict__ gain, int nb, double kmax)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= P.nv) return;
    const int c[3] = {(int)(v / ((long)P.n[1] * P.n[2])), (int)((v / P.n[2]) % P.n[1]), (int)(v % P.n[2])};
    double k2 = 0.0;
    int m[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int f = c[d] <= P.n[d] / 2 ? c[d] : c[d] - P.n[d];
        const double kd = 2.0 * M_PI * (double)f / (double)P.n[d];
        k2 += kd * kd;
        m[d] = c[d] == 0 ? 0 : P.n[d] - c[d];
    }
    const double x = sqrt(k2) / kmax * (double)(nb - 1);
    double gk;
    if (x >= (double)(nb - 1)) {
        gk = gain[nb - 1];
    } else {
        const int b = (int)x;
        const double t = x - (double)b;
        gk = gain[b] + t * (gain[b + 1] - gain[b]);
    }
    const Cx<R> zp = z[v], zm = z[((long)m[0] * P.n[1] + m[1]) * P.n[2] + m[2]];
    const R ax = (R)0.5 * (zp.x + zm.x), ay = (R)0.5 * (zp.y - zm.y);     // A(k)
    const R bx = (R)0.5 * (zp.y + zm.y), by = (R)0.5 * (zm.x - zp.x);     // B(k)
    const R f = (R)(1.0 - gk);
    out[v] = Cx<R>{ax + f * bx, ay + f * by};
}

__global__ void k_copy_d(const double* __restrict__ a, double* __restrict__ b, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

int log2_exact(int n)
{
    int l = 0;
    while ((1 << l) < n) l++;
    return (1 << l) == n ? l : -1;
}

// work layout in g->proj (doubles): complex volume (2 nv), twiddles (n_d / 2 complex values per
// dimension), eigenvalues (n0 + n1 + n2), then, at 2 nv + 2 (n0 + n1 + n2), the divergence of the
// reference current (nv)
double* jref_of(pk_grid* g, const Phys& P) { return g->proj + 2 * P.nv + 2 * (P.n[0] + P.n[1] + P.n[2]); }

int ensure_proj(pk_grid* g, const Phys& P)
{
    // complex volume, twiddles and eigenvalues, the reference divergence, and a second complex
    // volume for the charge filter
    const size_t need = 5 * (size_t)P.nv + 2 * (size_t)(P.n[0] + P.n[1] + P.n[2]) + 16;
    if (g->proj && g->proj_n >= need) return PK_OK;
    PK_HIP(g->ctx, hipStreamSynchronize(g->ctx->stream));
    if (g->proj) (void)hipFree(g->proj);
    g->proj = nullptr;
    g->proj_n = 0;
    if (hipMalloc((void**)&g->proj, sizeof(double) * need) != hipSuccess) {
        g->proj = nullptr;
        (void)hipGetLastError();
        return pk_fail(g->ctx, PK_E_ALLOC, "projection: allocation failed");
    }
    // twiddles exp(-2 pi i k / n), k < n / 2, and eigenvalues per dimension (host FP64)
    std::vector<double> tab;
    for (int d = 0; d < 3; d++)
        for (int k = 0; k < P.n[d] / 2; k++) {
            const double a = -2.0 * M_PI * k / P.n[d];
            tab.push_back(cos(a));
            tab.push_back(sin(a));
        }
    for (int d = 0; d < 3; d++)
        for (int k = 0; k < P.n[d]; k++) {
            const double s = sin(M_PI * k / P.n[d]);
            tab.push_back(-4.0 * s * s / (g->step[d] * g->step[d]));
        }
    PK_HIP(g->ctx, hipMemcpy(g->proj + 2 * P.nv, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
    g->proj_n = need;
    return PK_OK;
}

// forward: z, y, x passes; inverse: x, y, z, the first of them with the Poisson solve when
// lam != null.  Edges of one cell (1D / 2D grids) need no pass.
template <class R>
int fft3(pk_grid* g, const Phys& P, Cx<R>* w, const double2* tw[3], int inverse, const double* lam,
         ResidualSrc<R> res = ResidualSrc<R>())
{
    bool first = true;
    hipStream_t s = g->ctx->stream;
    const long n12 = (long)P.n[1] * P.n[2];
    // pencils along x (stride n1 n2), y (stride n2), z (contiguous)
    const long inner[3] = {n12, P.n[2], 1};
    const long ostride[3] = {P.nv, n12, P.n[2]};
    const long stride[3] = {n12, P.n[2], 1};
    for (int dd = 0; dd < 3; dd++) {
        const int d = inverse ? dd : 2 - dd;
        const int n = P.n[d];
        if (n == 1) continue;
        const long npencil = P.nv / n;
        const int pb = n <= 256 ? 16 : 4096 / n;          // at most 82 KB of LDS (FP64, n = 1024)
        hipLaunchKernelGGL(k_fft<R>, dim3((unsigned)((npencil + pb - 1) / pb)), dim3(256),
                           sizeof(Cx<R>) * (pb * (n + n / 8 + 1) + n / 2), s, w, log2_exact(n), npencil, log2_exact(pb),
                           log2_exact((int)inner[d]), ostride[d], stride[d], d == 2 ? 1 : 0, tw[d], inverse,
                           (inverse && first) ? lam : nullptr, P, first ? res : ResidualSrc<R>());
        first = false;
    }
    if (first && res.d0)   // a grid of one cell: no pass ran, so the residual is formed alone
        hipLaunchKernelGGL(k_residual_ref<R>, dim3(nblk(P.nv, 256)), dim3(256), 0, s, res.g, P, res.d0, w, res.pack);
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}

// the charge filter of the residual: gain table (device), its size and range; n = 0: none
struct ChargeGain {
    const double* g = nullptr;
    int n = 0;
    double kmax = 0.0;
};

// rho1 != null: residual D J + (rho1 - rho_prev) / dt; else D J - S d0 with d0 the divergence of
// the reference current saved after the complex volume in g->proj and S the radial charge filter
// cg (none when cg.n = 0)
template <class T>
int project(pk_grid* g, const Phys& P, const double* rho1, double dt, ChargeGain cg = ChargeGain())
{
    hipStream_t s = g->ctx->stream;
    Cx<T>* w = (Cx<T>*)g->proj;
    const double2* tw[3];
    const double* tab = g->proj + 2 * P.nv;
    tw[0] = (const double2*)tab;
    tw[1] = tw[0] + P.n[0] / 2;
    tw[2] = tw[1] + P.n[1] / 2;
    // the twiddle block holds n / 2 complex values per dimension (none for an edge of one cell)
    const double* lam = tab + 2 * (P.n[0] / 2 + P.n[1] / 2 + P.n[2] / 2);
    // the reference-divergence residual is formed by the first forward pass as it loads
    ResidualSrc<T> res;
    if (rho1) {
        hipLaunchKernelGGL(k_residual<T>, dim3(nblk(P.nv, 256)), dim3(256), 0, s, grid_view<T>(g), P, rho1,
                           g->rho_prev, 1.0 / dt, w);
    } else {
        res.g = grid_view<T>(g);
        res.d0 = jref_of(g, P);
        res.pack = cg.n > 0 ? 1 : 0;
    }
    int r;
    if ((r = fft3(g, P, w, tw, 0, nullptr, res))) return r;
    if (!rho1 && cg.n > 0) {
        // both real fields went through the one transform; their filtered combination goes to a
        // second volume (the combination reads Z at k and -k)
        Cx<T>* w2 = (Cx<T>*)(jref_of(g, P) + P.nv);
        hipLaunchKernelGGL(k_gain_combine<T>, dim3(nblk(P.nv, 256)), dim3(256), 0, s, w, w2, P, cg.g, cg.n, cg.kmax);
        w = w2;
    }
    if ((r = fft3(g, P, w, tw, 1, lam))) return r;
    hipLaunchKernelGGL(k_correct<T>, dim3(nblk(P.nv, 256)), dim3(256), 0, s, grid_view<T>(g), P, w);
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}

int phys_of(pk_grid* g, Phys& P)
{
    for (int d = 0; d < 3; d++) {
        P.n[d] = g->s[d] - 2 * g->ghost;
        if (P.n[d] < 1 || P.n[d] > 1024 || log2_exact(P.n[d]) < 0)
            return pk_fail(g->ctx, PK_E_UNSUPPORTED, "continuity projection: physical edges must be powers of two");
    }
    P.nv = (long)P.n[0] * P.n[1] * P.n[2];
    if (P.nv < 2) return pk_fail(g->ctx, PK_E_UNSUPPORTED, "continuity projection: grid of a single cell");
    return PK_OK;
}

int save_rho(pk_ens* e, pk_grid* g, int shape)
{
    int r;
    if ((r = pk_rho_fp64(e, g, shape))) return r;
    if (!g->rho_prev && hipMalloc((void**)&g->rho_prev, sizeof(double) * g->n) != hipSuccess) {
        g->rho_prev = nullptr;
        (void)hipGetLastError();
        return pk_fail(g->ctx, PK_E_ALLOC, "pk_rho_save: allocation failed");
    }
    hipLaunchKernelGGL(k_copy_d, dim3(nblk((long)g->n, 256)), dim3(256), 0, g->ctx->stream, g->scratch, g->rho_prev,
                       (long)g->n);
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}

}  // namespace

extern "C" {

int pk_rho_save(pk_ens* e, pk_grid* g, int shape)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if ((r = pk_ens_resort(e))) return r;
    return save_rho(e, g, shape);
}

int pk_project_continuity(pk_ens* e, pk_grid* g, int shape, double dt)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if (!(dt > 0.0)) return pk_fail(e->ctx, PK_E_ARG, "pk_project_continuity: dt must be positive");
    if (!g->rho_prev) return pk_fail(e->ctx, PK_E_STATE, "pk_project_continuity: no saved charge density");
    Phys P;
    if ((r = phys_of(g, P))) return r;
    if ((r = pk_ens_resort(e))) return r;
    if ((r = ensure_proj(g, P))) return r;
    if ((r = pk_rho_fp64(e, g, shape))) return r;
    if ((r = pk_field_ghost(g, PK_FILL_J))) return r;
    {
        pk_ktimer tm(e->ctx, "project");
        r = g->prec == PK_F64 ? project<double>(g, P, g->scratch, dt) : project<float>(g, P, g->scratch, dt);
        if (r) return r;
    }
    if ((r = pk_field_ghost(g, PK_FILL_J))) return r;
    hipLaunchKernelGGL(k_copy_d, dim3(nblk((long)g->n, 256)), dim3(256), 0, g->ctx->stream, g->scratch, g->rho_prev,
                       (long)g->n);
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}

int pk_step_denoised(pk_ens* e, pk_grid* g, pk_dn* d, double dt, int shape, int dep)
{
    int r = pk_check_pair(e, g);
    if (r) return r;
    if (!d) return pk_fail(e->ctx, PK_E_ARG, "pk_step_denoised: no denoiser");
    Phys P;
    if ((r = phys_of(g, P))) return r;
    if ((r = pk_ens_resort(e))) return r;
    // the direct current needs the charge density before the push; the Esirkepov current is its
    // own continuity reference (D J_esirkepov = -(rho1 - rho0) / dt exactly)
    if (dep != PK_ESIRKEPOV && (r = save_rho(e, g, shape))) return r;
    if ((r = pk_field_zero_j(g))) return r;
    if ((r = pk_particle_step(e, g, dt, shape, dep))) return r;
    // Esirkepov: the saved divergence, the denoiser and the residual read J's physical cells only,
    // so its ghosts are filled once, after the projection
    if ((r = pk_field_fold_fill_j(g, dep != PK_ESIRKEPOV))) return r;
    if (dep == PK_ESIRKEPOV) {
        if ((r = ensure_proj(g, P))) return r;
        // the denoiser's input pass also saves D J_esirkepov
        if ((r = pk_dn_apply_impl(d, g, false, jref_of(g, P)))) return r;
        {
            pk_ktimer tm(e->ctx, "project");
            ChargeGain cg;
            pk_dn_charge_gain(d, &cg.g, &cg.n, &cg.kmax);
            r = g->prec == PK_F64 ? project<double>(g, P, nullptr, dt, cg) : project<float>(g, P, nullptr, dt, cg);
            if (r) return r;
        }
        if ((r = pk_field_ghost(g, PK_FILL_J))) return r;
        // the re-sort's insert goes ahead of the field update (disjoint data); its flags are
        // read while the GPU runs the field update
        if ((r = pk_ens_resort_begin(e))) return r;
    } else {
        if ((r = pk_dn_apply(d, g))) return r;
        if ((r = pk_project_continuity(e, g, shape, dt))) return r;
    }
    if (!(r = pk_field_yee_periodic(g, dt * 0.5, false)) && !(r = pk_field_yee_periodic(g, dt, true)))
        r = pk_field_yee_periodic(g, dt * 0.5, false);
    const int rs = pk_ens_resort_end(e);
    return r ? r : rs;
}

}  // extern "C"
