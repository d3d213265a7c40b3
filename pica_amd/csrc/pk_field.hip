// pk_field.hip -- Yee FDTD and periodic ghost kernels (gfx950).
//
// Yee: YeeSolver::Implementation<Three>::updateE / updateB (YeeSolver.h:132-213).  pica's main
// loop plus its three edge loops update every component wherever its curl stencil is inside
// the array, with one formula; each thread here does exactly that for one cell, in pica's
// operation order (this file is compiled with -ffp-contract=off, so no FMA is formed and the
// FP64 result is bit-identical to pica's).  The kernels are streaming stencils: consecutive
// threads walk the z-fastest storage, the +1/-1 neighbours of a row come from L1/L2.
#include "pk_internal.h"

namespace {

template <class T>
struct YeeCoef {
    T cc, kx, ky, kz;
};

template <class T>
__device__ __forceinline__ void yee_e_at(const GridView<T>& g, const YeeCoef<T>& k, long c)
{
    const int s2 = g.s2;
    const long s12 = (long)g.s1 * g.s2;
    const int kk = (int)(c % s2);
    const int jj = (int)((c / s2) % g.s1);
    const int ii = (int)(c / s12);
    T* ex = g.f[PK_EX];
    T* ey = g.f[PK_EY];
    T* ez = g.f[PK_EZ];
    const T* bx = g.f[PK_BX];
    const T* by = g.f[PK_BY];
    const T* bz = g.f[PK_BZ];
    if (jj < g.s1 - 1 && kk < s2 - 1)
        ex[c] += k.cc * g.f[PK_JX][c] + k.ky * (bz[c + s2] - bz[c]) - k.kz * (by[c + 1] - by[c]);
    if (ii < g.s0 - 1 && kk < s2 - 1)
        ey[c] += k.cc * g.f[PK_JY][c] + k.kz * (bx[c + 1] - bx[c]) - k.kx * (bz[c + s12] - bz[c]);
    if (ii < g.s0 - 1 && jj < g.s1 - 1)
        ez[c] += k.cc * g.f[PK_JZ][c] + k.kx * (by[c + s12] - by[c]) - k.ky * (bx[c + s2] - bx[c]);
}

template <class T>
__device__ __forceinline__ void yee_b_at(const GridView<T>& g, const YeeCoef<T>& k, long c)
{
    const int s2 = g.s2;
    const long s12 = (long)g.s1 * g.s2;
    const int kk = (int)(c % s2);
    const int jj = (int)((c / s2) % g.s1);
    const int ii = (int)(c / s12);
    const T* ex = g.f[PK_EX];
    const T* ey = g.f[PK_EY];
    const T* ez = g.f[PK_EZ];
    if (jj >= 1 && kk >= 1)
        g.f[PK_BX][c] += k.kz * (ey[c] - ey[c - 1]) - k.ky * (ez[c] - ez[c - s2]);
    if (ii >= 1 && kk >= 1)
        g.f[PK_BY][c] += k.kx * (ez[c] - ez[c - s12]) - k.kz * (ex[c] - ex[c - 1]);
    if (ii >= 1 && jj >= 1)
        g.f[PK_BZ][c] += k.ky * (ex[c] - ex[c - s2]) - k.kx * (ey[c] - ey[c - s12]);
}

template <class T>
__global__ void __launch_bounds__(256) k_yee_e(GridView<T> g, YeeCoef<T> k, long n)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) yee_e_at(g, k, c);
}

template <class T>
__global__ void __launch_bounds__(256) k_yee_b(GridView<T> g, YeeCoef<T> k, long n)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) yee_b_at(g, k, c);
}

// ---- the periodic step's Yee updates: each thread updates one physical cell with the expressions
// above (at a physical cell every stencil is inside the array) and writes the new value into the
// cell and into every ghost cell that images it, which is what the whole-array update followed by
// a ghost fill leaves there; so the fill launch is not needed.  The update reads only its own cell
// of the component it writes, and neighbours of the other field, whose ghosts are filled.
constexpr int MAX_IMAGES = 1 + 2 * PK_MAX_PERIODIC_GHOST;

// array indices along one dimension whose periodic image is physical index p (n physical cells,
// G ghost layers a side): p itself, then the low ghosts, then the high ghosts
__device__ __forceinline__ int image_list(int p, int n, int G, int out[MAX_IMAGES])
{
    int m = 0;
    out[m++] = G + p;
    for (int a = p + G - n; a >= 0; a -= n) out[m++] = a;
    for (int b = p; b < G; b += n) out[m++] = G + n + b;
    return m;
}

template <class T>
__device__ __forceinline__ void write_images(const GridView<T>& g, int n0, int n1, int n2, int i, int j, int k, int c0,
                                             T v0, T v1, T v2)
{
    int li[MAX_IMAGES], lj[MAX_IMAGES], lk[MAX_IMAGES];
    const int mi = image_list(i, n0, g.G, li), mj = image_list(j, n1, g.G, lj), mk = image_list(k, n2, g.G, lk);
    for (int a = 0; a < mi; a++)
        for (int b = 0; b < mj; b++)
            for (int d = 0; d < mk; d++) {
                const long idx = ((long)li[a] * g.s1 + lj[b]) * g.s2 + lk[d];
                g.f[c0][idx] = v0;
                g.f[c0 + 1][idx] = v1;
                g.f[c0 + 2][idx] = v2;
            }
}

template <class T>
__device__ __forceinline__ void yee_e_periodic_at(const GridView<T>& g, const YeeCoef<T>& k, int n0, int n1, int n2, long v)
{
    if (v >= (long)n0 * n1 * n2) return;
    const int i = (int)(v / ((long)n1 * n2)), j = (int)((v / n2) % n1), kz = (int)(v % n2);
    const int s2 = g.s2;
    const long s12 = (long)g.s1 * g.s2;
    const long c = ((long)(i + g.G) * g.s1 + (j + g.G)) * s2 + (kz + g.G);
    const T* bx = g.f[PK_BX];
    const T* by = g.f[PK_BY];
    const T* bz = g.f[PK_BZ];
    const T ex = g.f[PK_EX][c] + (k.cc * g.f[PK_JX][c] + k.ky * (bz[c + s2] - bz[c]) - k.kz * (by[c + 1] - by[c]));
    const T ey = g.f[PK_EY][c] + (k.cc * g.f[PK_JY][c] + k.kz * (bx[c + 1] - bx[c]) - k.kx * (bz[c + s12] - bz[c]));
    const T ez = g.f[PK_EZ][c] + (k.cc * g.f[PK_JZ][c] + k.kx * (by[c + s12] - by[c]) - k.ky * (bx[c + s2] - bx[c]));
    write_images(g, n0, n1, n2, i, j, kz, PK_EX, ex, ey, ez);
}

template <class T>
__device__ __forceinline__ void yee_b_periodic_at(const GridView<T>& g, const YeeCoef<T>& k, int n0, int n1, int n2, long v)
{
    if (v >= (long)n0 * n1 * n2) return;
    const int i = (int)(v / ((long)n1 * n2)), j = (int)((v / n2) % n1), kz = (int)(v % n2);
    const int s2 = g.s2;
    const long s12 = (long)g.s1 * g.s2;
    const long c = ((long)(i + g.G) * g.s1 + (j + g.G)) * s2 + (kz + g.G);
    const T* ex = g.f[PK_EX];
    const T* ey = g.f[PK_EY];
    const T* ez = g.f[PK_EZ];
    const T bx = g.f[PK_BX][c] + (k.kz * (ey[c] - ey[c - 1]) - k.ky * (ez[c] - ez[c - s2]));
    const T by = g.f[PK_BY][c] + (k.kx * (ez[c] - ez[c - s12]) - k.kz * (ex[c] - ex[c - 1]));
    const T bz = g.f[PK_BZ][c] + (k.ky * (ex[c] - ex[c - s2]) - k.kx * (ey[c] - ey[c - s12]));
    write_images(g, n0, n1, n2, i, j, kz, PK_BX, bx, by, bz);
}

template <class T>
__global__ void __launch_bounds__(256) k_yee_e_periodic(GridView<T> g, YeeCoef<T> k, int n0, int n1, int n2)
{
    yee_e_periodic_at(g, k, n0, n1, n2, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <class T>
__global__ void __launch_bounds__(256) k_yee_b_periodic(GridView<T> g, YeeCoef<T> k, int n0, int n1, int n2)
{
    yee_b_periodic_at(g, k, n0, n1, n2, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- the periodic step's J fold in one launch: every physical cell gathers the ghost cells that
// image it, in the order and association of the fold dimension by dimension (x, then y, then z;
// along a dimension the cell itself, then the low ghosts, then the high ghosts, ascending), so its
// value is that fold's bit for bit.  Ghost cells keep their values (the step fills them next, or
// reads only physical cells until it does); a cell with no image along any dimension is left alone.
// ordered list of the array indices that fold into physical index p along one dimension
__device__ __forceinline__ int fold_list(int p, int n, int G, int out[MAX_IMAGES])
{
    int m = 0;
    out[m++] = G + p;
    for (int a = 0; a < G; a++) {              // low ghost a images (a - G) mod n
        int r = (a - G) % n;
        if (r < 0) r += n;
        if (r == p) out[m++] = a;
    }
    for (int b = 0; b < G; b++)                // high ghost G + n + b images b mod n
        if (b % n == p) out[m++] = G + n + b;
    return m;
}

// the physical cells within G of a face, each once: slab x (i near a face; any j, k), then slab y
// (i inside; j near a face), then slab z (i, j inside; k near a face).  nb = cells near a face
// along a dimension (all n when n <= 2G)
__device__ __forceinline__ int near_face(int q, int n, int G) { return n <= 2 * G || q < G ? q : n - 2 * G + q; }
__device__ __forceinline__ int inside(int r, int G) { return G + r; }   // (only when n > 2G)

// cells near a face of an n0 x n1 x n2 grid with G ghost layers
__host__ __device__ __forceinline__ long near_face_cells(int n0, int n1, int n2, int G)
{
    const int b0 = n0 <= 2 * G ? n0 : 2 * G, b1 = n1 <= 2 * G ? n1 : 2 * G, b2 = n2 <= 2 * G ? n2 : 2 * G;
    return (long)b0 * n1 * n2 + (long)(n0 - b0) * b1 * n2 + (long)(n0 - b0) * (n1 - b1) * b2;
}

template <class T>
__device__ __forceinline__ void fold_j_at(const GridView<T>& g, int n0, int n1, int n2, long t)
{
    const int G = g.G;
    const int b0 = n0 <= 2 * G ? n0 : 2 * G, b1 = n1 <= 2 * G ? n1 : 2 * G, b2 = n2 <= 2 * G ? n2 : 2 * G;
    const long sx = (long)b0 * n1 * n2, sy = (long)(n0 - b0) * b1 * n2, sz = (long)(n0 - b0) * (n1 - b1) * b2;
    int i, j, k;
    if (t < sx) {
        i = near_face((int)(t / ((long)n1 * n2)), n0, G);
        j = (int)((t / n2) % n1);
        k = (int)(t % n2);
    } else if ((t -= sx) < sy) {
        i = inside((int)(t / ((long)b1 * n2)), G);
        j = near_face((int)((t / n2) % b1), n1, G);
        k = (int)(t % n2);
    } else if ((t -= sy) < sz) {
        i = inside((int)(t / ((long)(n1 - b1) * b2)), G);
        j = inside((int)((t / b2) % (n1 - b1)), G);
        k = near_face((int)(t % b2), n2, G);
    } else {
        return;
    }
    int li[MAX_IMAGES], lj[MAX_IMAGES], lk[MAX_IMAGES];
    const int mi = fold_list(i, n0, G, li), mj = fold_list(j, n1, G, lj), mk = fold_list(k, n2, G, lk);
    for (int comp = PK_JX; comp <= PK_JZ; comp++) {
        T* f = g.f[comp];
        T acc_z = T(0);
        for (int c = 0; c < mk; c++) {
            T acc_y = T(0);
            for (int b = 0; b < mj; b++) {
                T acc_x = T(0);
                for (int a = 0; a < mi; a++) {
                    const T x = f[((long)li[a] * g.s1 + lj[b]) * g.s2 + lk[c]];
                    acc_x = a == 0 ? x : acc_x + x;
                }
                acc_y = b == 0 ? acc_x : acc_y + acc_x;
            }
            acc_z = c == 0 ? acc_y : acc_z + acc_y;
        }
        f[((long)(i + G) * g.s1 + (j + G)) * g.s2 + (k + G)] = acc_z;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_fold_j_gather(GridView<T> g, int n0, int n1, int n2)
{
    fold_j_at(g, n0, n1, n2, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Periodic fill: every ghost cell takes the value of its wrapped physical image.  The ghost
// cells are enumerated as x slabs (full y,z), then y slabs (physical x, full z), then z slabs
// (physical x,y); the closed-form image equals the dimension-by-dimension copy.
// array index of the periodic image of index i (physical indices [G, G + N)); an edge of N < G
// cells (1D / 2D grids: N = 1) wraps more than once
__device__ __forceinline__ int ghost_image(int i, int G, int N)
{
    if (i >= G && i < G + N) return i;
    int r = (i - G) % N;
    if (r < 0) r += N;
    return G + r;
}

template <class T>
__device__ __forceinline__ void ghost_fill_at(const GridView<T>& g, unsigned mask, long nx_, long ny_, long nz_, long t)
{
    const int G = g.G;
    const int N0 = g.s0 - 2 * G, N1 = g.s1 - 2 * G, N2 = g.s2 - 2 * G;
    int i, j, k;
    if (t < nx_) {
        const int a = (int)(t / ((long)g.s1 * g.s2));
        const long r = t % ((long)g.s1 * g.s2);
        i = a < G ? a : N0 + a;
        j = (int)(r / g.s2);
        k = (int)(r % g.s2);
    } else if ((t -= nx_) < ny_) {
        const int b = (int)((t / g.s2) % (2 * G));
        i = G + (int)(t / ((long)2 * G * g.s2));
        j = b < G ? b : N1 + b;
        k = (int)(t % g.s2);
    } else if ((t -= ny_) < nz_) {
        const int cc = (int)(t % (2 * G));
        const long r = t / (2 * G);
        j = G + (int)(r % N1);
        i = G + (int)(r / N1);
        k = cc < G ? cc : N2 + cc;
    } else {
        return;
    }
    const int si = ghost_image(i, G, N0), sj = ghost_image(j, G, N1), sk = ghost_image(k, G, N2);
    const long dst = ((long)i * g.s1 + j) * g.s2 + k;
    const long src = ((long)si * g.s1 + sj) * g.s2 + sk;
    for (int c = 0; c < PK_NCOMP; c++)
        if (mask & (1u << c)) g.f[c][dst] = g.f[c][src];
}

template <class T>
__global__ void __launch_bounds__(256) k_ghost_fill(GridView<T> g, unsigned mask, long nx_, long ny_, long nz_)
{
    ghost_fill_at(g, mask, nx_, ny_, nz_, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Periodic fold along dimension d: physical += ghost image, low ghosts first then high ones,
// over the full extent of the other two dimensions (the oracle's order exactly).
template <class T>
__device__ __forceinline__ void ghost_fold_at(const GridView<T>& g, unsigned mask, int d, long t)
{
    const int s[3] = {g.s0, g.s1, g.s2};
    const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
    if (t >= (long)s[o1] * s[o2]) return;
    const int p = (int)(t / s[o2]), q = (int)(t % s[o2]);
    const int G = g.G;
    const int N = s[d] - 2 * G;
    long stride[3];
    stride[2] = 1;
    stride[1] = g.s2;
    stride[0] = (long)g.s1 * g.s2;
    const long base = p * stride[o1] + q * stride[o2];
    for (int c = 0; c < PK_NCOMP; c++) {
        if (!(mask & (1u << c))) continue;
        T* f = g.f[c];
        for (int a = 0; a < G; a++) f[base + ghost_image(a, G, N) * stride[d]] += f[base + a * stride[d]];
        for (int a = 0; a < G; a++) f[base + ghost_image(G + N + a, G, N) * stride[d]] += f[base + (G + N + a) * stride[d]];
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_ghost_fold(GridView<T> g, unsigned mask, int d)
{
    ghost_fold_at(g, mask, d, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

template <class T>
__global__ void __launch_bounds__(256) k_fill_zero(T* p, long n)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = T(0);
}

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

template <class T>
YeeCoef<T> yee_coef(const pk_grid* g, double dt_in, bool e_update)
{
    const T dt = (T)dt_in;
    YeeCoef<T> k;
    k.cc = e_update ? -(T)4 * (T)PK_PI * dt : (T)0;    // YeeSolver.h:135
    const T cdt = (T)PK_C * dt;                        // :137, :181-182
    k.kx = cdt / (T)g->step[0];
    k.ky = cdt / (T)g->step[1];
    k.kz = cdt / (T)g->step[2];
    return k;
}

template <class T>
int yee_e(pk_grid* g, double dt)
{
    const long n = (long)g->n;
    pk_ktimer tm(g->ctx, "yee_e");
    hipLaunchKernelGGL(k_yee_e<T>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream, grid_view<T>(g),
                       yee_coef<T>(g, dt, true), n);
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}

template <class T>
int yee_b(pk_grid* g, double dt)
{
    const long n = (long)g->n;
    pk_ktimer tm(g->ctx, "yee_b");
    hipLaunchKernelGGL(k_yee_b<T>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream, grid_view<T>(g),
                       yee_coef<T>(g, dt, false), n);
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}

template <class T>
int ghost(pk_grid* g, int ops)
{
    const int G = g->ghost;
    if (G <= 0) return pk_fail(g->ctx, PK_E_ARG, "pk_ghost: grid has no ghost layers");
    for (int d = 0; d < 3; d++)
        if (g->s[d] - 2 * G < 1) return pk_fail(g->ctx, PK_E_ARG, "pk_ghost: no physical cells");
    unsigned fold = 0, fill = 0;
    if (ops & PK_FOLD_J) fold |= (1u << PK_JX) | (1u << PK_JY) | (1u << PK_JZ);
    if (ops & PK_FOLD_RHO) fold |= (1u << PK_RHO);
    if (ops & PK_FILL_E) fill |= (1u << PK_EX) | (1u << PK_EY) | (1u << PK_EZ);
    if (ops & PK_FILL_B) fill |= (1u << PK_BX) | (1u << PK_BY) | (1u << PK_BZ);
    if (ops & PK_FILL_J) fill |= (1u << PK_JX) | (1u << PK_JY) | (1u << PK_JZ);
    if (ops & PK_FOLD_RHO) fill |= (1u << PK_RHO);
    GridView<T> v = grid_view<T>(g);
    pk_ktimer tm(g->ctx, "ghost");
    if (fold) {
        for (int d = 0; d < 3; d++) {
            const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
            const long n = (long)g->s[o1] * g->s[o2];
            hipLaunchKernelGGL(k_ghost_fold<T>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream, v, fold, d);
        }
    }
    if (fill) {
        const long nx = (long)2 * G * g->s[1] * g->s[2];
        const long ny = (long)(g->s[0] - 2 * G) * 2 * G * g->s[2];
        const long nz = (long)(g->s[0] - 2 * G) * (g->s[1] - 2 * G) * 2 * G;
        hipLaunchKernelGGL(k_ghost_fill<T>, dim3(nblk(nx + ny + nz, 256)), dim3(256), 0, g->ctx->stream, v, fill,
                           nx, ny, nz);
    }
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}

// ---- many grids per launch ----
template <class T>
struct FieldJob {
    GridView<T> g;
    YeeCoef<T> ke, kb;      // updateE(dt), updateB(dt / 2)
    long ncell, nx, ny, nz;
    int n0, n1, n2;         // physical cells
};

// ensemble of block b: the last i with start[i] <= b (empty ranges are skipped)
__device__ __forceinline__ int job_of_block(const long* start, int n, long b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <class T, int OP>
__global__ void __launch_bounds__(256) k_batch_field(const FieldJob<T>* __restrict__ jobs, const long* __restrict__ start,
                                                     int n, unsigned mask, int d, const int* __restrict__ halt)
{
    if (halt[PK_HALT_STOP]) return;
    const int i = job_of_block(start, n, blockIdx.x);
    const FieldJob<T>& J = jobs[i];
    const long t = (blockIdx.x - start[i]) * 256 + threadIdx.x;
    if (OP == PK_BOP_ZERO_J) {
        if (t < J.ncell)
            for (int c = PK_JX; c <= PK_JZ; c++) J.g.f[c][t] = T(0);
    } else if (OP == PK_BOP_FILL) {
        ghost_fill_at(J.g, mask, J.nx, J.ny, J.nz, t);
    } else if (OP == PK_BOP_YEE_E_P) {
        yee_e_periodic_at(J.g, J.ke, J.n0, J.n1, J.n2, t);
    } else if (OP == PK_BOP_YEE_B_P) {
        yee_b_periodic_at(J.g, J.kb, J.n0, J.n1, J.n2, t);
    } else {
        fold_j_at(J.g, J.n0, J.n1, J.n2, t);
    }
}

template <class T>
int batch_prepare(pk_batch* b, double dt)
{
    const int n = b->n();
    std::vector<FieldJob<T>> jobs(n);
    std::vector<long> start((size_t)PK_BSTAGES * (n + 1), 0);
    for (int i = 0; i < n; i++) {
        const pk_grid* g = b->grids[i];
        FieldJob<T>& J = jobs[i];
        J.g = grid_view<T>(g);
        J.ke = yee_coef<T>(g, dt, true);
        J.kb = yee_coef<T>(g, dt * 0.5, false);
        const int G = g->ghost;
        if (G < 1 || G > PK_MAX_PERIODIC_GHOST)
            return pk_fail(b->ctx, PK_E_ARG, "batch: grids need 1 or 2 periodic ghost layers");
        J.ncell = (long)g->n;
        J.nx = (long)2 * G * g->s[1] * g->s[2];
        J.ny = (long)(g->s[0] - 2 * G) * 2 * G * g->s[2];
        J.nz = (long)(g->s[0] - 2 * G) * (g->s[1] - 2 * G) * 2 * G;
        J.n0 = g->s[0] - 2 * G;
        J.n1 = g->s[1] - 2 * G;
        J.n2 = g->s[2] - 2 * G;
        const long work[PK_BSTAGES] = {J.ncell, J.nx + J.ny + J.nz, (long)J.n0 * J.n1 * J.n2,
                                       near_face_cells(J.n0, J.n1, J.n2, G)};
        for (int st = 0; st < PK_BSTAGES; st++)
            start[(size_t)st * (n + 1) + i + 1] = start[(size_t)st * (n + 1) + i] + (long)nblk(work[st], 256);
    }
    for (int st = 0; st < PK_BSTAGES; st++) b->fblocks[st] = start[(size_t)st * (n + 1) + n];
    int r = pk_batch_upload(b, &b->fjobs, jobs.data(), sizeof(FieldJob<T>) * n);
    if (r) return r;
    return pk_batch_upload(b, (void**)&b->fstart, start.data(), sizeof(long) * start.size());
}

template <class T>
int batch_launch(pk_batch* b, int op, unsigned mask, int d)
{
    const int n = b->n();
    const int st = op == PK_BOP_FILL                             ? PK_BSTAGE_FILL
                   : op == PK_BOP_YEE_E_P || op == PK_BOP_YEE_B_P ? PK_BSTAGE_PHYS
                   : op == PK_BOP_FOLD_J_P                       ? PK_BSTAGE_FACE
                                                                 : PK_BSTAGE_CELLS;
    const long nb = b->fblocks[st];
    if (nb == 0) return PK_OK;
    const FieldJob<T>* jobs = (const FieldJob<T>*)b->fjobs;
    const long* start = b->fstart + (size_t)st * (n + 1);
    hipStream_t s = b->ctx->stream;
    switch (op) {
    case PK_BOP_ZERO_J:
        hipLaunchKernelGGL((k_batch_field<T, PK_BOP_ZERO_J>), dim3(nb), dim3(256), 0, s, jobs, start, n, mask, d, b->halt);
        break;
    case PK_BOP_FILL:
        hipLaunchKernelGGL((k_batch_field<T, PK_BOP_FILL>), dim3(nb), dim3(256), 0, s, jobs, start, n, mask, d, b->halt);
        break;
    case PK_BOP_YEE_E_P:
        hipLaunchKernelGGL((k_batch_field<T, PK_BOP_YEE_E_P>), dim3(nb), dim3(256), 0, s, jobs, start, n, mask, d, b->halt);
        break;
    case PK_BOP_YEE_B_P:
        hipLaunchKernelGGL((k_batch_field<T, PK_BOP_YEE_B_P>), dim3(nb), dim3(256), 0, s, jobs, start, n, mask, d, b->halt);
        break;
    default:
        hipLaunchKernelGGL((k_batch_field<T, PK_BOP_FOLD_J_P>), dim3(nb), dim3(256), 0, s, jobs, start, n, mask, d, b->halt);
        break;
    }
    PK_HIP(b->ctx, hipGetLastError());
    return PK_OK;
}

}  // namespace

int pk_field_zero_j(pk_grid* g)
{
    const long n = (long)(2 * g->pitch + g->n);
    pk_ktimer tm(g->ctx, "zero");
    if (g->prec == PK_F64)
        hipLaunchKernelGGL(k_fill_zero<double>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream,
                           (double*)g->comp(PK_JX), n);
    else
        hipLaunchKernelGGL(k_fill_zero<float>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream,
                           (float*)g->comp(PK_JX), n);
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}

int pk_field_zero(pk_grid* g, int comp)
{
    const long n = (long)g->n;
    pk_ktimer tm(g->ctx, "zero");
    if (g->prec == PK_F64)
        hipLaunchKernelGGL(k_fill_zero<double>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream,
                           (double*)g->comp(comp), n);
    else
        hipLaunchKernelGGL(k_fill_zero<float>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream,
                           (float*)g->comp(comp), n);
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}

int pk_field_yee_e(pk_grid* g, double dt) { return g->prec == PK_F64 ? yee_e<double>(g, dt) : yee_e<float>(g, dt); }

namespace {
template <class T>
int yee_periodic(pk_grid* g, double dt, bool e_update)
{
    const int G = g->ghost;
    const int n0 = g->s[0] - 2 * G, n1 = g->s[1] - 2 * G, n2 = g->s[2] - 2 * G;
    const long n = (long)n0 * n1 * n2;
    pk_ktimer tm(g->ctx, e_update ? "yee_e" : "yee_b");
    if (e_update)
        hipLaunchKernelGGL(k_yee_e_periodic<T>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream, grid_view<T>(g),
                           yee_coef<T>(g, dt, true), n0, n1, n2);
    else
        hipLaunchKernelGGL(k_yee_b_periodic<T>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream, grid_view<T>(g),
                           yee_coef<T>(g, dt, false), n0, n1, n2);
    PK_HIP(g->ctx, hipGetLastError());
    return PK_OK;
}
}  // namespace

int pk_field_fold_fill_j(pk_grid* g, bool fill)
{
    const int G = g->ghost;
    if (G < 1 || G > PK_MAX_PERIODIC_GHOST) return pk_field_ghost(g, fill ? PK_FOLD_J | PK_FILL_J : PK_FOLD_J);
    const int n0 = g->s[0] - 2 * G, n1 = g->s[1] - 2 * G, n2 = g->s[2] - 2 * G;
    if (n0 < 1 || n1 < 1 || n2 < 1) return pk_fail(g->ctx, PK_E_ARG, "pk_ghost: no physical cells");
    const long n = near_face_cells(n0, n1, n2, G);
    {
        pk_ktimer tm(g->ctx, "ghost");
        if (g->prec == PK_F64)
            hipLaunchKernelGGL(k_fold_j_gather<double>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream,
                               grid_view<double>(g), n0, n1, n2);
        else
            hipLaunchKernelGGL(k_fold_j_gather<float>, dim3(nblk(n, 256)), dim3(256), 0, g->ctx->stream,
                               grid_view<float>(g), n0, n1, n2);
        PK_HIP(g->ctx, hipGetLastError());
    }
    return fill ? pk_field_ghost(g, PK_FILL_J) : PK_OK;
}

int pk_field_yee_periodic(pk_grid* g, double dt, bool e_update)
{
    const int G = g->ghost;
    if (G < 1 || G > PK_MAX_PERIODIC_GHOST) {
        // not the periodic step's layout: pica's whole-array update, then the fill
        const int r = e_update ? pk_field_yee_e(g, dt) : pk_field_yee_b(g, dt);
        return r ? r : pk_field_ghost(g, e_update ? PK_FILL_E : PK_FILL_B);
    }
    for (int d = 0; d < 3; d++)
        if (g->s[d] - 2 * G < 1) return pk_fail(g->ctx, PK_E_ARG, "Yee update: no physical cells");
    return g->prec == PK_F64 ? yee_periodic<double>(g, dt, e_update) : yee_periodic<float>(g, dt, e_update);
}
int pk_field_yee_b(pk_grid* g, double dt) { return g->prec == PK_F64 ? yee_b<double>(g, dt) : yee_b<float>(g, dt); }
int pk_field_ghost(pk_grid* g, int ops) { return g->prec == PK_F64 ? ghost<double>(g, ops) : ghost<float>(g, ops); }

int pk_batch_fields_prepare(pk_batch* b, double dt)
{
    return b->prec == PK_F64 ? batch_prepare<double>(b, dt) : batch_prepare<float>(b, dt);
}

int pk_batch_fields_launch(pk_batch* b, int op, unsigned mask, int d)
{
    return b->prec == PK_F64 ? batch_launch<double>(b, op, mask, d) : batch_launch<float>(b, op, mask, d);
}
