// pk_particles_split.hip -- FP32 split-position particle step (gfx950).
//
// A particle is an int32 linear physical cell index, a float offset in [0, 1) per dimension,
// float u = p/(mc) and a float weight (32 bytes).  Every stencil index is exact small-integer
// arithmetic relative to the particle's cell; only the in-cell offset carries rounding, so
// the position resolution is 2^-24 of a cell everywhere in the domain.  Fields and J in HBM
// are FP32.  The kernel is the FP64 kernel's pipeline (pk_particles_f64.hip): stencil in LDS
// -> interpolate -> Boris push (BorisPusher.h:32-47) -> deposit into an LDS J tile -> movers
// to the outbox, tail stayers fill holes -> one flush of global FP32 atomics.
//
// The LDS J tile is accumulated in FP64: on MI355X a native ds_add_f64 wave instruction costs
// ~9 cycles per CU and a native ds_add_f32 ~190 (measured 7.1e10 against 3.2e9 wave
// instructions/s on the whole chip, whatever the bank stride or denormal mode), and the extra
// precision is free.  Esirkepov deposition uses a per-wavefront window: 4 (TSC) or 3 (CIC) nodes in a
// dimension only when some lane of the wavefront changes its nearest node there, else 3 / 2.
// Particles that change node go to a workgroup change queue deposited by full wavefronts after
// the push, so most wavefronts issue only the small window.  The push loop has no workgroup
// barrier: each wavefront owns every fourth 128-slot group and the four drift apart, which
// lets one wavefront's LDS atomics overlap another's arithmetic.  A lane pushes two adjacent
// slots; when both particles keep their node and share a window (the loader and the periodic
// interleave place two particles of one cell side by side) their Esirkepov terms are summed
// before the LDS atomics, which halves the atomics of such a pair.
//
// Stencil conventions (grid origin -2h, two ghost layers; l = cell - supercell start):
//   interpolation tile base = array index m+1 (pica's baseIdx), CIC n+2 / TSC n+3 wide;
//   a staggered node of cell c sits at tile index l+1, the collocated node c at tile l+1;
//   deposition tile base = array index m (n+4 wide) except pica's CIC direct (m+1, n+3).
#include "pk_internal.h"

namespace {

constexpr int WAVES = PK_BLOCK / 64;
constexpr int REC_WORDS = 8;     // deposit record: packed cell/shift/class, o[3], on[3], q w
constexpr int QCAP = 256;        // change-queue records per workgroup
// x-stride padding of the FP64 J tile (doubles).  With 8 x 8 x 8 nodes unpadded, the 4^3 cells a
// wavefront's particles sit in fall on 8 of the 16 double-wide LDS bank pairs; +4 spreads them 4
// per pair, which lowers the bank conflicts of the deposit atomics.
constexpr int JPAD = 4;
constexpr int IPAD = 4;          // x-stride padding of the field tile (floats), same purpose

__device__ __forceinline__ bool colloc(int comp, int d)
{
    const unsigned m[6] = {1u, 2u, 4u, 6u, 5u, 3u};
    return (m[comp] >> d) & 1u;
}

__device__ __forceinline__ void tsc_w(float d, float w[3])
{
    const float a = 0.5f - d, b = 0.5f + d;
    w[0] = 0.5f * a * a;
    w[1] = 0.75f - d * d;
    w[2] = 0.5f * b * b;
}

// per-dimension interpolation stencil of one particle: first tile index and weights of the
// collocated and the staggered component grids
template <int SHAPE>
__device__ __forceinline__ void interp_stencil(int l, float o, int& bc, float wc[3], int& bs, float ws[3])
{
    if (SHAPE == PK_TSC) {
        const int h = o >= 0.5f ? 1 : 0;
        bc = l + h;                  // nearest node c+h, nodes at tile l+h .. l+h+2
        tsc_w(o - (float)h, wc);
        bs = l;                      // nearest staggered node is the cell's own centre (tile l+1)
        tsc_w(o - 0.5f, ws);
    } else {
        bc = l + 1;                  // nodes c, c+1 at tile l+1, l+2
        wc[0] = 1.0f - o; wc[1] = o; wc[2] = 0.0f;
        const float s = o - 0.5f;
        if (s < 0.0f) { bs = l; ws[0] = -s; ws[1] = 1.0f + s; }   // nodes c-1, c at tile l, l+1
        else { bs = l + 1; ws[0] = 1.0f - s; ws[1] = s; }          // nodes c, c+1
        ws[2] = 0.0f;
    }
}

template <int SHAPE>
__device__ __forceinline__ float gather(const float* t, int sx, int sy, const int b[3], const float w[3][3])
{
    constexpr int K = SHAPE == PK_TSC ? 3 : 2;
    const float* base = t + b[0] * sx + b[1] * sy + b[2];
    float sa = 0.0f;
#pragma unroll
    for (int a = 0; a < K; a++) {
        float sb = 0.0f;
#pragma unroll
        for (int bb = 0; bb < K; bb++) {
            const float* row = base + a * sx + bb * sy;
            float sc = 0.0f;
#pragma unroll
            for (int c = 0; c < K; c++) sc += w[2][c] * row[c];
            sb += w[1][bb] * sc;
        }
        sa += w[0][a] * sb;
    }
    return sa;
}

// shape on the cell-centre (rho) grid, 4-entry window starting at node wb, of a particle whose
// own cell centre is node nc (deposition tile coordinates) and offset o.  The particle's lowest
// node is wb or wb + 1 (the window of one step covers the old and the new shape).
template <int SHAPE>
__device__ __forceinline__ void rho_shape(int nc, float o, int wb, float S[4])
{
    float ww[3];
    int lo;
    if (SHAPE == PK_TSC) {
        tsc_w(o - 0.5f, ww);          // nodes nc-1, nc, nc+1
        lo = nc - 1;
    } else {
        const float s = o - 0.5f;     // lower node nc-1 (s < 0) or nc
        lo = s < 0.0f ? nc - 1 : nc;
        const float f = s < 0.0f ? s + 1.0f : s;
        ww[0] = 1.0f - f;
        ww[1] = f;
        ww[2] = 0.0f;
    }
    const bool up = lo != wb;
    S[0] = up ? 0.0f : ww[0];
    S[1] = up ? ww[0] : ww[1];
    S[2] = up ? ww[1] : ww[2];
    S[3] = up ? ww[2] : 0.0f;
}

// Esirkepov W decomposition (CPC 135 (2001) 144) on a KX x KY x KZ window, accumulated in
// FP64 LDS.  W_c(a,p,r) = DS_c[a] (S0_1 S0_2 + DS_1 S0_2/2 + S0_1 DS_2/2 + DS_1 DS_2/3), written
// as DS_c[a] (A[p] S0_2[r] + B[p] DS_2[r]); J_c on face a+1 is the running sum over a.
template <int KX, int KY, int KZ>
__device__ __forceinline__ void esirkepov(double* jt, int dtn, int sx, int sy, const int wb[3], const float S0[3][4],
                                          const float DS[3][4], const float kq[3])
{
    constexpr int K[3] = {KX, KY, KZ};
    const int st[3] = {sx, sy, 1};
    const int orig = wb[0] * sx + wb[1] * sy + wb[2];
    const float third = 1.0f / 3.0f;
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
        const int d1 = (comp + 1) % 3, d2 = (comp + 2) % 3;
        float A[4], B[4];
#pragma unroll
        for (int p = 0; p < 4; p++) {
            A[p] = kq[comp] * (S0[d1][p] + 0.5f * DS[d1][p]);
            B[p] = kq[comp] * (0.5f * S0[d1][p] + third * DS[d1][p]);
        }
        double* t = jt + comp * dtn + orig + st[comp];
#pragma unroll
        for (int p = 0; p < K[d1]; p++)
#pragma unroll
            for (int r = 0; r < K[d2]; r++) {
                const float coef = A[p] * S0[d2][r] + B[p] * DS[d2][r];
                float cum = 0.0f;
#pragma unroll
                for (int a = 0; a < K[comp] - 1; a++) {
                    cum += DS[comp][a] * coef;
                    atomicAdd(&t[a * st[comp] + p * st[d1] + r * st[d2]], (double)cum);
                }
            }
    }
}

// fields at a particle from the LDS stencil tile (tile strides sx, sy; n per component)
template <int SHAPE>
__device__ __forceinline__ void interp_fields(const float* itile, int n, int sx, int sy, const int l[3],
                                              const float o[3], float e[3], float b[3])
{
    int bc[3], bs[3];
    float wc[3][3], ws[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++) interp_stencil<SHAPE>(l[d], o[d], bc[d], wc[d], bs[d], ws[d]);
#pragma unroll
    for (int comp = 0; comp < 6; comp++) {
        int bi[3];
        float wt[3][3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const bool col = colloc(comp, d);
            bi[d] = col ? bc[d] : bs[d];
#pragma unroll
            for (int k = 0; k < 3; k++) wt[d][k] = col ? wc[d][k] : ws[d][k];
        }
        const float val = gather<SHAPE>(itile + comp * n, sx, sy, bi, wt);
        if (comp < 3) e[comp] = val; else b[comp - 3] = val;
    }
}

// Boris push (BorisPusher.h:32-47) in FP32; coeff = q dt / (2 m c)
__device__ __forceinline__ void boris_f32(float u[3], const float e[3], const float b[3], float coeff)
{
    float eM[3], um[3], t[3], up[3], sv[3];
#pragma unroll
    for (int d = 0; d < 3; d++) eM[d] = e[d] * coeff;
#pragma unroll
    for (int d = 0; d < 3; d++) um[d] = u[d] + eM[d];
    const float tk = coeff * rsqrtf(1.0f + um[0] * um[0] + um[1] * um[1] + um[2] * um[2]);
#pragma unroll
    for (int d = 0; d < 3; d++) t[d] = b[d] * tk;
    up[0] = um[0] + (um[1] * t[2] - um[2] * t[1]);
    up[1] = um[1] + (um[2] * t[0] - um[0] * t[2]);
    up[2] = um[2] + (um[0] * t[1] - um[1] * t[0]);
    const float sk = 2.0f * __builtin_amdgcn_rcpf(1.0f + t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
#pragma unroll
    for (int d = 0; d < 3; d++) sv[d] = t[d] * sk;
    u[0] = um[0] + (up[1] * sv[2] - up[2] * sv[1]) + eM[0];
    u[1] = um[1] + (up[2] * sv[0] - up[0] * sv[2]) + eM[1];
    u[2] = um[2] + (up[0] * sv[1] - up[1] * sv[0]) + eM[2];
}

// new offset and cell shift in {-1, 0, 1} from velocity v (units of c); cdt_h = c dt / h
__device__ __forceinline__ void move_split(const float o[3], const float v[3], const float cdt_h[3], float on[3],
                                           int sh[3])
{
#pragma unroll
    for (int d = 0; d < 3; d++) {
        float x = o[d] + v[d] * cdt_h[d];
        int k = x < 0.0f ? -1 : (x >= 1.0f ? 1 : 0);
        x -= (float)k;
        if (x >= 1.0f) { x = 0.0f; k += 1; }     // a tiny negative offset rounded up to 1
        if (x < 0.0f) x = 0.0f;
        on[d] = x;
        sh[d] = k;
    }
}

// does the particle's Esirkepov window in one dimension need the larger size (TSC: its cell
// changes; CIC: its lower cell-centre node changes)?
template <int SHAPE>
__device__ __forceinline__ bool node_change(float o, float on, int sh)
{
    if (SHAPE == PK_TSC) return sh != 0;
    const int lo0 = o < 0.5f ? -1 : 0;
    const int lo1 = sh + (on < 0.5f ? -1 : 0);
    return lo0 != lo1;
}

// window base wb and the old shape S0 / shape change DS of one particle (local cell l, shift sh,
// offsets o -> on) on the base-m deposition tile; big[d] selects the window size per dimension
template <int SHAPE>
__device__ __forceinline__ void esk_shapes(const int l[3], const int sh[3], const float o[3], const float on[3],
                                           const bool big[3], int wb[3], float S0[3][4], float DS[3][4])
{
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const int nc0 = l[d] + 2, nc1 = l[d] + sh[d] + 2;
        if (SHAPE == PK_TSC) {
            wb[d] = (big[d] ? (nc0 < nc1 ? nc0 : nc1) : nc0) - 1;
        } else {
            const int lo0 = o[d] < 0.5f ? nc0 - 1 : nc0;
            const int lo1 = on[d] < 0.5f ? nc1 - 1 : nc1;
            wb[d] = big[d] ? (lo0 < lo1 ? lo0 : lo1) : lo0;
        }
        float S1[4];
        rho_shape<SHAPE>(nc0, o[d], wb[d], S0[d]);
        rho_shape<SHAPE>(nc1, on[d], wb[d], S1);
#pragma unroll
        for (int k = 0; k < 4; k++) DS[d][k] = S1[k] - S0[d][k];
    }
}

// Esirkepov deposit of one particle on the deposition tile; big[d] selects the window size per
// dimension (lanes with different choices take different branches of the switch)
template <int SHAPE>
__device__ __forceinline__ void esirkepov_window(double* jt, int dtn, int dsx, int dsy, const int l[3],
                                                 const int sh[3], const float o[3], const float on[3],
                                                 const bool big[3], float qw, const float hdt[3])
{
    float S0[3][4], DS[3][4];
    int wb[3];
    esk_shapes<SHAPE>(l, sh, o, on, big, wb, S0, DS);
    const float kq[3] = {-qw * hdt[0], -qw * hdt[1], -qw * hdt[2]};
    constexpr int KS = SHAPE == PK_TSC ? 3 : 2;    // window without a node change
    const int sel = (big[0] ? 4 : 0) | (big[1] ? 2 : 0) | (big[2] ? 1 : 0);
    switch (sel) {
    case 0: esirkepov<KS, KS, KS>(jt, dtn, dsx, dsy, wb, S0, DS, kq); break;
    case 1: esirkepov<KS, KS, KS + 1>(jt, dtn, dsx, dsy, wb, S0, DS, kq); break;
    case 2: esirkepov<KS, KS + 1, KS>(jt, dtn, dsx, dsy, wb, S0, DS, kq); break;
    case 3: esirkepov<KS, KS + 1, KS + 1>(jt, dtn, dsx, dsy, wb, S0, DS, kq); break;
    case 4: esirkepov<KS + 1, KS, KS>(jt, dtn, dsx, dsy, wb, S0, DS, kq); break;
    case 5: esirkepov<KS + 1, KS, KS + 1>(jt, dtn, dsx, dsy, wb, S0, DS, kq); break;
    case 6: esirkepov<KS + 1, KS + 1, KS>(jt, dtn, dsx, dsy, wb, S0, DS, kq); break;
    default: esirkepov<KS + 1, KS + 1, KS + 1>(jt, dtn, dsx, dsy, wb, S0, DS, kq); break;
    }
}

// Esirkepov deposit of two particles P and Q on one K x K x K window (neither changes its node):
// every tile node receives the sum of both particles' terms with one atomic.  kqQ = 0 (and
// finite Q shapes) deposits P alone.
template <int K>
__device__ __forceinline__ void esirkepov_pair(double* jt, int dtn, int sx, int sy, const int wb[3], const float S0P[3][4],
                                               const float DSP[3][4], const float kqP[3], const float S0Q[3][4],
                                               const float DSQ[3][4], const float kqQ[3])
{
    const int st[3] = {sx, sy, 1};
    const int orig = wb[0] * sx + wb[1] * sy + wb[2];
    const float third = 1.0f / 3.0f;
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
        const int d1 = (comp + 1) % 3, d2 = (comp + 2) % 3;
        float AP[K], BP[K], AQ[K], BQ[K];
#pragma unroll
        for (int p = 0; p < K; p++) {
            AP[p] = kqP[comp] * (S0P[d1][p] + 0.5f * DSP[d1][p]);
            BP[p] = kqP[comp] * (0.5f * S0P[d1][p] + third * DSP[d1][p]);
            AQ[p] = kqQ[comp] * (S0Q[d1][p] + 0.5f * DSQ[d1][p]);
            BQ[p] = kqQ[comp] * (0.5f * S0Q[d1][p] + third * DSQ[d1][p]);
        }
        double* t = jt + comp * dtn + orig + st[comp];
#pragma unroll
        for (int p = 0; p < K; p++)
#pragma unroll
            for (int r = 0; r < K; r++) {
                const float cp = AP[p] * S0P[d2][r] + BP[p] * DSP[d2][r];
                const float cq = AQ[p] * S0Q[d2][r] + BQ[p] * DSQ[d2][r];
                float cum = 0.0f;
#pragma unroll
                for (int a = 0; a < K - 1; a++) {
                    cum += DSP[comp][a] * cp + DSQ[comp][a] * cq;
                    atomicAdd(&t[a * st[comp] + p * st[d1] + r * st[d2]], (double)cum);
                }
            }
    }
}

// bit k of x moved to bit 2k
__device__ __forceinline__ uint64_t spread_bits(uint32_t x)
{
    uint64_t v = x;
    v = (v | (v << 16)) & 0x0000FFFF0000FFFFull;
    v = (v | (v << 8)) & 0x00FF00FF00FF00FFull;
    v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0Full;
    v = (v | (v << 2)) & 0x3333333333333333ull;
    v = (v | (v << 1)) & 0x5555555555555555ull;
    return v;
}

struct SplitParams {
    float coeff;      // q dt / (2 m c) (BorisPusher.h:29)
    float cdt_h[3];   // c dt / h per dimension (position update in cells)
    float q;          // species charge
    float hdt[3];     // h / dt per dimension (Esirkepov)
    float norm;       // 1 / V
    int mask_words;
};

// all arguments of one ensemble's particle step (one kernel argument, or one entry of a batch)
struct SplitArgs {
    GridView<float> g;
    EnsGeom eg;
    PartSplit pm, po;
    uint8_t* odir;
    const int64_t* off;
    const int32_t* cnt;
    int32_t *nstay, *nmove, *mvc, *flags;
    SplitParams sp;
};

// One supercell s of one ensemble.  NC > 0: cubic supercells of NC cells (compile-time tile
// strides); NC == 0: any shape.
template <int SHAPE, int DEP, int NC, int BS = PK_BLOCK>
__device__ __forceinline__ void particles_split_sc(const SplitArgs& A, long s)
{
    constexpr int WV = BS / 64;      // wavefronts of the workgroup
    const GridView<float>& g = A.g;
    const EnsGeom& eg = A.eg;
    const PartSplit& pm = A.pm;
    const PartSplit& po = A.po;
    uint8_t* __restrict__ odir = A.odir;
    const int64_t* __restrict__ off = A.off;
    const int32_t* __restrict__ cnt = A.cnt;
    int32_t* __restrict__ nstay = A.nstay;
    int32_t* __restrict__ nmove = A.nmove;
    int32_t* __restrict__ mvc = A.mvc;
    int32_t* __restrict__ flags = A.flags;
    const SplitParams& sp = A.sp;
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr bool PICA_CIC = SHAPE == PK_CIC && DEP == PK_DIRECT;
    const int nn[3] = {NC ? NC : eg.ncps[0], NC ? NC : eg.ncps[1], NC ? NC : eg.ncps[2]};
    const int sidx[3] = {(int)(s / ((long)eg.nsc[1] * eg.nsc[2])), (int)((s / eg.nsc[2]) % eg.nsc[1]),
                         (int)(s % eg.nsc[2])};
    int its[3], dts[3], ibase[3], dbase[3], m0[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        m0[d] = sidx[d] * nn[d];
        its[d] = nn[d] + (SHAPE == PK_TSC ? 3 : 2);
        ibase[d] = m0[d] + 1;
        dts[d] = nn[d] + (PICA_CIC ? 3 : 4);
        dbase[d] = PICA_CIC ? m0[d] + 1 : m0[d];
    }
    const int isx = its[1] * its[2] + IPAD, isy = its[2];
    const int dsx = dts[1] * dts[2] + JPAD, dsy = dts[2];
    const int itn = its[0] * isx;
    const int dtn = dts[0] * dsx;
    // every LDS pointer is smem + a byte offset, so that the compiler keeps DS instructions
    float* itile = (float*)smem;
    const int joff = (int)((sizeof(float) * 6 * itn + 7) & ~(size_t)7);
    double* jtile = (double*)(smem + joff);
    uint64_t* mask = (uint64_t*)(smem + joff + sizeof(double) * 3 * dtn);   // mover bit per particle
    int* dircnt = (int*)(mask + sp.mask_words);                                 // 27 direction counts
    int* bad = dircnt + 27;
    int* qlen = dircnt + 28;                                                    // change-queue length
    long* wpre = (long*)(dircnt + 32);                                          // movers before each mask word
    // Esirkepov only: the change queue, QCAP deposit records stored word-major (conflict-free)
    const int rec_off = (int)(((size_t)((unsigned char*)(wpre + sp.mask_words + 1) - smem) + 15) & ~(size_t)15);
    int* rec = (int*)(smem + rec_off);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    // field tile: a lane holds one (j, k) node of a slice, a wavefront walks the slices (comp, i).
    // The grid's components are equally spaced in one allocation, so a slice's address is
    // E_x + comp * pitch + offset, with no per-slice load of a component pointer.
    const int uwave = __builtin_amdgcn_readfirstlane(wave);
    {
        const float* __restrict__ f0 = g.f[PK_EX];
        const long pitch = g.f[PK_EY] - g.f[PK_EX];
        const long plane = (long)g.s1 * g.s2;
        if constexpr (NC > 0) {
            // every load of a lane is issued before its first LDS store: one memory latency per
            // workgroup instead of one per slice
            constexpr int IT = NC + (SHAPE == PK_TSC ? 3 : 2), NSL = 6 * IT, PER = (NSL + WV - 1) / WV;
            for (int jk0 = 0; jk0 < IT * IT; jk0 += 64) {
                const int jk = jk0 + lane, j = jk / IT, k = jk - j * IT;
                const bool on = jk < IT * IT;
                const long gofs = (long)(ibase[0]) * plane + (long)(ibase[1] + j) * g.s2 + (ibase[2] + k);
                float v[PER];
#pragma unroll
                for (int q = 0; q < PER; q++) {
                    const int sl = uwave + q * WV, comp = sl / IT, i = sl - comp * IT;
                    v[q] = on && sl < NSL ? f0[comp * pitch + i * plane + gofs] : 0.0f;
                }
                float* const lt = itile + j * isy + k;
#pragma unroll
                for (int q = 0; q < PER; q++) {
                    const int sl = uwave + q * WV, comp = sl / IT, i = sl - comp * IT;
                    if (on && sl < NSL) lt[comp * itn + i * isx] = v[q];
                }
            }
        } else {
            for (int jk0 = 0; jk0 < its[1] * its[2]; jk0 += 64) {
                const int jk = jk0 + lane, j = jk / its[2], k = jk - j * its[2];
                if (jk < its[1] * its[2]) {
                    const long gofs = (long)(ibase[1] + j) * g.s2 + (ibase[2] + k);
                    float* const lt = itile + j * isy + k;
                    for (int sl = uwave; sl < 6 * its[0]; sl += WV) {
                        const int comp = sl / its[0], i = sl - comp * its[0];
                        lt[comp * itn + i * isx] = f0[comp * pitch + (long)(ibase[0] + i) * plane + gofs];
                    }
                }
            }
        }
    }
    for (int t = tid; t < 3 * dtn; t += BS) jtile[t] = 0.0;
    for (int t = tid; t < 27; t += BS) dircnt[t] = 0;
    if (tid == 0) {
        bad[0] = 0;
        qlen[0] = 0;
    }
    __syncthreads();

    const long base = off[s];
    const int n = cnt[s];
    const int N1N2 = eg.ncell[1] * eg.ncell[2];
    // Each wavefront owns the 128-slot groups wave, wave + WV, ...; a lane pushes the two
    // adjacent slots c0 + 2 lane and c0 + 2 lane + 1.  No workgroup barrier until every particle
    // is pushed, so the wavefronts drift apart and one's LDS atomics overlap another's
    // arithmetic.  Every pushed particle is written back in place; movers are copied to the
    // outbox afterwards, in index order.  Esirkepov: when both particles of a lane keep their
    // node in every dimension and share a deposition window (the interleaved order puts two
    // particles of one cell in adjacent slots), their terms are summed and deposited together.
    struct PState {
        bool ok, queued, mover;
        int cls, l[3], sh[3];
        float o[3], on[3], qw;
    };
    auto push = [&](int i, PState& P) {
        P.cls = 0;
        P.queued = false;
            const bool act = i < n;
            int cell = 0;
            int c[3] = {0, 0, 0}, l[3] = {0, 0, 0}, sh[3] = {0, 0, 0};
            float o[3] = {0, 0, 0}, on[3] = {0, 0, 0}, u[3] = {0, 0, 0}, v[3] = {0, 0, 0}, w = 0.0f;
            bool ok = act;
            // ---- load, interpolate, push, move (lane-divergent)
            if (act) {
                const long p = base + i;
                cell = pm.cell[p];
                o[0] = pm.ox[p]; o[1] = pm.oy[p]; o[2] = pm.oz[p];
                u[0] = pm.ux[p]; u[1] = pm.uy[p]; u[2] = pm.uz[p];
                w = pm.w[p];
                cell_coords(cell, eg, c);
    #pragma unroll
                for (int d = 0; d < 3; d++) {
                    l[d] = c[d] - m0[d];
                    if (l[d] < 0 || l[d] >= nn[d] || !(o[d] >= 0.0f && o[d] < 1.0f)) ok = false;
                }
                ok = ok && (unsigned)cell < (unsigned)(eg.ncell[0] * N1N2);
            }
            if (ok) {
                float e[3], b[3];
                interp_fields<SHAPE>(itile, itn, isx, isy, l, o, e, b);
                boris_f32(u, e, b, sp.coeff);
                ok = isfinite(u[0]) && isfinite(u[1]) && isfinite(u[2]);
                const float invGamma = rsqrtf(1.0f + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    #pragma unroll
                for (int d = 0; d < 3; d++) v[d] = u[d] * invGamma;    // in units of c
                move_split(o, v, sp.cdt_h, on, sh);
            }
            if (act && !ok) bad[0] = 1;                     // left unchanged and reported
            // ---- deposition
            const float qw = sp.q * w;
            if (DEP == PK_ESIRKEPOV) {
                // A wavefront deposits on the window its widest lane needs.  Particles that change
                // their node in some dimension (the larger window there) go to the workgroup's change
                // queue, deposited after the loop by full wavefronts; the rest deposit now on the
                // small window.  A full queue sends the overflow back to its own wavefront.
                int cls = 0;                                  // bit d: node change in dimension d
                if (ok)
                    cls = (node_change<SHAPE>(o[0], on[0], sh[0]) ? 4 : 0) |
                          (node_change<SHAPE>(o[1], on[1], sh[1]) ? 2 : 0) | (node_change<SHAPE>(o[2], on[2], sh[2]) ? 1 : 0);
                const uint64_t cb = __ballot(cls != 0);
                bool queued = false;
                if (cb) {
                    int q0 = 0;
                    if (lane == 0) q0 = atomicAdd(qlen, __popcll(cb));
                    q0 = __shfl(q0, 0);
                    const int slot = q0 + __popcll(cb & below);
                    if (cls != 0 && slot < QCAP) {
                        queued = true;
                        rec[0 * QCAP + slot] = l[0] | (l[1] << 4) | (l[2] << 8) | ((sh[0] + 1) << 12) |
                                               ((sh[1] + 1) << 14) | ((sh[2] + 1) << 16) | (cls << 18);
                        rec[1 * QCAP + slot] = __float_as_int(o[0]);
                        rec[2 * QCAP + slot] = __float_as_int(o[1]);
                        rec[3 * QCAP + slot] = __float_as_int(o[2]);
                        rec[4 * QCAP + slot] = __float_as_int(on[0]);
                        rec[5 * QCAP + slot] = __float_as_int(on[1]);
                        rec[6 * QCAP + slot] = __float_as_int(on[2]);
                        rec[7 * QCAP + slot] = __float_as_int(qw);
                    }
                }
                P.cls = cls;
                P.queued = queued;
            } else if (ok) {
                // direct deposition of q w v at the new position: the CIC tile is pica's (base m+1,
                // the interpolation indices), the TSC tile has base m (indices + 1)
                constexpr int KK = SHAPE == PK_TSC ? 3 : 2;
                int bc[3], bs[3];
                float wc[3][3], ws[3][3];
                bool fits = true;
    #pragma unroll
                for (int d = 0; d < 3; d++) {
                    interp_stencil<SHAPE>(l[d] + sh[d], on[d], bc[d], wc[d], bs[d], ws[d]);
                    if (SHAPE == PK_TSC) {
                        bc[d] += 1;
                        bs[d] += 1;
                    }
                    if (bc[d] < 0 || bc[d] + KK > dts[d] || bs[d] < 0 || bs[d] + KK > dts[d]) fits = false;
                }
                if (fits) {
    #pragma unroll
                    for (int comp = 0; comp < 3; comp++) {
                        const float val = v[comp] * (float)PK_C * qw;
                        double* t = jtile + comp * dtn;
                        int bi[3];
                        float wt[3][3];
    #pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const bool col = colloc(comp, d);
                            bi[d] = col ? bc[d] : bs[d];
    #pragma unroll
                            for (int k = 0; k < 3; k++) wt[d][k] = col ? wc[d][k] : ws[d][k];
                        }
                        double* tb = t + bi[0] * dsx + bi[1] * dsy + bi[2];
    #pragma unroll
                        for (int a = 0; a < KK; a++)
    #pragma unroll
                            for (int bb = 0; bb < KK; bb++)
    #pragma unroll
                                for (int cc = 0; cc < KK; cc++)
                                    atomicAdd(&tb[a * dsx + bb * dsy + cc], (double)(wt[0][a] * wt[1][bb] * wt[2][cc] * val));
                    }
                } else {
                    ok = false;
                    bad[0] = 1;
                }
            }
            // ---- periodic wrap of the cell; a mover leaves the supercell
            bool mover = false;
            if (ok) {
                int cn[3];
    #pragma unroll
                for (int d = 0; d < 3; d++) {
                    cn[d] = c[d] + sh[d];
                    if (cn[d] < 0) cn[d] += eg.ncell[d];
                    if (cn[d] >= eg.ncell[d]) cn[d] -= eg.ncell[d];
                    mover = mover || cn[d] / nn[d] != sidx[d];
                }
                const long p = base + i;
                pm.cell[p] = (cn[0] * eg.ncell[1] + cn[1]) * eg.ncell[2] + cn[2];
                pm.ox[p] = on[0]; pm.oy[p] = on[1]; pm.oz[p] = on[2];
                pm.ux[p] = u[0]; pm.uy[p] = u[1]; pm.uz[p] = u[2];
            }
        P.ok = ok;
        P.mover = mover;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            P.l[d] = l[d];
            P.sh[d] = sh[d];
            P.o[d] = o[d];
            P.on[d] = on[d];
        }
        P.qw = qw;
    };
    // a particle left over by a full change queue deposits alone, on the window the widest such
    // lane of the wavefront needs
    auto overflow = [&](const PState& P) {
        const bool direct = P.ok && !P.queued && P.cls != 0;
        if (__ballot(direct)) {
            bool big[3];
#pragma unroll
            for (int d = 0; d < 3; d++) big[d] = __ballot(direct && ((P.cls >> (2 - d)) & 1)) != 0ull;
            if (direct) esirkepov_window<SHAPE>(jtile, dtn, dsx, dsy, P.l, P.sh, P.o, P.on, big, P.qw, sp.hdt);
        }
    };
    for (int c0 = 128 * wave; c0 < n; c0 += 128 * WV) {
        PState A, B;
        push(c0 + 2 * lane, A);
        push(c0 + 2 * lane + 1, B);
        const uint64_t ma = __ballot(A.mover), mbm = __ballot(B.mover);
        if (lane == 0) {
            mask[c0 >> 6] = spread_bits((uint32_t)ma) | (spread_bits((uint32_t)mbm) << 1);
            if (c0 + 64 < n) mask[(c0 >> 6) + 1] = spread_bits((uint32_t)(ma >> 32)) | (spread_bits((uint32_t)(mbm >> 32)) << 1);
        }
        if (DEP == PK_ESIRKEPOV) {
            constexpr int KS = SHAPE == PK_TSC ? 3 : 2;    // window without a node change
            overflow(A);
            overflow(B);
            const bool sa = A.ok && A.cls == 0, sb = B.ok && B.cls == 0;
            const bool small[3] = {false, false, false};
            float S0a[3][4] = {}, DSa[3][4] = {}, S0b[3][4] = {}, DSb[3][4] = {};
            int wa[3] = {0, 0, 0}, wbb[3] = {0, 0, 0};
            if (sa) esk_shapes<SHAPE>(A.l, A.sh, A.o, A.on, small, wa, S0a, DSa);
            if (sb) esk_shapes<SHAPE>(B.l, B.sh, B.o, B.on, small, wbb, S0b, DSb);
            const bool merge = sa && sb && wa[0] == wbb[0] && wa[1] == wbb[1] && wa[2] == wbb[2];
            if (sa || sb) {
                // A (with B when they share the window), or B alone when A has no small window
                int wp[3];
                float S0p[3][4], DSp[3][4], kp[3], kqq[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    wp[d] = sa ? wa[d] : wbb[d];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        S0p[d][k] = sa ? S0a[d][k] : S0b[d][k];
                        DSp[d][k] = sa ? DSa[d][k] : DSb[d][k];
                    }
                    kp[d] = -(sa ? A.qw : B.qw) * sp.hdt[d];
                    kqq[d] = merge ? -B.qw * sp.hdt[d] : 0.0f;
                }
                esirkepov_pair<KS>(jtile, dtn, dsx, dsy, wp, S0p, DSp, kp, S0b, DSb, kqq);
            }
            const bool second = sa && sb && !merge;
            if (__ballot(second)) {
                if (second) {
                    const float kb[3] = {-B.qw * sp.hdt[0], -B.qw * sp.hdt[1], -B.qw * sp.hdt[2]};
                    esirkepov<KS, KS, KS>(jtile, dtn, dsx, dsy, wbb, S0b, DSb, kb);
                }
            }
        }
    }
    __syncthreads();
    if (DEP == PK_ESIRKEPOV) {
        // the change queue, by full wavefronts
        const int nq = qlen[0] < QCAP ? qlen[0] : QCAP;
        for (int t = tid; t - lane < nq; t += BS) {
            const bool has = t < nq;
            int dl[3] = {0, 0, 0}, dsh[3] = {0, 0, 0}, dcls = 0;
            float d_o[3] = {0.5f, 0.5f, 0.5f}, d_on[3] = {0.5f, 0.5f, 0.5f}, dqw = 0.0f;
            if (has) {
                const int w0 = rec[t];
                dl[0] = w0 & 15; dl[1] = (w0 >> 4) & 15; dl[2] = (w0 >> 8) & 15;
                dsh[0] = ((w0 >> 12) & 3) - 1; dsh[1] = ((w0 >> 14) & 3) - 1; dsh[2] = ((w0 >> 16) & 3) - 1;
                dcls = (w0 >> 18) & 7;
                d_o[0] = __int_as_float(rec[1 * QCAP + t]);
                d_o[1] = __int_as_float(rec[2 * QCAP + t]);
                d_o[2] = __int_as_float(rec[3 * QCAP + t]);
                d_on[0] = __int_as_float(rec[4 * QCAP + t]);
                d_on[1] = __int_as_float(rec[5 * QCAP + t]);
                d_on[2] = __int_as_float(rec[6 * QCAP + t]);
                dqw = __int_as_float(rec[7 * QCAP + t]);
            }
            // each record on its own window: the lanes of one window shape run together (a union
            // window per wavefront made most x-crossers pay for the few y/z-crossers' window)
            bool big[3];
#pragma unroll
            for (int d = 0; d < 3; d++) big[d] = ((dcls >> (2 - d)) & 1) != 0;
            if (has) esirkepov_window<SHAPE>(jtile, dtn, dsx, dsy, dl, dsh, d_o, d_on, big, dqw, sp.hdt);
        }
    }
    // ---- movers: prefix over the mask words, copy to the outbox in index order, fill the holes
    const int words = (n + 63) >> 6;
    if (tid == 0) {
        long acc = 0;
        for (int wd = 0; wd < words; wd++) {
            wpre[wd] = acc;
            acc += __popcll(mask[wd]);
        }
        wpre[words] = acc;
    }
    __syncthreads();
    const int nmv = (int)wpre[words];
    const int nfinal = n - nmv;
    if (nmv > 0) {
        auto movers_below = [&](int y) -> long {
            const int wd = y >> 6, bt = y & 63;
            return wpre[wd] + (bt ? __popcll(mask[wd] & ((1ull << bt) - 1ull)) : 0);
        };
        // slot of the r-th mover (r < nmv): its mask word is the last one with wpre <= r
        auto mover_slot = [&](long r) -> int {
            int lo = 0, hi = words - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (wpre[mid] <= r) lo = mid; else hi = mid - 1;
            }
            return (lo << 6) + pk_select_bit(mask[lo], (int)(r - wpre[lo]));
        };
        // one lane per mover (not per slot): the t-th mover goes to outbox slot t
        for (int t = tid; t < nmv; t += BS) {
            const int i = mover_slot(t);
            const long src = base + i, q = base + t;
            const int cl = pm.cell[src];
            po.cell[q] = cl;
            po.ox[q] = pm.ox[src]; po.oy[q] = pm.oy[src]; po.oz[q] = pm.oz[src];
            po.ux[q] = pm.ux[src]; po.uy[q] = pm.uy[src]; po.uz[q] = pm.uz[src];
            po.w[q] = pm.w[src];
            int cn[3], code = 0;
            cell_coords(cl, eg, cn);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                int dd = cn[d] / nn[d] - sidx[d];
                if (dd > 1) dd -= eg.nsc[d];
                if (dd < -1) dd += eg.nsc[d];
                code = code * 3 + (dd + 1);
            }
            odir[q] = (uint8_t)code;
            atomicAdd(&dircnt[code], 1);
        }
        __syncthreads();
        // tail stayers fill the holes the movers left
        const long below_final = movers_below(nfinal);
        for (int t = tid; t < nmv; t += BS) {
            const int j = nfinal + t;
            if ((mask[j >> 6] >> (j & 63)) & 1ull) continue;
            const long r = (long)t - (movers_below(j) - below_final);   // its rank among the tail stayers
            if (r < 0 || r >= below_final) {                             // the r-th hole is the r-th mover
                bad[0] = 1;
                continue;
            }
            const long src = base + j, dst = base + mover_slot(r);
            pm.cell[dst] = pm.cell[src];
            pm.ox[dst] = pm.ox[src]; pm.oy[dst] = pm.oy[src]; pm.oz[dst] = pm.oz[src];
            pm.ux[dst] = pm.ux[src]; pm.uy[dst] = pm.uy[src]; pm.uz[dst] = pm.uz[src];
            pm.w[dst] = pm.w[src];
        }
    }
    __syncthreads();
    // J flush: a lane holds one (j, k) node of a slice, a wavefront walks the slices (comp, i)
    {
        float* const j0 = g.f[PK_JX];
        const long pitch = g.f[PK_JY] - g.f[PK_JX], plane = (long)g.s1 * g.s2;
        for (int jk0 = 0; jk0 < dts[1] * dts[2]; jk0 += 64) {
            const int jk = jk0 + lane, j = jk / dts[2], k = jk - j * dts[2];
            if (jk < dts[1] * dts[2]) {
                const long gofs = (long)dbase[0] * plane + (long)(dbase[1] + j) * g.s2 + (dbase[2] + k);
                const double* const lt = jtile + j * dsy + k;
                for (int sl = uwave; sl < 3 * dts[0]; sl += WV) {
                    const int comp = sl / dts[0], i = sl - comp * dts[0];
                    const double v = lt[comp * dtn + i * dsx];
                    if (v != 0.0) atomicAdd(&j0[comp * pitch + i * plane + gofs], (float)(v * (double)sp.norm));
                }
            }
        }
    }
    if (tid < 27) mvc[s * 27 + tid] = dircnt[tid];
    if (tid == 0) {
        nstay[s] = nfinal;
        nmove[s] = nmv;
        if (bad[0]) atomicOr(&flags[PK_FLAG_ERR], 1);
    }
}

// At least 5 wavefronts per SIMD: the LDS tiles allow 5 workgroups per CU, and the TSC
// Esirkepov body (two particles per lane) needs ~111 VGPRs, i.e. 4 wavefronts, unless the
// compiler keeps it within 96 (a few spilled dwords; 14.5 -> 14.1 ms per step at 256^3 x 16 ppc).
// BS threads per workgroup (256, or 128 for the batched form's block-shape knob)
template <int SHAPE, int DEP, int NC, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(5))) k_particles_split(SplitArgs A)
{
    const long s = pk_xcd_supercell(blockIdx.x, A.eg.xcd_chunk);
    if (s >= A.eg.nsct) return;
    particles_split_sc<SHAPE, DEP, NC, BS>(A, s);
}

// many ensembles: a list of (ensemble, supercell) items; gridDim.x < nitems makes it persistent
template <int SHAPE, int DEP, int NC, int BS>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(5))) k_particles_split_batch(const SplitArgs* __restrict__ jobs,
                                                                    const int2* __restrict__ items, long nitems,
                                                                    const int* __restrict__ halt)
{
    if (halt[PK_HALT_STOP]) return;
    for (long w = blockIdx.x; w < nitems; w += gridDim.x) {
        __syncthreads();
        const int2 it = items[w];
        particles_split_sc<SHAPE, DEP, NC, BS>(jobs[it.x], it.y);     // arguments read in place (a copy would live in scratch)
    }
}

// ---------------------------------------------------------------------------------------------
// Esirkepov step with per-cell register accumulation (cubic supercells of NC cells).
//
// The workgroup stages its supercell's particles in LDS (coalesced), counting-sorts their slot
// indices by cell, and gives every lane a run of consecutive particles in cell order.  A particle
// that stays in its cell deposits on the fixed window of the 3 cell-centre nodes c-1, c, c+1 in
// every dimension; the lane sums those 54 current contributions in registers and adds them to the
// FP64 LDS tile once per cell it visits.  A particle that changes cell is queued in LDS and
// deposited afterwards by full wavefronts on the wavefront-uniform windows of esirkepov_window.
// The new state goes back to the particle's own slot (leavers to the outbox in slot order, tail
// stayers into holes), so storage order and the re-sort rule are those of k_particles_split.
// ---------------------------------------------------------------------------------------------
constexpr int HQ_WORDS = 4;    // queue entry: slot | local cell << 16 | (shift + 1) bits << 24, o[3]

// shape of a particle staying in its cell on the window of the cell-centre nodes c-1, c, c+1:
// TSC's quadratic spline, or CIC's linear one (on the lower two nodes when o < 1/2)
template <int SHAPE>
__device__ __forceinline__ void stay_shape(float o, float S[3])
{
    if (SHAPE == PK_TSC) {
        tsc_w(o - 0.5f, S);
    } else {
        const bool lo = o < 0.5f;
        const float f = lo ? o + 0.5f : o - 0.5f;
        S[0] = lo ? 1.0f - f : 0.0f;
        S[1] = lo ? f : 1.0f - f;
        S[2] = lo ? 0.0f : f;
    }
}

// Esirkepov contributions (same decomposition as esirkepov()) of a staying particle, summed into
// acc[comp * 18 + face * 9 + p * 3 + r] for the two inner faces of the current component
template <int SHAPE>
__device__ __forceinline__ void accumulate_stay(float acc[54], const float o[3], const float on[3], float qw,
                                                const float hdt[3])
{
    float S0[3][3], DS[3][3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        float S1[3];
        stay_shape<SHAPE>(o[d], S0[d]);
        stay_shape<SHAPE>(on[d], S1);
#pragma unroll
        for (int k = 0; k < 3; k++) DS[d][k] = S1[k] - S0[d][k];
    }
    const float third = 1.0f / 3.0f;
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
        const int d1 = (comp + 1) % 3, d2 = (comp + 2) % 3;
        const float kq = -qw * hdt[comp];
#pragma unroll
        for (int p = 0; p < 3; p++) {
            const float A = kq * (S0[d1][p] + 0.5f * DS[d1][p]);
            const float B = kq * (0.5f * S0[d1][p] + third * DS[d1][p]);
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const float coef = A * S0[d2][r] + B * DS[d2][r];
                const float c0 = DS[comp][0] * coef;
                acc[comp * 18 + p * 3 + r] += c0;
                acc[comp * 18 + 9 + p * 3 + r] += c0 + DS[comp][1] * coef;
            }
        }
    }
}

// add a lane's register sums for local cell lc to the LDS J tile: window base l + 1 in the base-m
// deposition tile (DTS per side), faces one further along the current component
template <int NC, int DTS>
__device__ __forceinline__ void flush_stay(double* jt, const float acc[54], int lc)
{
    constexpr int DTN = DTS * DTS * DTS;
    constexpr int st[3] = {DTS * DTS, DTS, 1};
    const int l0 = lc / (NC * NC), l1 = (lc / NC) % NC, l2 = lc % NC;
    const int orig = (l0 + 1) * st[0] + (l1 + 1) * st[1] + (l2 + 1);
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
        const int d1 = (comp + 1) % 3, d2 = (comp + 2) % 3;
        double* t = jt + comp * DTN + orig + st[comp];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int r = 0; r < 3; r++)
                    atomicAdd(&t[a * st[comp] + p * st[d1] + r * st[d2]],
                              (double)acc[comp * 18 + a * 9 + p * 3 + r]);
    }
}

// LDS layout of k_esk_sorted (byte offsets; scap is a multiple of 8)
template <int SHAPE, int NC>
struct SortedLayout {
    static constexpr int NCELLS = NC * NC * NC;
    static constexpr int ITS = NC + (SHAPE == PK_TSC ? 3 : 2), DTS = NC + 4;
    static constexpr int ITN = ITS * ITS * ITS, DTN = DTS * DTS * DTS;
    int scap, hq_cap, mask_words;
    __host__ __device__ int o_itile() const { return 8 * 3 * DTN; }
    __host__ __device__ int o_stage() const { return (o_itile() + 4 * 6 * ITN + 7) & ~7; }
    __host__ __device__ int o_perm() const { return o_stage() + 28 * scap; }       // cell + 6 floats
    __host__ __device__ int o_lcell() const { return o_perm() + 2 * scap; }
    __host__ __device__ int o_code() const { return o_lcell() + scap; }
    __host__ __device__ int o_cnt() const { return o_code() + scap; }              // 2 x (NCELLS + 2) ints
    __host__ __device__ int o_hq() const { return o_cnt() + 8 * (NCELLS + 2); }
    __host__ __device__ int o_mask() const { return (o_hq() + 4 * HQ_WORDS * hq_cap + 7) & ~7; }
    __host__ __device__ int o_misc() const { return o_mask() + 8 * mask_words; }    // 40 ints
    __host__ __device__ int o_wpre() const { return o_misc() + 4 * 40; }
    __host__ __device__ int bytes() const { return o_wpre() + 8 * (mask_words + 1); }
};

template <int SHAPE, int NC>
__global__ void __launch_bounds__(PK_BLOCK) k_esk_sorted(GridView<float> g, EnsGeom eg, PartSplit pm, PartSplit po,
                                                         uint8_t* __restrict__ odir, const int64_t* __restrict__ off,
                                                         const int32_t* __restrict__ cnt, int32_t* __restrict__ nstay,
                                                         int32_t* __restrict__ nmove, int32_t* __restrict__ mvc,
                                                         int32_t* __restrict__ flags, SplitParams sp,
                                                         SortedLayout<SHAPE, NC> L)
{
    typedef SortedLayout<SHAPE, NC> SL;
    constexpr int NCELLS = SL::NCELLS, ITS = SL::ITS, DTS = SL::DTS, ITN = SL::ITN, DTN = SL::DTN;
    constexpr int ISX = ITS * ITS, ISY = ITS, DSX = DTS * DTS, DSY = DTS;
    extern __shared__ __align__(16) unsigned char smem[];
    const long s = pk_xcd_supercell(blockIdx.x, eg.xcd_chunk);
    if (s >= eg.nsct) return;
    const int scap = L.scap;
    double* jtile = (double*)smem;
    float* itile = (float*)(smem + L.o_itile());
    int* st_cell = (int*)(smem + L.o_stage());
    float* st_f = (float*)(smem + L.o_stage() + 4 * scap);     // ox oy oz ux uy uz, scap each
    unsigned short* perm = (unsigned short*)(smem + L.o_perm());
    uint8_t* lcell = smem + L.o_lcell();
    uint8_t* code = smem + L.o_code();
    int* cellcnt = (int*)(smem + L.o_cnt());
    int* cellcur = (int*)(smem + L.o_cnt() + 4 * (NCELLS + 2));
    int* hq = (int*)(smem + L.o_hq());
    uint64_t* mask = (uint64_t*)(smem + L.o_mask());
    int* wtot = (int*)(smem + L.o_misc());          // WAVES
    int* dircnt = wtot + WAVES;                      // 27
    int* bad = dircnt + 27;
    int* hqn = bad + 1;
    long* wpre = (long*)(smem + L.o_wpre());

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const long base = off[s];
    const int n = cnt[s];
    const int sidx[3] = {(int)(s / ((long)eg.nsc[1] * eg.nsc[2])), (int)((s / eg.nsc[2]) % eg.nsc[1]),
                         (int)(s % eg.nsc[2])};
    const int m0[3] = {sidx[0] * NC, sidx[1] * NC, sidx[2] * NC};
    if (n > scap) {      // the host sizes scap from the largest capacity; never taken
        if (tid < 27) mvc[s * 27 + tid] = 0;
        if (tid == 0) {
            nstay[s] = n;
            nmove[s] = 0;
            atomicOr(&flags[PK_FLAG_ERR], 1);
        }
        return;
    }
    const int N1N2 = eg.ncell[1] * eg.ncell[2];

    // ---- phase 0: field stencil, zeroed tiles and counters, particles staged (coalesced)
    for (int t = tid; t < 6 * ITN; t += PK_BLOCK) {
        const int comp = t / ITN, r = t % ITN;
        const int i = r / ISX, j = (r / ISY) % ITS, k = r % ISY;
        itile[t] = g.f[PK_EX + comp][((long)(m0[0] + 1 + i) * g.s1 + (m0[1] + 1 + j)) * g.s2 + (m0[2] + 1 + k)];
    }
    for (int t = tid; t < 3 * DTN; t += PK_BLOCK) jtile[t] = 0.0;
    for (int t = tid; t < NCELLS + 1; t += PK_BLOCK) cellcnt[t] = 0;
    if (tid < 27) dircnt[tid] = 0;
    if (tid == 0) {
        bad[0] = 0;
        hqn[0] = 0;
    }
    for (int i = tid; i < n; i += PK_BLOCK) {
        const long p = base + i;
        st_cell[i] = pm.cell[p];
        st_f[i] = pm.ox[p];
        st_f[scap + i] = pm.oy[p];
        st_f[2 * scap + i] = pm.oz[p];
        st_f[3 * scap + i] = pm.ux[p];
        st_f[4 * scap + i] = pm.uy[p];
        st_f[5 * scap + i] = pm.uz[p];
        code[i] = 13;
    }
    __syncthreads();

    // ---- phase 1: counting sort of the slots by local cell (NCELLS = not in this supercell)
    for (int i = tid; i < n; i += PK_BLOCK) {
        const int c = st_cell[i];
        int cc[3];
        cell_coords(c, eg, cc);
        const int l0 = cc[0] - m0[0], l1 = cc[1] - m0[1], l2 = cc[2] - m0[2];
        const bool in = (unsigned)c < (unsigned)(eg.ncell[0] * N1N2) && (unsigned)l0 < (unsigned)NC &&
                        (unsigned)l1 < (unsigned)NC && (unsigned)l2 < (unsigned)NC;
        const int lc = in ? (l0 * NC + l1) * NC + l2 : NCELLS;
        lcell[i] = (uint8_t)lc;
        atomicAdd(&cellcnt[lc], 1);
    }
    __syncthreads();
    if (wave == 0) {      // exclusive scan of the NCELLS + 1 counters by one wavefront
        constexpr int PER = (NCELLS + 1 + 63) / 64;
        int v[PER], sum = 0;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int idx = lane * PER + k;
            v[k] = idx <= NCELLS ? cellcnt[idx] : 0;
            sum += v[k];
        }
        int incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        int run = incl - sum;
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int idx = lane * PER + k;
            if (idx <= NCELLS) cellcur[idx] = run;
            run += v[k];
        }
    }
    __syncthreads();
    for (int i = tid; i < n; i += PK_BLOCK) perm[atomicAdd(&cellcur[lcell[i]], 1)] = (unsigned short)i;
    __syncthreads();

    // ---- phase 2: every lane walks its run of particles in cell order
    constexpr int KS = SHAPE == PK_TSC ? 3 : 2;
    const int r0 = (int)(((long)tid * n) / PK_BLOCK), r1 = (int)(((long)(tid + 1) * n) / PK_BLOCK);
    float acc[54];
#pragma unroll
    for (int k = 0; k < 54; k++) acc[k] = 0.0f;
    int cur = -1;
    for (int r = r0; r < r1; r++) {
        const int i = perm[r];
        const int lc = lcell[i];
        if (lc == NCELLS) {          // outside this supercell: left unchanged, reported
            bad[0] = 1;
            continue;
        }
        const int l[3] = {lc / (NC * NC), (lc / NC) % NC, lc % NC};
        const float o[3] = {st_f[i], st_f[scap + i], st_f[2 * scap + i]};
        float u[3] = {st_f[3 * scap + i], st_f[4 * scap + i], st_f[5 * scap + i]};
        if (!(o[0] >= 0.0f && o[0] < 1.0f && o[1] >= 0.0f && o[1] < 1.0f && o[2] >= 0.0f && o[2] < 1.0f)) {
            bad[0] = 1;
            continue;
        }
        float e[3], b[3];
        interp_fields<SHAPE>(itile, ITN, ISX, ISY, l, o, e, b);
        boris_f32(u, e, b, sp.coeff);
        if (!(isfinite(u[0]) && isfinite(u[1]) && isfinite(u[2]))) {
            bad[0] = 1;
            continue;
        }
        const float invGamma = rsqrtf(1.0f + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        const float v[3] = {u[0] * invGamma, u[1] * invGamma, u[2] * invGamma};
        float on[3];
        int sh[3];
        move_split(o, v, sp.cdt_h, on, sh);
        if (sh[0] == 0 && sh[1] == 0 && sh[2] == 0) {
            if (lc != cur) {
                if (cur >= 0) flush_stay<NC, DTS>(jtile, acc, cur);
#pragma unroll
                for (int k = 0; k < 54; k++) acc[k] = 0.0f;
                cur = lc;
            }
            accumulate_stay<SHAPE>(acc, o, on, sp.q * pm.w[base + i], sp.hdt);
        } else {
            const int q = atomicAdd(hqn, 1);
            if (q < L.hq_cap) {
                int* en = hq + q * HQ_WORDS;
                en[0] = i | (lc << 16) | ((sh[0] + 1) << 24) | ((sh[1] + 1) << 26) | ((sh[2] + 1) << 28);
                en[1] = __float_as_int(o[0]);
                en[2] = __float_as_int(o[1]);
                en[3] = __float_as_int(o[2]);
            } else {         // queue full: this lane deposits on the largest window now
                const bool big[3] = {true, true, true};
                esirkepov_window<SHAPE>(jtile, DTN, DSX, DSY, l, sh, o, on, big, sp.q * pm.w[base + i], sp.hdt);
            }
        }
        // new state in the staging slot; direction class of the supercell it lands in
        int cls = 0, cn[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int x = l[d] + sh[d];      // -1 .. NC
            cn[d] = m0[d] + x;
            if (cn[d] < 0) cn[d] += eg.ncell[d];
            if (cn[d] >= eg.ncell[d]) cn[d] -= eg.ncell[d];
            int ns = sidx[d] + (x < 0 ? -1 : (x >= NC ? 1 : 0));
            if (ns < 0) ns += eg.nsc[d];
            if (ns >= eg.nsc[d]) ns -= eg.nsc[d];
            int dd = ns - sidx[d];
            if (dd > 1) dd -= eg.nsc[d];
            if (dd < -1) dd += eg.nsc[d];
            cls = cls * 3 + (dd + 1);
        }
        st_cell[i] = (cn[0] * eg.ncell[1] + cn[1]) * eg.ncell[2] + cn[2];
        st_f[i] = on[0];
        st_f[scap + i] = on[1];
        st_f[2 * scap + i] = on[2];
        st_f[3 * scap + i] = u[0];
        st_f[4 * scap + i] = u[1];
        st_f[5 * scap + i] = u[2];
        code[i] = (uint8_t)cls;
    }
    if (cur >= 0) flush_stay<NC, DTS>(jtile, acc, cur);
    __syncthreads();

    // ---- phase 3: cell changers by full wavefronts
    const int nq = hqn[0] < L.hq_cap ? hqn[0] : L.hq_cap;
    for (int q0 = 0; q0 < nq; q0 += PK_BLOCK) {
        const int q = q0 + tid;
        const bool ok = q < nq;
        int i = 0, lc = 0, sh[3] = {0, 0, 0};
        float o[3] = {0.5f, 0.5f, 0.5f}, on[3] = {0.5f, 0.5f, 0.5f};
        if (ok) {
            const int* en = hq + q * HQ_WORDS;
            i = en[0] & 0xffff;
            lc = (en[0] >> 16) & 0xff;
            sh[0] = ((en[0] >> 24) & 3) - 1;
            sh[1] = ((en[0] >> 26) & 3) - 1;
            sh[2] = ((en[0] >> 28) & 3) - 1;
            o[0] = __int_as_float(en[1]);
            o[1] = __int_as_float(en[2]);
            o[2] = __int_as_float(en[3]);
            on[0] = st_f[i];
            on[1] = st_f[scap + i];
            on[2] = st_f[2 * scap + i];
        }
        const int l[3] = {lc / (NC * NC), (lc / NC) % NC, lc % NC};
        bool big[3];
#pragma unroll
        for (int d = 0; d < 3; d++) big[d] = __ballot(ok && node_change<SHAPE>(o[d], on[d], sh[d])) != 0ull;
        if (ok) esirkepov_window<SHAPE>(jtile, DTN, DSX, DSY, l, sh, o, on, big, sp.q * pm.w[base + i], sp.hdt);
    }

    // ---- phase 4: leavers to the outbox in slot order
    int running = 0;
    for (int c0 = 0; c0 < n; c0 += PK_BLOCK) {
        const int i = c0 + tid;
        const bool act = i < n;
        const int cls = act ? code[i] : 13;
        const bool mover = cls != 13;
        const uint64_t mb = __ballot(mover);
        if (lane == 0) {
            wtot[wave] = __popcll(mb);
            mask[(c0 >> 6) + wave] = mb;
        }
        __syncthreads();
        int pre = running, tot = 0;
#pragma unroll
        for (int vv = 0; vv < WAVES; vv++) {
            pre += (vv < wave) ? wtot[vv] : 0;
            tot += wtot[vv];
        }
        if (mover) {
            const long q = base + pre + __popcll(mb & ((1ull << lane) - 1ull));
            po.cell[q] = st_cell[i];
            po.ox[q] = st_f[i];
            po.oy[q] = st_f[scap + i];
            po.oz[q] = st_f[2 * scap + i];
            po.ux[q] = st_f[3 * scap + i];
            po.uy[q] = st_f[4 * scap + i];
            po.uz[q] = st_f[5 * scap + i];
            po.w[q] = pm.w[base + i];
            odir[q] = (uint8_t)cls;
            atomicAdd(&dircnt[cls], 1);
        }
        running += tot;
        __syncthreads();
    }
    const int nmv = running;
    const int nfinal = n - nmv;
    const int words = (n + 63) >> 6;
    if (tid == 0) {
        long a = 0;
        for (int wd = 0; wd < words; wd++) {
            wpre[wd] = a;
            a += __popcll(mask[wd]);
        }
        wpre[words] = a;
    }
    __syncthreads();
    // ---- stayers back to HBM: in place below nfinal, tail stayers into the holes in order
    auto movers_below = [&](int y) -> long {
        const int wd = y >> 6, bt = y & 63;
        return wpre[wd] + (bt ? __popcll(mask[wd] & ((1ull << bt) - 1ull)) : 0);
    };
    const long below_final = movers_below(nfinal);
    for (int i = tid; i < n; i += PK_BLOCK) {
        if ((mask[i >> 6] >> (i & 63)) & 1ull) continue;
        long dst = i;
        const bool tail = i >= nfinal;
        if (tail) {
            const long r = (long)(i - nfinal) - (movers_below(i) - below_final);
            int lo = 0, hi = words - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (wpre[mid] <= r) lo = mid; else hi = mid - 1;
            }
            uint64_t m = mask[lo];
            for (long z = r - wpre[lo]; z > 0; z--) m &= m - 1ull;
            if (m == 0ull) {
                bad[0] = 1;
                continue;
            }
            dst = (lo << 6) + (__ffsll((long long)m) - 1);
        }
        const long p = base + dst;
        const float w = tail ? pm.w[base + i] : 0.0f;
        pm.cell[p] = st_cell[i];
        pm.ox[p] = st_f[i];
        pm.oy[p] = st_f[scap + i];
        pm.oz[p] = st_f[2 * scap + i];
        pm.ux[p] = st_f[3 * scap + i];
        pm.uy[p] = st_f[4 * scap + i];
        pm.uz[p] = st_f[5 * scap + i];
        if (tail) pm.w[p] = w;
    }
    __syncthreads();
    // ---- J tile to HBM (FP32 global atomics of the FP64 sums times 1/V)
    for (int t = tid; t < 3 * DTN; t += PK_BLOCK) {
        const double v = jtile[t];
        if (v == 0.0) continue;
        const int comp = t / DTN, r = t % DTN;
        const int i = r / DSX, j = (r / DSY) % DTS, k = r % DSY;
        atomicAdd(&g.f[PK_JX + comp][((long)(m0[0] + i) * g.s1 + (m0[1] + j)) * g.s2 + (m0[2] + k)],
                  (float)(v * (double)sp.norm));
    }
    if (tid < 27) mvc[s * 27 + tid] = dircnt[tid];
    if (tid == 0) {
        nstay[s] = nfinal;
        nmove[s] = nmv;
        if (bad[0]) atomicOr(&flags[PK_FLAG_ERR], 1);
    }
}

// dynamic LDS of k_particles_split: field tile, FP64 J tile, mover mask, counters, mask prefix
size_t split_lds(const pk_ens* e, int shape, int dep, int mask_words)
{
    int its[3], dts[3];
    for (int d = 0; d < 3; d++) {
        its[d] = e->ncps[d] + (shape == PK_TSC ? 3 : 2);
        dts[d] = e->ncps[d] + ((shape == PK_CIC && dep == PK_DIRECT) ? 3 : 4);
    }
    const size_t joff = (sizeof(float) * 6 * its[0] * (its[1] * its[2] + IPAD) + 7) & ~(size_t)7;
    const size_t base = joff + sizeof(double) * 3 * dts[0] * (dts[1] * dts[2] + JPAD) + sizeof(uint64_t) * mask_words +
                        sizeof(int) * 32 + sizeof(long) * (mask_words + 1);
    if (dep != PK_ESIRKEPOV) return base + 16;
    return ((base + 15) & ~(size_t)15) + sizeof(int) * REC_WORDS * QCAP + 16;
}

SplitArgs split_args(pk_ens* e, pk_grid* g, const EnsGeom& eg, const SplitParams& sp)
{
    SplitArgs A;
    A.g = grid_view<float>(g);
    A.eg = eg;
    A.pm = part_split(e, e->main);
    A.po = part_split(e, e->outbox);
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

int split_params(pk_ens* e, double dt, SplitParams& sp)
{
    const double q = e->ctx->charge[e->type], m = e->ctx->mass[e->type];
    sp.coeff = (float)(q * dt / (2.0 * m * PK_C));
    sp.q = (float)q;
    for (int d = 0; d < 3; d++) {
        sp.cdt_h[d] = (float)(PK_C * dt / e->step[d]);
        sp.hdt[d] = (float)(e->step[d] / dt);
        if (PK_C * dt / e->step[d] >= 1.0)
            return pk_fail(e->ctx, PK_E_ARG, "split-position step needs c dt < h in every dimension");
    }
    sp.norm = (float)(1.0 / (e->step[0] * e->step[1] * e->step[2]));
    sp.mask_words = (e->maxcap + PK_BLOCK + 63) / 64 + 1;
    return PK_OK;
}

template <int SHAPE, int DEP, int NC>
int launch(pk_ens* e, pk_grid* g, double dt)
{
    EnsGeom eg = ens_geom(e);
    SplitParams sp;
    if (int r = split_params(e, dt, sp)) return r;
    const size_t lds = split_lds(e, SHAPE, DEP, sp.mask_words);
    if (lds > 160 * 1024) return pk_fail(e->ctx, PK_E_CAPACITY, "particle step: supercell too large for LDS");
    const unsigned nblocks = (unsigned)(PK_XCD * (long)eg.xcd_chunk);
    pk_ktimer tm(e->ctx, "particles");
    const SplitArgs A = split_args(e, g, eg, sp);
    hipLaunchKernelGGL((k_particles_split<SHAPE, DEP, NC, PK_BLOCK>), dim3(nblocks), dim3(PK_BLOCK), lds,
                       e->ctx->stream, A);
    PK_HIP(e->ctx, hipGetLastError());
    return PK_OK;
}

constexpr int LDS_PER_CU = 160 * 1024;

// LDS layout of the cell-sorted kernel for this ensemble; the change queue takes what is left of
// `budget` bytes (at most one entry per staged particle, at most the context's cap when one is
// set, else at least 64 entries).  False if it does not fit.
template <int SHAPE, int NC>
bool sorted_layout(const pk_ens* e, int mask_words, int budget, SortedLayout<SHAPE, NC>& L)
{
    const int qcap = e->ctx->opt[PK_OPT_CHANGE_QUEUE];
    L.scap = (e->maxcap + 7) & ~7;
    L.mask_words = mask_words;
    L.hq_cap = 0;
    const int rest = L.bytes() + 8;
    L.hq_cap = (budget - rest) / (4 * HQ_WORDS);
    if (L.hq_cap > L.scap) L.hq_cap = L.scap;
    if (qcap > 0 && L.hq_cap > qcap) L.hq_cap = qcap;
    if (L.hq_cap < 1 || (qcap == 0 && L.hq_cap < 64 && L.hq_cap < L.scap)) return false;
    return L.bytes() <= budget;
}

template <int SHAPE, int NC>
int launch_sorted(pk_ens* e, pk_grid* g, double dt, const SortedLayout<SHAPE, NC>& L)
{
    EnsGeom eg = ens_geom(e);
    SplitParams sp;
    if (int r = split_params(e, dt, sp)) return r;
    if (sp.mask_words != L.mask_words) return pk_fail(e->ctx, PK_E_STATE, "particle step: stale LDS layout");
    const unsigned nblocks = (unsigned)(PK_XCD * (long)eg.xcd_chunk);
    PK_HIP(e->ctx, hipFuncSetAttribute((const void*)k_esk_sorted<SHAPE, NC>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, L.bytes()));
    pk_ktimer tm(e->ctx, "particles");
    hipLaunchKernelGGL((k_esk_sorted<SHAPE, NC>), dim3(nblocks), dim3(PK_BLOCK), (size_t)L.bytes(), e->ctx->stream,
                       grid_view<float>(g), eg, part_split(e, e->main), part_split(e, e->outbox), e->odir, e->off,
                       e->cnt, e->nstay, e->nmove, e->mvc, e->flags, sp, L);
    PK_HIP(e->ctx, hipGetLastError());
    return PK_OK;
}

// Esirkepov: the cell-sorted kernel when asked for, else the per-particle kernel.  Measured on
// MI355X (256^3, 16 ppc, TSC, 4^3 supercells) the cell-sorted kernel issues ~4x fewer LDS
// atomics but runs at 2 workgroups per CU (its LDS staging) and is 1.5x slower, so it is not
// the default.
template <int SHAPE, int NC>
int launch_esirkepov(pk_ens* e, pk_grid* g, double dt)
{
    if (e->ctx->opt[PK_OPT_PARTICLE_KERNEL] == PK_KERNEL_CELL_SORTED) {
        const int mask_words = (e->maxcap + PK_BLOCK + 63) / 64 + 1;
        SortedLayout<SHAPE, NC> L;
        if (!sorted_layout<SHAPE, NC>(e, mask_words, LDS_PER_CU, L))
            return pk_fail(e->ctx, PK_E_CAPACITY, "cell-sorted particle kernel: supercell too large for LDS");
        return launch_sorted<SHAPE, NC>(e, g, dt, L);
    }
    return launch<SHAPE, PK_ESIRKEPOV, NC>(e, g, dt);
}

template <int SHAPE, int DEP>
int launch_nc(pk_ens* e, pk_grid* g, double dt)
{
    const bool cube = e->ncps[0] == e->ncps[1] && e->ncps[1] == e->ncps[2];
    if (DEP == PK_ESIRKEPOV) {
        if (cube && e->ncps[0] == 4) return launch_esirkepov<SHAPE, 4>(e, g, dt);
        if (cube && e->ncps[0] == 2) return launch_esirkepov<SHAPE, 2>(e, g, dt);
        if (e->ctx->opt[PK_OPT_PARTICLE_KERNEL] == PK_KERNEL_CELL_SORTED)
            return pk_fail(e->ctx, PK_E_UNSUPPORTED, "cell-sorted particle kernel: needs 2^3 or 4^3 supercells");
    }
    if (cube && e->ncps[0] == 4) return launch<SHAPE, DEP, 4>(e, g, dt);
    if (cube && e->ncps[0] == 2) return launch<SHAPE, DEP, 2>(e, g, dt);
    if (cube && e->ncps[0] == 8) return launch<SHAPE, DEP, 8>(e, g, dt);
    return launch<SHAPE, DEP, 0>(e, g, dt);
}

// the items of one form factor (SHAPE) in one launch
template <int SHAPE, int DEP, int NC>
int batch_go(pk_batch* b, bool prepare)
{
    const bool small = b->knob[PK_BATCH_BLOCK] == 128;
    const void* fn = small ? (const void*)k_particles_split_batch<SHAPE, DEP, NC, 128>
                           : (const void*)k_particles_split_batch<SHAPE, DEP, NC, PK_BLOCK>;
    const size_t lds = b->cplds[SHAPE];
    const long nitems = b->cnitems[SHAPE];
    if (nitems == 0) return PK_OK;
    if (prepare) {
        PK_HIP(b->ctx, hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        return PK_OK;
    }
    const long wg = b->knob[PK_BATCH_WG_PER_CU];
    const long nb = wg > 0 && wg * 256 < nitems ? wg * 256 : nitems;
    if (small)
        hipLaunchKernelGGL((k_particles_split_batch<SHAPE, DEP, NC, 128>), dim3((unsigned)nb), dim3(128), lds,
                           b->ctx->stream, (const SplitArgs*)b->pjobs, b->citems[SHAPE], nitems, b->halt);
    else
        hipLaunchKernelGGL((k_particles_split_batch<SHAPE, DEP, NC, PK_BLOCK>), dim3((unsigned)nb), dim3(PK_BLOCK),
                           lds, b->ctx->stream, (const SplitArgs*)b->pjobs, b->citems[SHAPE], nitems, b->halt);
    PK_HIP(b->ctx, hipGetLastError());
    return PK_OK;
}

template <int SHAPE, int DEP>
int batch_go_nc(pk_batch* b, bool prepare)
{
    const pk_ens* e = b->ens[0];
    const bool cube = e->ncps[0] == e->ncps[1] && e->ncps[1] == e->ncps[2];
    if (cube && e->ncps[0] == 4) return batch_go<SHAPE, DEP, 4>(b, prepare);
    if (cube && e->ncps[0] == 2) return batch_go<SHAPE, DEP, 2>(b, prepare);
    if (cube && e->ncps[0] == 8) return batch_go<SHAPE, DEP, 8>(b, prepare);
    return batch_go<SHAPE, DEP, 0>(b, prepare);
}

// every form factor present, one launch each
int batch_dispatch(pk_batch* b, int dep, bool prepare)
{
    int r;
    if (dep == PK_DIRECT) {
        if ((r = batch_go_nc<PK_CIC, PK_DIRECT>(b, prepare))) return r;
        return batch_go_nc<PK_TSC, PK_DIRECT>(b, prepare);
    }
    if (dep == PK_ESIRKEPOV) {
        if ((r = batch_go_nc<PK_CIC, PK_ESIRKEPOV>(b, prepare))) return r;
        return batch_go_nc<PK_TSC, PK_ESIRKEPOV>(b, prepare);
    }
    return pk_fail(b->ctx, PK_E_ARG, "particle step: unknown depositor");
}

}  // namespace

int pk_particles_split(pk_ens* e, pk_grid* g, double dt, int shape, int dep)
{
    if (shape == PK_CIC && dep == PK_DIRECT) return launch_nc<PK_CIC, PK_DIRECT>(e, g, dt);
    if (shape == PK_CIC && dep == PK_ESIRKEPOV) return launch_nc<PK_CIC, PK_ESIRKEPOV>(e, g, dt);
    if (shape == PK_TSC && dep == PK_DIRECT) return launch_nc<PK_TSC, PK_DIRECT>(e, g, dt);
    if (shape == PK_TSC && dep == PK_ESIRKEPOV) return launch_nc<PK_TSC, PK_ESIRKEPOV>(e, g, dt);
    return pk_fail(e->ctx, PK_E_ARG, "particle step: unknown shape/depositor");
}

int pk_batch_split_prepare(pk_batch* b, double dt, int shape, int dep)
{
    const int n = b->n();
    std::vector<SplitArgs> jobs(n);
    b->cplds[0] = b->cplds[1] = 0;
    for (int i = 0; i < n; i++) {
        pk_ens* e = b->ens[i];
        SplitParams sp;
        if (int r = split_params(e, dt, sp)) return r;
        jobs[i] = split_args(e, b->grids[i], ens_geom(e), sp);
        const int sh = pk_batch_shape_of(b, i, shape);
        const size_t l = split_lds(e, sh, dep, sp.mask_words);
        if (l > b->cplds[sh]) b->cplds[sh] = l;
    }
    if (b->cplds[0] > 160 * 1024 || b->cplds[1] > 160 * 1024)
        return pk_fail(b->ctx, PK_E_CAPACITY, "particle step: supercell too large for LDS");
    if (int r = pk_batch_upload(b, &b->pjobs, jobs.data(), sizeof(SplitArgs) * n)) return r;
    return batch_dispatch(b, dep, true);
}

int pk_batch_split_launch(pk_batch* b, int dep) { return batch_dispatch(b, dep, false); }
