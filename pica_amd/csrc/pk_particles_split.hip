// pk_particles_split.hip -- FP32 split-position particle step (gfx950).
//
// A particle is an int32 linear physical cell index, a float offset in [0, 1) per dimension,
// float u = p/(mc) and a float weight (32 bytes).  Every stencil index is exact small-integer
// arithmetic relative to the particle's cell; only the in-cell offset carries rounding, so
// the position resolution is 2^-24 of a cell everywhere in the domain.  Fields, J and the LDS
// tiles are FP32.  The kernel is the FP64 kernel's pipeline (pk_particles_f64.hip): stencil
// in LDS -> interpolate -> Boris push (BorisPusher.h:32-47) -> deposit into an LDS J tile ->
// movers to the outbox, tail stayers fill holes -> one flush of global atomics.
//
// Stencil conventions (grid origin -2h, two ghost layers; l = cell - supercell start):
//   interpolation tile base = array index m+1 (pica's baseIdx), CIC n+2 / TSC n+3 wide;
//   a staggered node of cell c sits at tile index l+1, the collocated node c at tile l+1;
//   deposition tile base = array index m (n+4 wide) except pica's CIC direct (m+1, n+3).
#include "pk_internal.h"

namespace {

constexpr int WAVES = PK_BLOCK / 64;

__device__ __forceinline__ int tix(const int ts[3], int i, int j, int k) { return (i * ts[1] + j) * ts[2] + k; }

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
__device__ __forceinline__ float gather(const float* t, const int ts[3], const int b[3], const float w[3][3])
{
    constexpr int K = SHAPE == PK_TSC ? 3 : 2;
    float sa = 0.0f;
#pragma unroll
    for (int a = 0; a < K; a++) {
        float sb = 0.0f;
#pragma unroll
        for (int bb = 0; bb < K; bb++) {
            const float* row = t + tix(ts, b[0] + a, b[1] + bb, b[2]);
            float sc = 0.0f;
#pragma unroll
            for (int c = 0; c < K; c++) sc += w[2][c] * row[c];
            sb += w[1][bb] * sc;
        }
        sa += w[0][a] * sb;
    }
    return sa;
}

// shape of the cell-centre (rho) grid on a K-node window [w, w+K) for a particle in tile
// coordinates: node index of its cell centre nc = l + 2 (deposition tile base m), offset o
template <int SHAPE>
__device__ __forceinline__ void rho_shape(int nc, float o, int w, float S[4])
{
    S[0] = S[1] = S[2] = S[3] = 0.0f;
    if (SHAPE == PK_TSC) {
        float ww[3];
        tsc_w(o - 0.5f, ww);          // nodes nc-1, nc, nc+1
        const int a = nc - 1 - w;
#pragma unroll
        for (int k = 0; k < 4; k++) S[k] = (k == a) ? ww[0] : (k == a + 1) ? ww[1] : (k == a + 2) ? ww[2] : 0.0f;
    } else {
        const float s = o - 0.5f;     // lower node nc-1 (s < 0) or nc
        const int lo = s < 0.0f ? nc - 1 : nc;
        const float f = s < 0.0f ? s + 1.0f : s;
        const int a = lo - w;
#pragma unroll
        for (int k = 0; k < 4; k++) S[k] = (k == a) ? 1.0f - f : (k == a + 1) ? f : 0.0f;
    }
}

struct SplitParams {
    float coeff;      // q dt / (2 m c) (BorisPusher.h:29)
    float cdt_h[3];   // c dt / h per dimension (position update in cells)
    float q;          // species charge
    float hdt[3];     // h / dt per dimension (Esirkepov)
    float norm;       // 1 / V
    int mask_words;
};

template <int SHAPE, int DEP>
__global__ void __launch_bounds__(PK_BLOCK) k_particles_split(GridView<float> g, EnsGeom eg, PartSplit pm, PartSplit po,
                                                              uint8_t* __restrict__ odir,
                                                              const int64_t* __restrict__ off,
                                                              const int32_t* __restrict__ cnt,
                                                              int32_t* __restrict__ nstay, int32_t* __restrict__ nmove,
                                                              int32_t* __restrict__ mvc, int32_t* __restrict__ flags,
                                                              SplitParams sp)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const long s = pk_xcd_supercell(blockIdx.x, eg.xcd_chunk);
    if (s >= eg.nsct) return;
    const int sidx[3] = {(int)(s / ((long)eg.nsc[1] * eg.nsc[2])), (int)((s / eg.nsc[2]) % eg.nsc[1]),
                         (int)(s % eg.nsc[2])};
    int its[3], dts[3], ibase[3], dbase[3], m0[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        m0[d] = sidx[d] * eg.ncps[d];
        its[d] = eg.ncps[d] + (SHAPE == PK_TSC ? 3 : 2);
        ibase[d] = m0[d] + 1;
        const bool pica_cic = SHAPE == PK_CIC && DEP == PK_DIRECT;
        dts[d] = eg.ncps[d] + (pica_cic ? 3 : 4);
        dbase[d] = pica_cic ? m0[d] + 1 : m0[d];
    }
    const int itn = its[0] * its[1] * its[2];
    const int dtn = dts[0] * dts[1] * dts[2];
    float* itile = (float*)smem;
    float* jtile = itile + 6 * itn;
    uint64_t* mask = (uint64_t*)(((uintptr_t)(jtile + 3 * dtn) + 7) & ~(uintptr_t)7);
    int* wtot = (int*)(mask + sp.mask_words);
    int* dircnt = wtot + WAVES;
    int* bad = dircnt + 27;
    long* wpre = (long*)(wtot + 32);

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    for (int t = tid; t < 6 * itn; t += PK_BLOCK) {
        const int comp = t / itn, r = t % itn;
        const int i = r / (its[1] * its[2]), j = (r / its[2]) % its[1], k = r % its[2];
        itile[t] = g.f[PK_EX + comp][((long)(ibase[0] + i) * g.s1 + (ibase[1] + j)) * g.s2 + (ibase[2] + k)];
    }
    for (int t = tid; t < 3 * dtn; t += PK_BLOCK) jtile[t] = 0.0f;
    for (int t = tid; t < 27; t += PK_BLOCK) dircnt[t] = 0;
    if (tid == 0) bad[0] = 0;
    __syncthreads();

    const long base = off[s];
    const int n = cnt[s];
    const int N1N2 = eg.ncell[1] * eg.ncell[2];
    int running = 0;
    for (int c0 = 0; c0 < n; c0 += PK_BLOCK) {
        const int i = c0 + tid;
        const bool act = i < n;
        int cell = 0, ncell_lin = 0, code = 13;
        float o[3], u[3], w = 0.0f;
        bool ok = true;
        if (act) {
            const long p = base + i;
            cell = pm.cell[p];
            o[0] = pm.ox[p]; o[1] = pm.oy[p]; o[2] = pm.oz[p];
            u[0] = pm.ux[p]; u[1] = pm.uy[p]; u[2] = pm.uz[p];
            w = pm.w[p];
            int c[3] = {cell / N1N2, (cell / eg.ncell[2]) % eg.ncell[1], cell % eg.ncell[2]};
            int l[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                l[d] = c[d] - m0[d];
                if (l[d] < 0 || l[d] >= eg.ncps[d] || !(o[d] >= 0.0f && o[d] < 1.0f)) ok = false;
            }
            ok = ok && (unsigned)cell < (unsigned)(eg.ncell[0] * N1N2);
            float e[3] = {0, 0, 0}, b[3] = {0, 0, 0};
            if (ok) {
                // ---- interpolation
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
                    const float v = gather<SHAPE>(itile + comp * itn, its, bi, wt);
                    if (comp < 3) e[comp] = v; else b[comp - 3] = v;
                }
            }
            // ---- Boris push (BorisPusher.h:32-47) in FP32
            float v[3] = {0, 0, 0};
            if (ok) {
                float eM[3], um[3], t[3], up[3], sv[3];
#pragma unroll
                for (int d = 0; d < 3; d++) eM[d] = e[d] * sp.coeff;
#pragma unroll
                for (int d = 0; d < 3; d++) um[d] = u[d] + eM[d];
                const float tk = sp.coeff * rsqrtf(1.0f + um[0] * um[0] + um[1] * um[1] + um[2] * um[2]);
#pragma unroll
                for (int d = 0; d < 3; d++) t[d] = b[d] * tk;
                up[0] = um[0] + (um[1] * t[2] - um[2] * t[1]);
                up[1] = um[1] + (um[2] * t[0] - um[0] * t[2]);
                up[2] = um[2] + (um[0] * t[1] - um[1] * t[0]);
                const float sk = 2.0f / (1.0f + t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
#pragma unroll
                for (int d = 0; d < 3; d++) sv[d] = t[d] * sk;
                u[0] = um[0] + (up[1] * sv[2] - up[2] * sv[1]) + eM[0];
                u[1] = um[1] + (up[2] * sv[0] - up[0] * sv[2]) + eM[1];
                u[2] = um[2] + (up[0] * sv[1] - up[1] * sv[0]) + eM[2];
                const float invGamma = rsqrtf(1.0f + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
#pragma unroll
                for (int d = 0; d < 3; d++) v[d] = u[d] * invGamma;    // in units of c
                ok = isfinite(u[0]) && isfinite(u[1]) && isfinite(u[2]);
            }
            // ---- move: new offset, cell shift in {-1, 0, 1}
            float on[3];
            int sh[3];
            if (ok) {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    float x = o[d] + v[d] * sp.cdt_h[d];
                    int k = x < 0.0f ? -1 : (x >= 1.0f ? 1 : 0);
                    x -= (float)k;
                    if (x >= 1.0f) { x = 0.0f; k += 1; }     // a tiny negative offset rounded up to 1
                    if (x < 0.0f) x = 0.0f;
                    on[d] = x;
                    sh[d] = k;
                    if (k < -1 || k > 1) ok = false;
                }
            }
            // ---- deposition
            if (ok) {
                const float qw = sp.q * w;
                if (DEP == PK_ESIRKEPOV) {
                    constexpr int K = SHAPE == PK_TSC ? 4 : 3;
                    float S0[3][4], DS[3][4];
                    int wb[3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const int nc0 = l[d] + 2, nc1 = l[d] + sh[d] + 2;
                        if (SHAPE == PK_TSC) {
                            // nearest rho node is the cell's own centre; window of 4 from min - 1
                            wb[d] = (nc0 < nc1 ? nc0 : nc1) - 1;
                        } else {
                            // lower CIC node: the centre below the particle; window of 3 from the min
                            const int lo0 = o[d] < 0.5f ? nc0 - 1 : nc0;
                            const int lo1 = on[d] < 0.5f ? nc1 - 1 : nc1;
                            wb[d] = lo0 < lo1 ? lo0 : lo1;
                        }
                        float S1[4];
                        rho_shape<SHAPE>(nc0, o[d], wb[d], S0[d]);
                        rho_shape<SHAPE>(nc1, on[d], wb[d], S1);
#pragma unroll
                        for (int k = 0; k < 4; k++) DS[d][k] = S1[k] - S0[d][k];
                    }
                    const float third = 1.0f / 3.0f;
#pragma unroll
                    for (int comp = 0; comp < 3; comp++) {
                        const int d1 = (comp + 1) % 3, d2 = (comp + 2) % 3;
                        const float kq = -qw * sp.hdt[comp];
                        float* t = jtile + comp * dtn;
                        int stride[3];
                        stride[0] = dts[1] * dts[2];
                        stride[1] = dts[2];
                        stride[2] = 1;
                        const int orig = wb[0] * stride[0] + wb[1] * stride[1] + wb[2] * stride[2] + stride[comp];
#pragma unroll
                        for (int p = 0; p < K; p++)
#pragma unroll
                            for (int r = 0; r < K; r++) {
                                const float coef = kq * (S0[d1][p] * S0[d2][r] + 0.5f * DS[d1][p] * S0[d2][r] +
                                                         0.5f * S0[d1][p] * DS[d2][r] + third * DS[d1][p] * DS[d2][r]);
                                float cum = 0.0f;
                                const int rowbase = orig + p * stride[d1] + r * stride[d2];
#pragma unroll
                                for (int a = 0; a < K - 1; a++) {
                                    cum += DS[comp][a] * coef;
                                    if (cum != 0.0f) atomicAdd(&t[rowbase + a * stride[comp]], cum);
                                }
                            }
                    }
                } else {
                    // direct deposition of q w v at the new position: the CIC tile is pica's
                    // (base m+1, the interpolation indices), the TSC tile has base m (indices + 1);
                    // check the stencil fits before touching the tile
                    int bc[3], bs[3];
                    float wc[3][3], ws[3][3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        interp_stencil<SHAPE>(l[d] + sh[d], on[d], bc[d], wc[d], bs[d], ws[d]);
                        if (SHAPE == PK_TSC) {
                            bc[d] += 1;
                            bs[d] += 1;
                        }
                    }
                    constexpr int KK = SHAPE == PK_TSC ? 3 : 2;
#pragma unroll
                    for (int d = 0; d < 3; d++)
                        if (bc[d] < 0 || bc[d] + KK > dts[d] || bs[d] < 0 || bs[d] + KK > dts[d]) ok = false;
                }
                if (ok && DEP == PK_DIRECT) {
                    constexpr int KK = SHAPE == PK_TSC ? 3 : 2;
                    int bc[3], bs[3];
                    float wc[3][3], ws[3][3];
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        interp_stencil<SHAPE>(l[d] + sh[d], on[d], bc[d], wc[d], bs[d], ws[d]);
                        if (SHAPE == PK_TSC) {
                            bc[d] += 1;
                            bs[d] += 1;
                        }
                    }
#pragma unroll
                    for (int comp = 0; comp < 3; comp++) {
                        const float val = v[comp] * (float)PK_C * qw;
                        float* t = jtile + comp * dtn;
                        int bi[3];
                        float wt[3][3];
#pragma unroll
                        for (int d = 0; d < 3; d++) {
                            const bool col = colloc(comp, d);
                            bi[d] = col ? bc[d] : bs[d];
#pragma unroll
                            for (int k = 0; k < 3; k++) wt[d][k] = col ? wc[d][k] : ws[d][k];
                        }
#pragma unroll
                        for (int a = 0; a < KK; a++)
#pragma unroll
                            for (int bb = 0; bb < KK; bb++)
#pragma unroll
                                for (int cc = 0; cc < KK; cc++)
                                    atomicAdd(&t[tix(dts, bi[0] + a, bi[1] + bb, bi[2] + cc)],
                                              wt[0][a] * wt[1][bb] * wt[2][cc] * val);
                    }
                }
                // ---- periodic wrap of the cell and direction class
                int cn[3];
                code = 0;
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    cn[d] = c[d] + sh[d];
                    if (cn[d] < 0) cn[d] += eg.ncell[d];
                    if (cn[d] >= eg.ncell[d]) cn[d] -= eg.ncell[d];
                    int dd = cn[d] / eg.ncps[d] - sidx[d];
                    if (dd > 1) dd -= eg.nsc[d];
                    if (dd < -1) dd += eg.nsc[d];
                    code = code * 3 + (dd + 1);
                    o[d] = on[d];
                }
                ncell_lin = (cn[0] * eg.ncell[1] + cn[1]) * eg.ncell[2] + cn[2];
            }
            if (!ok) {
                // leave the particle unchanged and report
                cell = pm.cell[p];
                ncell_lin = cell;
                o[0] = pm.ox[p]; o[1] = pm.oy[p]; o[2] = pm.oz[p];
                u[0] = pm.ux[p]; u[1] = pm.uy[p]; u[2] = pm.uz[p];
                code = 13;
                bad[0] = 1;
            }
        }
        const bool mover = act && code != 13;
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
        if (act) {
            const long p = base + i;
            if (!mover) {
                pm.cell[p] = ncell_lin;
                pm.ox[p] = o[0]; pm.oy[p] = o[1]; pm.oz[p] = o[2];
                pm.ux[p] = u[0]; pm.uy[p] = u[1]; pm.uz[p] = u[2];
            } else {
                const int rank = pre + __popcll(mb & ((1ull << lane) - 1ull));
                const long q = base + rank;
                po.cell[q] = ncell_lin;
                po.ox[q] = o[0]; po.oy[q] = o[1]; po.oz[q] = o[2];
                po.ux[q] = u[0]; po.uy[q] = u[1]; po.uz[q] = u[2];
                po.w[q] = w;
                odir[q] = (uint8_t)code;
                atomicAdd(&dircnt[code], 1);
            }
        }
        running += tot;
        __syncthreads();
    }
    const int nmv = running;
    const int nfinal = n - nmv;
    const int words = (n + 63) >> 6;
    if (nmv > 0) {
        if (tid == 0) {
            long acc = 0;
            for (int wd = 0; wd < words; wd++) {
                wpre[wd] = acc;
                acc += __popcll(mask[wd]);
            }
            wpre[words] = acc;
        }
        __syncthreads();
        auto movers_below = [&](int y) -> long {
            const int wd = y >> 6, bt = y & 63;
            return wpre[wd] + (bt ? __popcll(mask[wd] & ((1ull << bt) - 1ull)) : 0);
        };
        const long below_final = movers_below(nfinal);
        for (int t = tid; t < nmv; t += PK_BLOCK) {
            const int j = nfinal + t;
            if ((mask[j >> 6] >> (j & 63)) & 1ull) continue;
            const long r = (long)t - (movers_below(j) - below_final);
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
            const int h = (lo << 6) + (__ffsll((long long)m) - 1);
            const long src = base + j, dst = base + h;
            pm.cell[dst] = pm.cell[src];
            pm.ox[dst] = pm.ox[src]; pm.oy[dst] = pm.oy[src]; pm.oz[dst] = pm.oz[src];
            pm.ux[dst] = pm.ux[src]; pm.uy[dst] = pm.uy[src]; pm.uz[dst] = pm.uz[src];
            pm.w[dst] = pm.w[src];
        }
    }
    __syncthreads();
    for (int t = tid; t < 3 * dtn; t += PK_BLOCK) {
        const float v = jtile[t];
        if (v == 0.0f) continue;
        const int comp = t / dtn, r = t % dtn;
        const int i = r / (dts[1] * dts[2]), j = (r / dts[2]) % dts[1], k = r % dts[2];
        atomicAdd(&g.f[PK_JX + comp][((long)(dbase[0] + i) * g.s1 + (dbase[1] + j)) * g.s2 + (dbase[2] + k)],
                  v * sp.norm);
    }
    if (tid < 27) mvc[s * 27 + tid] = dircnt[tid];
    if (tid == 0) {
        nstay[s] = nfinal;
        nmove[s] = nmv;
        if (bad[0]) atomicOr(&flags[PK_FLAG_ERR], 1);
    }
}

template <int SHAPE, int DEP>
int launch(pk_ens* e, pk_grid* g, double dt)
{
    EnsGeom eg = ens_geom(e);
    SplitParams sp;
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
    int its[3], dts[3];
    for (int d = 0; d < 3; d++) {
        its[d] = e->ncps[d] + (SHAPE == PK_TSC ? 3 : 2);
        dts[d] = e->ncps[d] + ((SHAPE == PK_CIC && DEP == PK_DIRECT) ? 3 : 4);
    }
    const size_t lds = sizeof(float) * (6 * its[0] * its[1] * its[2] + 3 * dts[0] * dts[1] * dts[2]) + 8 +
                       sizeof(uint64_t) * sp.mask_words + sizeof(int) * 32 + sizeof(long) * (sp.mask_words + 1) + 16;
    if (lds > 160 * 1024) return pk_fail(e->ctx, PK_E_CAPACITY, "particle step: supercell too large for LDS");
    const unsigned nblocks = (unsigned)(PK_XCD * (long)eg.xcd_chunk);
    pk_ktimer tm(e->ctx, "particles");
    hipLaunchKernelGGL((k_particles_split<SHAPE, DEP>), dim3(nblocks), dim3(PK_BLOCK), lds, e->ctx->stream,
                       grid_view<float>(g), eg, part_split(e, e->main), part_split(e, e->outbox), e->odir, e->off,
                       e->cnt, e->nstay, e->nmove, e->mvc, e->flags, sp);
    PK_HIP(e->ctx, hipGetLastError());
    return PK_OK;
}

}  // namespace

int pk_particles_split(pk_ens* e, pk_grid* g, double dt, int shape, int dep)
{
    if (shape == PK_CIC && dep == PK_DIRECT) return launch<PK_CIC, PK_DIRECT>(e, g, dt);
    if (shape == PK_CIC && dep == PK_ESIRKEPOV) return launch<PK_CIC, PK_ESIRKEPOV>(e, g, dt);
    if (shape == PK_TSC && dep == PK_DIRECT) return launch<PK_TSC, PK_DIRECT>(e, g, dt);
    if (shape == PK_TSC && dep == PK_ESIRKEPOV) return launch<PK_TSC, PK_ESIRKEPOV>(e, g, dt);
    return pk_fail(e->ctx, PK_E_ARG, "particle step: unknown shape/depositor");
}
