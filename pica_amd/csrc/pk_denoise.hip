// pk_denoise.hip -- learned current denoiser: 3 -> 16 -> 16 -> 3 periodic conv3d (3x3x3 taps,
// ReLU) on BF16 MFMA with FP32 accumulation, its backward pass, Adam, and the C ABI (pk_dn_*).
// No counterpart in pica.
//
// Each convolution is an implicit GEMM: M = voxels, N = output channels, K = taps x input
// channels.  A wavefront owns 16 consecutive voxels and chains gfx950's v_mfma_f32_16x16x32_bf16
// over K:
//   A (16 voxels x 32 k): lane l holds A[l % 16][8 (l / 16) .. +7]  -- 8 BF16 of neighbours of
//                         one voxel (one 16-byte LDS read, or two 8-byte ones)
//   B (32 k x 16 cout):   lane l holds B[8 (l / 16) .. +7][l % 16]  -- pre-packed weights,
//                         held in registers for the whole kernel
//   D (16 x 16, FP32):    lane l holds D[4 (l / 16) + j][l % 16], j = 0..3
// With 16 input channels a K-chunk is 2 taps of one x-offset dx x 16 channels: 5 chunks per dx
// (the 9 (dy, dz) offsets in pairs, the tenth zero), 15 in all.  A wavefront then computes the
// four x-rows of its brick together: one LDS read of a halo x-plane feeds every row that plane is
// a tap plane of (dx = plane - row), 30 reads for 60 MFMAs instead of 56 for 56.  With 3 (padded
// to 4) input channels a chunk is 8 taps x 4 channels (4 chunks).  The input gradient is the same forward
// kernel with the weights flipped (tap -> 26 - tap) and transposed; the weight gradient is a GEMM
// with K = voxels over LDS-staged bricks on the same 16x16x32 MFMA, with the bias gradient as one
// more tile.  Training stacks a mini-batch's samples so that each layer is one launch.
//
// Tap t = (dx+1)*9 + (dy+1)*3 + (dz+1) reads the input at (x+dx, y+dy, z+dz), periodic.
// Activations are channel-last BF16 ([voxel][16] or [voxel][4]); the network output and the
// training pairs are channel-first FP32 ([3][voxel]).
#include <cmath>
#include <type_traits>

#include "pk_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));   // operand type of the 16x16x32 MFMA
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int HID = 16;
constexpr int PK_DN_RADIUS = 3;   // receptive radius of the three 3x3x3 layers
constexpr int NPARAM = 27 * 3 * HID + HID + 27 * HID * HID + HID + 27 * HID * 3 + 3;   // 9539
constexpr int KC16 = 15, KC4 = 4;      // K-chunks of a 16- and a 4-channel layer
constexpr int WF16 = KC16 * 64 * 8, WF4 = KC4 * 64 * 8;   // packed B fragments (BF16 values) per layer
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + 27 * 3 * HID, OFF_W2 = OFF_B1 + HID, OFF_B2 = OFF_W2 + 27 * HID * HID,
              OFF_W3 = OFF_B2 + HID, OFF_B3 = OFF_W3 + 27 * HID * 3;

struct pk_dn_pair {
    float* low = nullptr;    // [3][nv] normalised input current
    float* high = nullptr;   // [3][nv] normalised target current
    int n[3] = {0, 0, 0};
    int margin = 0;          // voxels this close to a face are left out of the loss (crops: the
                             // network's receptive radius, so no output sees the wrapped edge)
};

struct pk_dn {
    pk_ctx* ctx = nullptr;
    double scale = 1.0;      // J (CGS) -> network units
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // parameters, gradient, Adam moments
    long adam_t = 0;
    uint16_t* wf[3] = {nullptr, nullptr, nullptr};   // packed forward weights of layers 1..3
    uint16_t* wt[2] = {nullptr, nullptr};            // packed flipped-transposed W2, W3 (input gradients)
    // work buffers sized for cap voxels
    long cap = 0;
    uint16_t *xb = nullptr, *a1 = nullptr, *a2 = nullptr, *g3 = nullptr, *dz2 = nullptr, *dz1 = nullptr;
    float *out = nullptr, *tgt = nullptr;
    float* part = nullptr;    // per-workgroup weight-gradient tiles (k_wgrad)
    double* loss = nullptr;   // device accumulator
    std::vector<pk_dn_pair> pairs;
    // charge filter of pk_step_denoised: radial gain g(|k|) on nbins points over [0, kmax]
    // (linear interpolation, device copy in cgain); cgain_n = 0: none
    int cgain_n = 0;
    double cgain_kmax = 0.0;
    double* cgain = nullptr;
};

namespace {

// ns volumes of n0 x n1 x n2 voxels each, stacked in memory (sample s holds voxels
// [s nv, (s + 1) nv)); each is periodic on its own.  Training stacks its batch this way, so one
// launch per layer covers every sample.
struct Vol {
    int n0, n1, n2;
    long nv;         // voxels per sample
    int ns = 1;      // samples
};

inline unsigned nblk(long n, int b) { return (unsigned)((n + b - 1) / b); }

__host__ __device__ __forceinline__ uint16_t to_bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// volumes hold fewer than 2^31 voxels (check_dims), so 32-bit unsigned division suffices
__device__ __forceinline__ void voxel_xyz(const Vol& V, long v, int& x, int& y, int& z)
{
    const unsigned n12 = (unsigned)V.n1 * (unsigned)V.n2, u = (unsigned)v;
    x = (int)(u / n12);
    const unsigned r = u - (unsigned)x * n12;
    y = (int)(r / (unsigned)V.n2);
    z = (int)(r - (unsigned)y * (unsigned)V.n2);
}

__device__ __forceinline__ long neighbour(const Vol& V, int x, int y, int z, int tap)
{
    int a = x + tap / 9 - 1, b = y + (tap / 3) % 3 - 1, c = z + tap % 3 - 1;
    a += a < 0 ? V.n0 : 0;
    a -= a >= V.n0 ? V.n0 : 0;
    b += b < 0 ? V.n1 : 0;
    b -= b >= V.n1 ? V.n1 : 0;
    c += c < 0 ? V.n2 : 0;
    c -= c >= V.n2 ? V.n2 : 0;
    return ((long)a * V.n1 + b) * V.n2 + c;
}

enum { OUT_RELU_BF16 = 0, OUT_F32 = 1, OUT_MASKED_BF16 = 2 };

// One convolution layer over a brick of BX x BY x BZ voxels per workgroup: the brick and its
// one-voxel periodic halo are staged in LDS (each input voxel leaves L2 about 2.5 times instead
// of 27), then every wavefront runs 16-voxel z-rows of the brick through the MFMA chain.  A 4 x 4
// x 16 brick stages 648 voxels for 256 (2 x 8 x 16: 720, and a 10% slower denoiser at 128^3).
// CIN = 16 (2 taps x 16 channels per K-chunk) or 4 (8 taps x 4 channels per chunk).  Five
// workgroups per CU (92 VGPRs for the 16-channel form, 104 KB of LDS): at four, the 16-channel
// layers took 63 / 58 us at 128^3 against 56 / 49 us.
constexpr int BX = 4, BY = 4, BZ = 16, HX = BX + 2, HY = BY + 2, HZ = BZ + 2;
constexpr int BRICK_TILES = BX * BY * (BZ / 16);
static_assert(BRICK_TILES % 8 == 0, "4 wavefronts x 2 rows per pass");

// periodic image of a >= -1 (a brick's halo can reach several periods past an edge shorter
// than the brick, e.g. the one-cell edge of a 2D grid)
__device__ __forceinline__ int wrap(int a, int n)
{
    a += a < 0 ? n : 0;
    return a < n ? a : a % n;
}

template <int CIN, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_conv(Vol V, const uint16_t* __restrict__ in, const uint16_t* __restrict__ wf,
                                              const float* __restrict__ bias, int cout, uint16_t* __restrict__ out_b,
                                              float* __restrict__ out_f, const uint16_t* __restrict__ mask_act)
{
    constexpr int KC = CIN == 16 ? KC16 : KC4;
    constexpr int UNITS = CIN == 16 ? 2 : 1;              // 16- or 8-byte units per voxel
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t* halo = (uint16_t*)smem;                     // [HX][HY][HZ][CIN]
    const int nbz = (V.n2 + BZ - 1) / BZ, nby = (V.n1 + BY - 1) / BY;
    // XCD-aware: the workgroups of one XCD (blockIdx % 8) take one contiguous run of bricks, so
    // the halos that neighbouring bricks share stay in that XCD's L2
    const long nbrick = (long)((V.n0 + BX - 1) / BX) * nby * nbz, chunk = gridDim.x / PK_XCD;
    long b = (blockIdx.x % PK_XCD) * chunk + blockIdx.x / PK_XCD;
    if (b >= nbrick * V.ns) return;
    const long sample = b / nbrick;
    b -= sample * nbrick;
    in += sample * V.nv * CIN;
    if (out_b) out_b += sample * V.nv * 16;
    if (out_f) out_f += sample * 3 * V.nv;
    if (mask_act) mask_act += sample * V.nv * 16;
    const int x0 = (int)(b / ((long)nby * nbz)) * BX, y0 = (int)((b / nbz) % nby) * BY, z0 = (int)(b % nbz) * BZ;
    // all of a thread's halo loads are issued before the first LDS write (one wait, not one per load)
    typedef typename std::conditional<CIN == 16, bf16x8, bf16x4>::type unit_t;
    constexpr int NUNIT = HX * HY * HZ * UNITS, PER = (NUNIT + 255) / 256;
    unit_t buf[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int u = threadIdx.x + 256 * k;
        if (u < NUNIT) {
            const int h = u / UNITS, part = u - h * UNITS;
            const int hz = h % HZ, hy = (h / HZ) % HY, hx = h / (HZ * HY);
            const long gv =
                ((long)wrap(x0 - 1 + hx, V.n0) * V.n1 + wrap(y0 - 1 + hy, V.n1)) * V.n2 + wrap(z0 - 1 + hz, V.n2);
            buf[k] = *(const unit_t*)(in + gv * CIN + (CIN / UNITS) * part);
        }
    }
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int u = threadIdx.x + 256 * k;
        if (u < NUNIT) *(unit_t*)(halo + u * (CIN / UNITS)) = buf[k];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, q = lane >> 4;
    bf16x8 B[KC];
#pragma unroll
    for (int kc = 0; kc < KC; kc++) B[kc] = *(const bf16x8*)(wf + ((size_t)kc * 64 + lane) * 8);
    const float bco = (bias && m < cout) ? bias[m] : 0.0f;
    __syncthreads();
    // the bias, activation and store of one 16-voxel z-row (brick row tx, ty, tz)
    const auto store = [&](const f32x4& acc, int tx, int ty, int tz) {
        const int co = m;
        const int gx = x0 + tx, gy = y0 + ty;
        if (co >= cout || gx >= V.n0 || gy >= V.n1) return;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int gz = z0 + tz + 4 * q + j;
            if (gz >= V.n2) continue;
            const long vo = ((long)gx * V.n1 + gy) * V.n2 + gz;
            const float zv = acc[j] + bco;
            if (MODE == OUT_RELU_BF16) out_b[vo * 16 + co] = to_bf16(zv > 0.0f ? zv : 0.0f);
            else if (MODE == OUT_F32) out_f[(long)co * V.nv + vo] = zv;
            else out_b[vo * 16 + co] = to_bf16(from_bf16(mask_act[vo * 16 + co]) > 0.0f ? zv : 0.0f);
        }
    };
    if constexpr (CIN == 16) {
        // wavefront w: the z-row at brick y = w for all four x.  Halo x-plane hx is tap plane
        // dx = hx - r - 1 of row r; chunk 5 (dx + 1) + c holds (dy, dz) offsets 2c, 2c + 1 of it
        // (offset o at halo (ty + o / 3, m + o % 3)), lanes q / 2 taking one offset each.
        static_assert(BY * (BZ / 16) == PK_BLOCK / 64 && BZ == 16, "one z-row of each brick y per wavefront");
        f32x4 acc[BX];
#pragma unroll
        for (int r = 0; r < BX; r++) acc[r] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int hx = 0; hx < HX; hx++) {
#pragma unroll
            for (int c = 0; c < 5; c++) {
                const int o = 2 * c + q / 2;
                bf16x8 a = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
                if (o < 9) {
                    const int h = (hx * HY + wave + o / 3) * HZ + m + o % 3;
                    a = *(const bf16x8*)(halo + h * 16 + 8 * (q % 2));
                }
#pragma unroll
                for (int r = 0; r < BX; r++) {
                    const int dxi = hx - r;
                    if (dxi >= 0 && dxi <= 2)
                        acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, B[5 * dxi + c]), acc[r], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < BX; r++) store(acc[r], r, wave, 0);
    } else {
        // two z-rows per pass: independent accumulators hide the MFMA dependency latency
        for (int t0 = wave; t0 < BRICK_TILES; t0 += 8) {
            int tx[2], ty[2], tz[2];
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int t = t0 + 4 * r;
                tx[r] = t / (BY * (BZ / 16));
                ty[r] = (t / (BZ / 16)) % BY;
                tz[r] = (t % (BZ / 16)) * 16;
            }
            f32x4 acc[2] = {{0.0f, 0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 0.0f}};
#pragma unroll
            for (int kc = 0; kc < KC; kc++) {
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    bf16x4 lo = bf16x4{0, 0, 0, 0}, hi = bf16x4{0, 0, 0, 0};   // 4 channels of two taps
                    const int t0 = 8 * kc + 2 * q, t1 = t0 + 1;
                    if (t0 < 27) {
                        const int h = ((tx[r] + t0 / 9) * HY + (ty[r] + (t0 / 3) % 3)) * HZ + (tz[r] + m + t0 % 3);
                        lo = *(const bf16x4*)(halo + h * 4);
                    }
                    if (t1 < 27) {
                        const int h = ((tx[r] + t1 / 9) * HY + (ty[r] + (t1 / 3) % 3)) * HZ + (tz[r] + m + t1 % 3);
                        hi = *(const bf16x4*)(halo + h * 4);
                    }
                    const bf16x8 a = bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                    acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a),
                                                                     __builtin_bit_cast(bf16v8, B[kc]), acc[r], 0, 0, 0);
                }
            }
#pragma unroll
            for (int r = 0; r < 2; r++) store(acc[r], tx[r], ty[r], tz[r]);
        }
    }
}

// dW[t][ci][co] = sum_v in[v + off(t)][ci] g[v][co], as a GEMM per 16 x 16 tile (rows: one tap x
// 16 channels with CIN = 16, four taps x 4 channels with CIN = 4; columns: co) with K = voxels.  A
// workgroup takes bricks of BX x BY x BZ voxels with a grid stride; each brick's input with its
// periodic one-voxel halo and its output gradient are staged in LDS as in k_conv, and each
// wavefront accumulates its tiles over the brick's 256 voxels on v_mfma_f32_16x16x32_bf16
// (K-block = 32 voxels: two 16-voxel z-rows; lane (m, q) holds voxels 8q .. 8q + 7 of the block for
// row / column m).  Tile NT is the bias gradient db[co] = sum_v g[v][co]: A is all ones, so every
// row holds it; it fills the slot that wavefront 3 has left over (27 = 3 + 4 x 6, 7 = 3 + 4 x 1).
// The workgroup's NT + 1 tiles go to part[blockIdx.x]; k_wgrad_sum adds them into dW and db.
constexpr int WG_TILES16 = 27, WG_TILES4 = 7, WGRAD_WG_MAX = 512;
static_assert(WG_TILES16 % 4 == 3 && WG_TILES4 % 4 == 3, "the bias tile takes wavefront 3's last slot");

template <int CIN>
__global__ void __launch_bounds__(256) k_wgrad(Vol V, const uint16_t* __restrict__ in, const uint16_t* __restrict__ g,
                                               int gpad, int cout, float* __restrict__ part)
{
    constexpr int NT = CIN == 16 ? WG_TILES16 : WG_TILES4, TPW = (NT + 1 + 3) / 4;
    constexpr int UNITS = CIN == 16 ? 2 : 1, NVOX = BX * BY * BZ;
    static_assert(NVOX == 256 && BZ == 16, "one brick voxel per thread, K-blocks of two z-rows");
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t* halo = (uint16_t*)smem;                     // [HX][HY][HZ][CIN]
    uint16_t* gt = halo + HX * HY * HZ * CIN;             // [NVOX][16], zero outside the volume and for co >= cout
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, q = lane >> 4;
    const int nbz = (V.n2 + BZ - 1) / BZ, nby = (V.n1 + BY - 1) / BY;
    const long nbrick = (long)((V.n0 + BX - 1) / BX) * nby * nbz, total = nbrick * V.ns;
    typedef typename std::conditional<CIN == 16, bf16x8, bf16x4>::type unit_t;
    constexpr int NUNIT = HX * HY * HZ * UNITS, PER = (NUNIT + 255) / 256;
    f32x4 acc[TPW];
#pragma unroll
    for (int i = 0; i < TPW; i++) acc[i] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (long bb = blockIdx.x; bb < total; bb += gridDim.x) {
        const long sample = bb / nbrick, b = bb - sample * nbrick;
        const int x0 = (int)(b / ((long)nby * nbz)) * BX, y0 = (int)((b / nbz) % nby) * BY, z0 = (int)(b % nbz) * BZ;
        const uint16_t* ins = in + sample * V.nv * CIN;
        unit_t buf[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int u = threadIdx.x + 256 * k;
            if (u < NUNIT) {
                const int h = u / UNITS, unit = u - h * UNITS;
                const int hz = h % HZ, hy = (h / HZ) % HY, hx = h / (HZ * HY);
                const long gv = ((long)wrap(x0 - 1 + hx, V.n0) * V.n1 + wrap(y0 - 1 + hy, V.n1)) * V.n2 +
                                wrap(z0 - 1 + hz, V.n2);
                buf[k] = *(const unit_t*)(ins + gv * CIN + (CIN / UNITS) * unit);
            }
        }
        bf16x8 glo = bf16x8{0, 0, 0, 0, 0, 0, 0, 0}, ghi = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
        {
            const int t = threadIdx.x;
            const int gx = x0 + t / (BY * BZ), gy = y0 + (t / BZ) % BY, gz = z0 + t % BZ;
            if (gx < V.n0 && gy < V.n1 && gz < V.n2) {
                const uint16_t* gp = g + (sample * V.nv + ((long)gx * V.n1 + gy) * V.n2 + gz) * gpad;
                if (gpad == 16) {
                    glo = *(const bf16x8*)gp;
                    ghi = *(const bf16x8*)(gp + 8);
                } else {
                    const bf16x4 v4 = *(const bf16x4*)gp;
                    glo = bf16x8{v4[0], v4[1], v4[2], v4[3], 0, 0, 0, 0};
                }
#pragma unroll
                for (int c = 0; c < 8; c++) {
                    if (c >= cout) glo[c] = 0;
                    if (c + 8 >= cout) ghi[c] = 0;
                }
            }
        }
        __syncthreads();   // the previous brick's LDS reads are done
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int u = threadIdx.x + 256 * k;
            if (u < NUNIT) *(unit_t*)(halo + u * (CIN / UNITS)) = buf[k];
        }
        *(bf16x8*)(gt + threadIdx.x * 16) = glo;
        *(bf16x8*)(gt + threadIdx.x * 16 + 8) = ghi;
        __syncthreads();
#pragma unroll 2
        for (int kb = 0; kb < NVOX / 32; kb++) {
            const int row = 2 * kb + (q >> 1), tx = row / BY, ty = row % BY, tz = 8 * (q & 1);
            bf16x8 B;
#pragma unroll
            for (int j = 0; j < 8; j++) B[j] = (short)gt[(row * BZ + tz + j) * 16 + m];
#pragma unroll
            for (int i = 0; i < TPW; i++) {
                const int tile = wave + 4 * i;
                if (tile > NT) break;
                const int tap = CIN == 16 ? tile : 4 * tile + m / 4, ci = CIN == 16 ? m : m % 4;
                constexpr short one = 0x3f80;   // BF16 1.0
                bf16x8 A = tile == NT ? bf16x8{one, one, one, one, one, one, one, one} : bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
                if (tile < NT && tap < 27) {
                    const int h0 = ((tx + tap / 9) * HY + (ty + (tap / 3) % 3)) * HZ + tz + tap % 3;
#pragma unroll
                    for (int j = 0; j < 8; j++) A[j] = (short)halo[(h0 + j) * CIN + ci];
                }
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, A),
                                                                 __builtin_bit_cast(bf16v8, B), acc[i], 0, 0, 0);
            }
        }
    }
    // D[4q + j][m]: row 4q + j of the tile, column co = m
    float* pp = part + (long)blockIdx.x * (NT + 1) * 256;
#pragma unroll
    for (int i = 0; i < TPW; i++) {
        const int tile = wave + 4 * i;
        if (tile <= NT)
#pragma unroll
            for (int j = 0; j < 4; j++) pp[tile * 256 + (4 * q + j) * 16 + m] = acc[i][j];
    }
}

// dW[t][ci][co] and db[co] += sum over the nwg workgroups' tiles; blockIdx.y takes every
// gridDim.y-th group of 4 workgroups, so a weight receives gridDim.y atomic adds
template <int CIN>
__global__ void __launch_bounds__(256) k_wgrad_sum(int nwg, const float* __restrict__ part, int cin, int cout,
                                                   float* __restrict__ dW, float* __restrict__ db)
{
    constexpr int NT = CIN == 16 ? WG_TILES16 : WG_TILES4, NE = (NT + 1) * 256;
    const int lane = threadIdx.x & 63, p = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    float s = 0.0f;
    if (i < NE)
        for (int w = 4 * blockIdx.y + p; w < nwg; w += 4 * gridDim.y) s += part[(long)w * NE + i];
    __shared__ float red[4][64];
    red[p][lane] = s;
    __syncthreads();
    if (p != 0 || i >= NE) return;
    s = red[0][lane] + red[1][lane] + red[2][lane] + red[3][lane];
    const int tile = i / 256, row = (i / 16) % 16, co = i % 16;
    if (s == 0.0f || co >= cout) return;
    if (tile == NT) {
        if (row == 0) atomicAdd(&db[co], s);
        return;
    }
    const int tap = CIN == 16 ? tile : 4 * tile + row / 4, ci = CIN == 16 ? row : row % 4;
    if (tap < 27 && ci < cin) atomicAdd(&dW[((long)tap * cin + ci) * cout + co], s);
}

// weights -> MFMA B fragments of v_mfma_f32_16x16x32_bf16, wf[kc][lane][8].  Effective layer:
// cin_e x cout_e with W_e(t, a, b) = W(t, a, b), or for the input gradient (flip) W(26 - t, b, a);
// cin_pad 16: chunk kc = 5 (dx + 1) + c holds the taps 9 (dx + 1) + 2c, + 2c + 1 (the tenth
// offset of an x-plane is zero) x 16 channels; cin_pad 4: taps 8kc .. 8kc + 7 x 4 channels.
__device__ void pack_one(const float* __restrict__ W, int cin, int cout, int flip, int cin_pad, uint16_t* __restrict__ wf)
{
    const int kcs = cin_pad == 16 ? KC16 : KC4;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kcs * 64 * 8) return;
    const int j = i & 7, lane = (i >> 3) & 63, kc = i >> 9;
    const int n = lane & 15, k = 8 * (lane >> 4) + j;
    int tap, a;
    if (cin_pad == 16) {
        const int o = 2 * (kc % 5) + k / 16;
        tap = o < 9 ? 9 * (kc / 5) + o : 27;
        a = k % 16;
    } else {
        tap = 8 * kc + k / 4;
        a = k % 4;
    }
    const int cin_e = flip ? cout : cin, cout_e = flip ? cin : cout;
    float w = 0.0f;
    if (tap < 27 && a < cin_e && n < cout_e)
        w = flip ? W[((long)(26 - tap) * cin + n) * cout + a] : W[((long)tap * cin + a) * cout + n];
    wf[i] = to_bf16(w);
}

// the five packed layouts (blockIdx.y): forward W1, W2, W3; flipped-transposed W2, W3
struct PackDst {
    uint16_t *wf0, *wf1, *wf2, *wt0, *wt1;
};

__global__ void k_pack(const float* __restrict__ p, PackDst o)
{
    switch (blockIdx.y) {
    case 0: pack_one(p + OFF_W1, 3, HID, 0, 4, o.wf0); break;
    case 1: pack_one(p + OFF_W2, HID, HID, 0, 16, o.wf1); break;
    case 2: pack_one(p + OFF_W3, HID, 3, 0, 16, o.wf2); break;
    case 3: pack_one(p + OFF_W2, HID, HID, 1, 16, o.wt0); break;
    default: pack_one(p + OFF_W3, HID, 3, 1, 4, o.wt1); break;
    }
}

// a training mini-batch: sample k is pair sel[k]; its input goes to xb (BF16, channel-last) and
// its target to tgt ([k][3][nv] FP32)
__global__ void k_gather(long nv, const float* const* __restrict__ lows, const float* const* __restrict__ highs,
                         const int* __restrict__ sel, uint16_t* __restrict__ xb, float* __restrict__ tgt)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
    const int k = blockIdx.y, pr = sel[k];
    const float *lo = lows[pr], *hi = highs[pr];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        xb[(k * nv + v) * 4 + c] = to_bf16(lo[(long)c * nv + v]);
        tgt[(3l * k + c) * nv + v] = hi[(long)c * nv + v];
    }
    xb[(k * nv + v) * 4 + 3] = 0;
}

// network input: x_b[v][0..3] = bf16(scale * src[c][v]) (channel 3 zero); src channel-first FP32
__global__ void k_input_f32(long nv, const float* __restrict__ src, double scale, uint16_t* __restrict__ xb)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nv) return;
#pragma unroll
    for (int c = 0; c < 3; c++) xb[v * 4 + c] = to_bf16((float)(scale * src[(long)c * nv + v]));
    xb[v * 4 + 3] = 0;
}

// physical J of a grid -> channel-first FP32 scaled (the network's units)
template <class T>
__global__ void k_grid_to_f32(GridView<T> g, Vol V, double scale, float* __restrict__ dst)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V.nv) return;
    int x, y, z;
    voxel_xyz(V, v, x, y, z);
    const long gi = ((long)(x + g.G) * g.s1 + (y + g.G)) * g.s2 + (z + g.G);
#pragma unroll
    for (int c = 0; c < 3; c++) dst[(long)c * V.nv + v] = (float)(scale * (double)g.f[PK_JX + c][gi]);
}

// an e0 x e1 x e2 crop of a grid's physical J at offset o (periodic wrap) -> channel-first FP32
// scaled; the same rounding as k_grid_to_f32
template <class T>
__global__ void k_grid_crop_to_f32(GridView<T> g, int n0, int n1, int n2, Vol V, int o0, int o1, int o2, double scale,
                                   float* __restrict__ dst)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V.nv) return;
    int x, y, z;
    voxel_xyz(V, v, x, y, z);
    x += o0; y += o1; z += o2;
    x -= x >= n0 ? n0 : 0; y -= y >= n1 ? n1 : 0; z -= z >= n2 ? n2 : 0;
    const long gi = ((long)(x + g.G) * g.s1 + (y + g.G)) * g.s2 + (z + g.G);
#pragma unroll
    for (int c = 0; c < 3; c++) dst[(long)c * V.nv + v] = (float)(scale * (double)g.f[PK_JX + c][gi]);
}

// physical J of a grid -> the network input xb[v][0..3] (BF16, scaled; the same roundings as
// k_grid_to_f32 then k_input_f32)
// div0 != null: also div0[v] = D J (pk_div_j) of the current as it is read
template <class T>
__global__ void k_grid_to_input(GridView<T> g, Vol V, double scale, uint16_t* __restrict__ xb,
                                double* __restrict__ div0)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V.nv) return;
    int x, y, z;
    voxel_xyz(V, v, x, y, z);
    const long gi = ((long)(x + g.G) * g.s1 + (y + g.G)) * g.s2 + (z + g.G);
#pragma unroll
    for (int c = 0; c < 3; c++) xb[v * 4 + c] = to_bf16((float)(scale * (double)g.f[PK_JX + c][gi]));
    xb[v * 4 + 3] = 0;
    if (div0) {
        const int n[3] = {V.n0, V.n1, V.n2};
        div0[v] = pk_div_j(g, n, x, y, z);
    }
}

template <class T>
__global__ void k_f32_to_grid(GridView<T> g, Vol V, double inv_scale, const float* __restrict__ src)
{
    const long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V.nv) return;
    int x, y, z;
    voxel_xyz(V, v, x, y, z);
    const long gi = ((long)(x + g.G) * g.s1 + (y + g.G)) * g.s2 + (z + g.G);
#pragma unroll
    for (int c = 0; c < 3; c++) g.f[PK_JX + c][gi] = (T)(inv_scale * (double)src[(long)c * V.nv + v]);
}

// MSE loss and its gradient: g3[v][c] = bf16(2 (out - y) / count), loss += sum (out - y)^2 / count,
// over the voxels at least `margin` from every face of their sample (the others get a zero gradient)
__global__ void __launch_bounds__(256) k_loss(Vol V, int margin, const float* __restrict__ out,
                                              const float* __restrict__ y, double inv_count, uint16_t* __restrict__ g3,
                                              double* __restrict__ loss)
{
    double s = 0.0;
    const long nv = V.nv;
    // v: stacked voxel; out, y are [ns][3][nv]
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < nv * V.ns; v += (long)gridDim.x * blockDim.x) {
        const long sample = v / nv, base = 2 * sample * nv + v;   // (3 sample + c) nv + (v - sample nv)
        bool in = true;
        if (margin > 0) {
            int x, yy, z;
            voxel_xyz(V, v - sample * nv, x, yy, z);
            in = x >= margin && x < V.n0 - margin && yy >= margin && yy < V.n1 - margin && z >= margin &&
                 z < V.n2 - margin;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float d = in ? out[base + (long)c * nv] - y[base + (long)c * nv] : 0.0f;
            g3[v * 4 + c] = to_bf16((float)(2.0 * inv_count) * d);
            s += (double)d * d;
        }
        g3[v * 4 + 3] = 0;
    }
    // one atomic per workgroup: every wavefront's sum to the same address would serialise
    __shared__ double part[4];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        s = part[0] + part[1] + part[2] + part[3];
        if (s != 0.0) atomicAdd(loss, s * inv_count);
    }
}

// one Adam step; the gradient is cleared for the next mini-batch
__global__ void k_adam(int n, float* __restrict__ p, float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    g[i] = 0.0f;
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_zero_f32(long n, float* p)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_zero_f64(double* p) { p[0] = 0.0; }

// ---------------------------------------------------------------- host side
int dalloc(pk_ctx* c, void** p, size_t bytes)
{
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    if (hipMalloc(p, bytes ? bytes : 16) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return pk_fail(c, PK_E_ALLOC, "denoiser: hipMalloc failed");
    }
    return PK_OK;
}

int ensure_cap(pk_dn* d, long nv)
{
    if (nv <= d->cap) return PK_OK;
    pk_ctx* c = d->ctx;
    PK_HIP(c, hipStreamSynchronize(c->stream));
    // no capacity while the buffers are replaced: a failed allocation leaves cap 0, so a later
    // call grows again instead of launching on a freed or null buffer
    d->cap = 0;
    int r;
    if ((r = dalloc(c, (void**)&d->xb, sizeof(uint16_t) * 4 * nv))) return r;
    if ((r = dalloc(c, (void**)&d->a1, sizeof(uint16_t) * 16 * nv))) return r;
    if ((r = dalloc(c, (void**)&d->a2, sizeof(uint16_t) * 16 * nv))) return r;
    if ((r = dalloc(c, (void**)&d->g3, sizeof(uint16_t) * 4 * nv))) return r;
    if ((r = dalloc(c, (void**)&d->dz2, sizeof(uint16_t) * 16 * nv))) return r;
    if ((r = dalloc(c, (void**)&d->dz1, sizeof(uint16_t) * 16 * nv))) return r;
    if ((r = dalloc(c, (void**)&d->out, sizeof(float) * 3 * nv))) return r;
    if ((r = dalloc(c, (void**)&d->tgt, sizeof(float) * 3 * nv))) return r;
    d->cap = nv;
    return PK_OK;
}

// pack all weight layouts from the FP32 parameters
int pack_weights(pk_dn* d)
{
    const PackDst o{d->wf[0], d->wf[1], d->wf[2], d->wt[0], d->wt[1]};
    hipLaunchKernelGGL(k_pack, dim3(nblk(WF16, 256), 5), dim3(256), 0, d->ctx->stream, d->p, o);
    PK_HIP(d->ctx, hipGetLastError());
    return PK_OK;
}

// launch size: the brick count rounded up to a multiple of the XCD count (k_conv maps blocks to bricks)
unsigned conv_blocks(const Vol& V)
{
    const long nbrick = (long)((V.n0 + BX - 1) / BX) * ((V.n1 + BY - 1) / BY) * ((V.n2 + BZ - 1) / BZ) * V.ns;
    return (unsigned)(PK_XCD * ((nbrick + PK_XCD - 1) / PK_XCD));
}

constexpr size_t conv_lds(int cin) { return sizeof(uint16_t) * HX * HY * HZ * cin; }

// layers 1..3 on d->xb; the output lands in d->out ([3][nv] FP32)
// the network on d->xb, output to d->out (channel-first FP32)
int forward(pk_dn* d, const Vol& V)
{
    hipStream_t s = d->ctx->stream;
    const unsigned nb = conv_blocks(V);
    hipLaunchKernelGGL((k_conv<4, OUT_RELU_BF16>), dim3(nb), dim3(256), conv_lds(4), s, V, d->xb, d->wf[0], d->p + OFF_B1, HID,
                       d->a1, (float*)nullptr, (const uint16_t*)nullptr);
    hipLaunchKernelGGL((k_conv<16, OUT_RELU_BF16>), dim3(nb), dim3(256), conv_lds(16), s, V, d->a1, d->wf[1], d->p + OFF_B2, HID,
                       d->a2, (float*)nullptr, (const uint16_t*)nullptr);
    hipLaunchKernelGGL((k_conv<16, OUT_F32>), dim3(nb), dim3(256), conv_lds(16), s, V, d->a2, d->wf[2], d->p + OFF_B3,
                       3, (uint16_t*)nullptr, d->out, (const uint16_t*)nullptr);
    PK_HIP(d->ctx, hipGetLastError());
    return PK_OK;
}

constexpr size_t wgrad_lds(int cin) { return sizeof(uint16_t) * (HX * HY * HZ * cin + BX * BY * BZ * 16); }

// weight and bias gradients of one layer: dW += sum_v in[v + off(t)] g[v], db += sum_v g[v]
int wgrad(pk_dn* d, const Vol& V, const uint16_t* in, int cin_pad, int cin, const uint16_t* g, int gpad, int cout,
          float* dW, float* db)
{
    hipStream_t s = d->ctx->stream;
    const long total = (long)((V.n0 + BX - 1) / BX) * ((V.n1 + BY - 1) / BY) * ((V.n2 + BZ - 1) / BZ) * V.ns;
    const int nwg = (int)(total < WGRAD_WG_MAX ? total : WGRAD_WG_MAX);
    if (cin_pad == 16) {
        hipLaunchKernelGGL(k_wgrad<16>, dim3(nwg), dim3(256), wgrad_lds(16), s, V, in, g, gpad, cout, d->part);
        hipLaunchKernelGGL(k_wgrad_sum<16>, dim3((WG_TILES16 + 1) * 256 / 64, 8), dim3(256), 0, s, nwg, d->part, cin,
                           cout, dW, db);
    } else {
        hipLaunchKernelGGL(k_wgrad<4>, dim3(nwg), dim3(256), wgrad_lds(4), s, V, in, g, gpad, cout, d->part);
        hipLaunchKernelGGL(k_wgrad_sum<4>, dim3((WG_TILES4 + 1) * 256 / 64, 8), dim3(256), 0, s, nwg, d->part, cin,
                           cout, dW, db);
    }
    PK_HIP(d->ctx, hipGetLastError());
    return PK_OK;
}

// loss of d->out against d->tgt (added to *loss) and gradients accumulated into d->g
int backward(pk_dn* d, const Vol& V, int margin, double inv_count, double* loss)
{
    hipStream_t s = d->ctx->stream;
    int r;
    const long nt = V.nv * V.ns;
    const unsigned nl = nblk(nt, 256) < 128 ? nblk(nt, 256) : 128;   // few workgroups: one loss atomic each
    hipLaunchKernelGGL(k_loss, dim3(nl), dim3(256), 0, s, V, margin, d->out, d->tgt, inv_count, d->g3, loss);
    if ((r = wgrad(d, V, d->a2, 16, HID, d->g3, 4, 3, d->g + OFF_W3, d->g + OFF_B3))) return r;
    const unsigned nb = conv_blocks(V);
    hipLaunchKernelGGL((k_conv<4, OUT_MASKED_BF16>), dim3(nb), dim3(256), conv_lds(4), s, V, d->g3, d->wt[1], (const float*)nullptr,
                       HID, d->dz2, (float*)nullptr, d->a2);
    if ((r = wgrad(d, V, d->a1, 16, HID, d->dz2, 16, HID, d->g + OFF_W2, d->g + OFF_B2))) return r;
    hipLaunchKernelGGL((k_conv<16, OUT_MASKED_BF16>), dim3(nb), dim3(256), conv_lds(16), s, V, d->dz2, d->wt[0],
                       (const float*)nullptr, HID, d->dz1, (float*)nullptr, d->a1);
    if ((r = wgrad(d, V, d->xb, 4, 3, d->dz1, 16, HID, d->g + OFF_W1, d->g + OFF_B1))) return r;
    PK_HIP(d->ctx, hipGetLastError());
    return PK_OK;
}

int check_dims(pk_dn* d, const int n[3], Vol& V)
{
    if (!n) return pk_fail(d->ctx, PK_E_ARG, "denoiser: null dimensions");
    for (int k = 0; k < 3; k++)
        if (n[k] < 1 || n[k] > 4096) return pk_fail(d->ctx, PK_E_ARG, "denoiser: volume edge out of range");
    V.n0 = n[0];
    V.n1 = n[1];
    V.n2 = n[2];
    V.nv = (long)n[0] * n[1] * n[2];
    if (V.nv >= (1l << 31)) return pk_fail(d->ctx, PK_E_ARG, "denoiser: volume above 2^31 voxels");
    return ensure_cap(d, V.nv);
}

// V becomes ns stacked samples of its size; the work buffers grow to hold them
int stack_samples(pk_dn* d, Vol& V, int ns)
{
    if ((long)ns * V.nv >= (1l << 31)) return pk_fail(d->ctx, PK_E_ARG, "denoiser: batch above 2^31 voxels");
    V.ns = ns;
    return ensure_cap(d, V.nv * ns);
}

// host [3][nv] FP32 -> d->xb (BF16, channel-last)
int load_input_host(pk_dn* d, const Vol& V, const float* x)
{
    pk_ctx* c = d->ctx;
    PK_HIP(c, hipMemcpyAsync(d->tgt, x, sizeof(float) * 3 * V.nv, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_input_f32, dim3(nblk(V.nv, 256)), dim3(256), 0, c->stream, V.nv, d->tgt, 1.0, d->xb);
    PK_HIP(c, hipGetLastError());
    return PK_OK;
}

uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int pk_dn_create(pk_ctx* c, pk_dn** out)
{
    if (!c || !out) return PK_E_ARG;
    *out = nullptr;
    pk_dn* d = new pk_dn();
    d->ctx = c;
    int r = 0;
    for (float** p : {&d->p, &d->g, &d->m, &d->v})
        if (!r) r = dalloc(c, (void**)p, sizeof(float) * NPARAM);
    if (!r) r = dalloc(c, (void**)&d->wf[0], sizeof(uint16_t) * WF4);
    if (!r) r = dalloc(c, (void**)&d->wf[1], sizeof(uint16_t) * WF16);
    if (!r) r = dalloc(c, (void**)&d->wf[2], sizeof(uint16_t) * WF16);
    if (!r) r = dalloc(c, (void**)&d->wt[0], sizeof(uint16_t) * WF16);
    if (!r) r = dalloc(c, (void**)&d->wt[1], sizeof(uint16_t) * WF4);
    if (!r) r = dalloc(c, (void**)&d->loss, sizeof(double));
    if (!r) r = dalloc(c, (void**)&d->part, sizeof(float) * WGRAD_WG_MAX * (WG_TILES16 + 1) * 256);
    if (r) {
        pk_dn_destroy(d);
        return r;
    }
    for (float* p : {d->p, d->g, d->m, d->v})
        hipLaunchKernelGGL(k_zero_f32, dim3(nblk(NPARAM, 256)), dim3(256), 0, c->stream, (long)NPARAM, p);
    r = pack_weights(d);
    if (r) {
        pk_dn_destroy(d);
        return r;
    }
    *out = d;
    return PK_OK;
}

void pk_dn_destroy(pk_dn* d)
{
    if (!d) return;
    (void)hipStreamSynchronize(d->ctx->stream);
    for (void* p : {(void*)d->p, (void*)d->g, (void*)d->m, (void*)d->v, (void*)d->wf[0], (void*)d->wf[1],
                    (void*)d->wf[2], (void*)d->wt[0], (void*)d->wt[1], (void*)d->xb, (void*)d->a1, (void*)d->a2,
                    (void*)d->g3, (void*)d->dz2, (void*)d->dz1, (void*)d->out, (void*)d->tgt, (void*)d->loss, (void*)d->part})
        if (p) (void)hipFree(p);
    for (auto& pr : d->pairs) {
        if (pr.low) (void)hipFree(pr.low);
        if (pr.high) (void)hipFree(pr.high);
    }
    if (d->cgain) (void)hipFree(d->cgain);
    delete d;
}

int pk_dn_num_params(void) { return NPARAM; }

int pk_dn_set_params(pk_dn* d, const float* p)
{
    if (!d || !p) return PK_E_ARG;
    pk_ctx* c = d->ctx;
    PK_HIP(c, hipMemcpyAsync(d->p, p, sizeof(float) * NPARAM, hipMemcpyHostToDevice, c->stream));
    for (float* q : {d->m, d->v})
        hipLaunchKernelGGL(k_zero_f32, dim3(nblk(NPARAM, 256)), dim3(256), 0, c->stream, (long)NPARAM, q);
    d->adam_t = 0;
    int r = pack_weights(d);
    if (r) return r;
    PK_HIP(c, hipStreamSynchronize(c->stream));
    return PK_OK;
}

int pk_dn_get_params(pk_dn* d, float* p)
{
    if (!d || !p) return PK_E_ARG;
    PK_HIP(d->ctx, hipMemcpyAsync(p, d->p, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, d->ctx->stream));
    PK_HIP(d->ctx, hipStreamSynchronize(d->ctx->stream));
    return PK_OK;
}

int pk_dn_init(pk_dn* d, uint64_t seed)
{
    if (!d) return PK_E_ARG;
    // He-normal weights (std sqrt(2 / fan_in), fan_in = 27 cin), zero biases; SplitMix64 counter
    // stream and Box-Muller, so the same seed gives the same parameters everywhere
    std::vector<float> p(NPARAM, 0.0f);
    const int woff[3] = {OFF_W1, OFF_W2, OFF_W3}, wn[3] = {27 * 3 * HID, 27 * HID * HID, 27 * HID * 3},
              fan[3] = {27 * 3, 27 * HID, 27 * HID};
    for (int l = 0; l < 3; l++) {
        const double sd = sqrt(2.0 / fan[l]);
        for (int i = 0; i < wn[l]; i++) {
            const uint64_t k = mix64(seed ^ mix64(((uint64_t)l << 40) + (uint64_t)i));
            const double u1 = ((k >> 11) + 0.5) * 0x1.0p-53, u2 = ((mix64(k) >> 11) + 0.5) * 0x1.0p-53;
            p[woff[l] + i] = (float)(sd * sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2));
        }
    }
    return pk_dn_set_params(d, p.data());
}

int pk_dn_set_charge_gain(pk_dn* d, int nbins, double kmax, const double* gain)
{
    if (!d) return PK_E_ARG;
    pk_ctx* c = d->ctx;
    if (nbins == 0) {
        d->cgain_n = 0;
        return PK_OK;
    }
    if (nbins < 2 || nbins > PK_DN_MAX_GAIN || !gain || !(kmax > 0.0) || !std::isfinite(kmax))
        return pk_fail(c, PK_E_ARG, "pk_dn_set_charge_gain: need 2..1024 bins, kmax > 0 and a gain table");
    for (int b = 0; b < nbins; b++)
        if (!std::isfinite(gain[b])) return pk_fail(c, PK_E_ARG, "pk_dn_set_charge_gain: gains must be finite");
    if (gain[0] != 1.0) return pk_fail(c, PK_E_ARG, "pk_dn_set_charge_gain: g(0) must be 1 (the filter keeps the charge)");
    if (!d->cgain && hipMalloc((void**)&d->cgain, sizeof(double) * PK_DN_MAX_GAIN) != hipSuccess) {
        d->cgain = nullptr;
        (void)hipGetLastError();
        return pk_fail(c, PK_E_ALLOC, "pk_dn_set_charge_gain: hipMalloc failed");
    }
    PK_HIP(c, hipStreamSynchronize(c->stream));
    PK_HIP(c, hipMemcpy(d->cgain, gain, sizeof(double) * nbins, hipMemcpyHostToDevice));
    d->cgain_n = nbins;
    d->cgain_kmax = kmax;
    return PK_OK;
}


int pk_dn_set_scale(pk_dn* d, double scale)
{
    if (!d || !(scale > 0.0)) return PK_E_ARG;
    d->scale = scale;
    return PK_OK;
}

int pk_dn_forward_host(pk_dn* d, const int n[3], const float* x, float* y)
{
    if (!d || !x || !y) return PK_E_ARG;
    Vol V;
    int r = check_dims(d, n, V);
    if (r) return r;
    if ((r = load_input_host(d, V, x))) return r;
    if ((r = forward(d, V))) return r;
    PK_HIP(d->ctx, hipMemcpyAsync(y, d->out, sizeof(float) * 3 * V.nv, hipMemcpyDeviceToHost, d->ctx->stream));
    PK_HIP(d->ctx, hipStreamSynchronize(d->ctx->stream));
    return PK_OK;
}

int pk_dn_layer_host(pk_dn* d, int layer, const int n[3], const float* in, float* out)
{
    if (!d || !in || !out || layer < 1 || layer > 3) return PK_E_ARG;
    Vol V;
    int r = check_dims(d, n, V);
    if (r) return r;
    pk_ctx* c = d->ctx;
    const int cin = layer == 1 ? 3 : HID, cout = layer == 3 ? 3 : HID;
    // input [cin][nv] FP32 -> BF16 channel-last via host rounding (exactly the device's rounding)
    const int pad = layer == 1 ? 4 : 16;
    std::vector<uint16_t> xb((size_t)pad * V.nv, 0);
    for (long v = 0; v < V.nv; v++)
        for (int ch = 0; ch < cin; ch++) xb[(size_t)v * pad + ch] = to_bf16(in[(size_t)ch * V.nv + v]);
    uint16_t* buf = layer == 1 ? d->xb : d->a1;
    PK_HIP(c, hipMemcpyAsync(buf, xb.data(), sizeof(uint16_t) * xb.size(), hipMemcpyHostToDevice, c->stream));
    float* res = nullptr;
    if ((r = dalloc(c, (void**)&res, sizeof(float) * (size_t)cout * V.nv))) return r;
    const unsigned nb = conv_blocks(V);
    const float* bias = d->p + (layer == 1 ? OFF_B1 : layer == 2 ? OFF_B2 : OFF_B3);
    if (layer == 1)
        hipLaunchKernelGGL((k_conv<4, OUT_F32>), dim3(nb), dim3(256), conv_lds(4), c->stream, V, buf, d->wf[0], bias, cout,
                           (uint16_t*)nullptr, res, (const uint16_t*)nullptr);
    else
        hipLaunchKernelGGL((k_conv<16, OUT_F32>), dim3(nb), dim3(256), conv_lds(16), c->stream, V, buf, d->wf[layer - 1], bias,
                           cout, (uint16_t*)nullptr, res, (const uint16_t*)nullptr);
    r = hipGetLastError() == hipSuccess ? PK_OK : pk_fail(c, PK_E_HIP, "pk_dn_layer_host: launch failed");
    if (!r && (hipMemcpyAsync(out, res, sizeof(float) * (size_t)cout * V.nv, hipMemcpyDeviceToHost, c->stream) !=
                   hipSuccess ||
               hipStreamSynchronize(c->stream) != hipSuccess))
        r = pk_fail(c, PK_E_HIP, "pk_dn_layer_host: copy failed");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(res);
    return r;
}

int pk_dn_grad_batch_host(pk_dn* d, const int n[3], int nsamples, const float* x, const float* y, float* grad,
                          double* loss)
{
    if (!d || !x || !y || !grad || nsamples < 1) return PK_E_ARG;
    Vol V;
    int r = check_dims(d, n, V);
    if (!r) r = stack_samples(d, V, nsamples);
    if (r) return r;
    pk_ctx* c = d->ctx;
    hipLaunchKernelGGL(k_zero_f32, dim3(nblk(NPARAM, 256)), dim3(256), 0, c->stream, (long)NPARAM, d->g);
    hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(1), 0, c->stream, d->loss);
    // inputs pass through d->tgt ([ns][3][nv] FP32) on their way to d->xb
    PK_HIP(c, hipMemcpyAsync(d->tgt, x, sizeof(float) * 3 * V.nv * V.ns, hipMemcpyHostToDevice, c->stream));
    for (int k = 0; k < V.ns; k++)
        hipLaunchKernelGGL(k_input_f32, dim3(nblk(V.nv, 256)), dim3(256), 0, c->stream, V.nv, d->tgt + 3 * V.nv * k,
                           1.0, d->xb + 4 * V.nv * k);
    if ((r = forward(d, V))) return r;
    PK_HIP(c, hipMemcpyAsync(d->tgt, y, sizeof(float) * 3 * V.nv * V.ns, hipMemcpyHostToDevice, c->stream));
    if ((r = backward(d, V, 0, 1.0 / (3.0 * V.nv * V.ns), d->loss))) return r;
    PK_HIP(c, hipMemcpyAsync(grad, d->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, c->stream));
    double l = 0.0;
    PK_HIP(c, hipMemcpyAsync(&l, d->loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PK_HIP(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = l;
    return PK_OK;
}

int pk_dn_grad_host(pk_dn* d, const int n[3], const float* x, const float* y, float* grad, double* loss)
{
    return pk_dn_grad_batch_host(d, n, 1, x, y, grad, loss);
}

int pk_dn_add_pair(pk_dn* d, pk_grid* low, pk_grid* high)
{
    if (!d || !low || !high) return PK_E_ARG;
    if (low->ctx != d->ctx || high->ctx != d->ctx) return pk_fail(d->ctx, PK_E_ARG, "pk_dn_add_pair: other context");
    pk_dn_pair pr;
    for (int k = 0; k < 3; k++) {
        pr.n[k] = low->s[k] - 2 * low->ghost;
        if (high->s[k] - 2 * high->ghost != pr.n[k] || low->prec != high->prec)
            return pk_fail(d->ctx, PK_E_ARG, "pk_dn_add_pair: grids differ in size or precision");
    }
    Vol V{pr.n[0], pr.n[1], pr.n[2], (long)pr.n[0] * pr.n[1] * pr.n[2]};
    int r;
    if ((r = dalloc(d->ctx, (void**)&pr.low, sizeof(float) * 3 * V.nv))) return r;
    if ((r = dalloc(d->ctx, (void**)&pr.high, sizeof(float) * 3 * V.nv))) {
        (void)hipFree(pr.low);
        return r;
    }
    hipStream_t s = d->ctx->stream;
    for (int k = 0; k < 2; k++) {
        pk_grid* g = k ? high : low;
        float* dst = k ? pr.high : pr.low;
        if (g->prec == PK_F64)
            hipLaunchKernelGGL(k_grid_to_f32<double>, dim3(nblk(V.nv, 256)), dim3(256), 0, s, grid_view<double>(g), V,
                               d->scale, dst);
        else
            hipLaunchKernelGGL(k_grid_to_f32<float>, dim3(nblk(V.nv, 256)), dim3(256), 0, s, grid_view<float>(g), V,
                               d->scale, dst);
    }
    PK_HIP(d->ctx, hipGetLastError());
    d->pairs.push_back(pr);
    return PK_OK;
}

int pk_dn_add_pair_crops(pk_dn* d, pk_grid* low, pk_grid* high, int ncrops, int edge, uint64_t seed, int* offsets)
{
    if (!d || !low || !high || ncrops < 0) return PK_E_ARG;
    if (low->ctx != d->ctx || high->ctx != d->ctx)
        return pk_fail(d->ctx, PK_E_ARG, "pk_dn_add_pair_crops: other context");
    int n[3];
    for (int k = 0; k < 3; k++) {
        n[k] = low->s[k] - 2 * low->ghost;
        if (high->s[k] - 2 * high->ghost != n[k] || low->prec != high->prec)
            return pk_fail(d->ctx, PK_E_ARG, "pk_dn_add_pair_crops: grids differ in size or precision");
        if (edge < 2 * PK_DN_RADIUS + 1 || edge > n[k])
            return pk_fail(d->ctx, PK_E_ARG, "pk_dn_add_pair_crops: crop edge outside [7, grid edge]");
    }
    Vol V{edge, edge, edge, (long)edge * edge * edge};
    hipStream_t s = d->ctx->stream;
    for (int i = 0; i < ncrops; i++) {
        pk_dn_pair pr;
        pr.n[0] = pr.n[1] = pr.n[2] = edge;
        pr.margin = PK_DN_RADIUS;
        int o[3];
        for (int k = 0; k < 3; k++)
            o[k] = (int)(mix64(seed ^ mix64(((uint64_t)i << 2) + (uint64_t)k)) % (uint64_t)n[k]);
        if (offsets)
            for (int k = 0; k < 3; k++) offsets[3 * i + k] = o[k];
        int r;
        if ((r = dalloc(d->ctx, (void**)&pr.low, sizeof(float) * 3 * V.nv))) return r;
        if ((r = dalloc(d->ctx, (void**)&pr.high, sizeof(float) * 3 * V.nv))) {
            (void)hipFree(pr.low);
            return r;
        }
        for (int k = 0; k < 2; k++) {
            pk_grid* g = k ? high : low;
            float* dst = k ? pr.high : pr.low;
            if (g->prec == PK_F64)
                hipLaunchKernelGGL(k_grid_crop_to_f32<double>, dim3(nblk(V.nv, 256)), dim3(256), 0, s,
                                   grid_view<double>(g), n[0], n[1], n[2], V, o[0], o[1], o[2], d->scale, dst);
            else
                hipLaunchKernelGGL(k_grid_crop_to_f32<float>, dim3(nblk(V.nv, 256)), dim3(256), 0, s,
                                   grid_view<float>(g), n[0], n[1], n[2], V, o[0], o[1], o[2], d->scale, dst);
        }
        PK_HIP(d->ctx, hipGetLastError());
        d->pairs.push_back(pr);
    }
    return PK_OK;
}

int pk_dn_num_pairs(pk_dn* d) { return d ? (int)d->pairs.size() : -1; }

int pk_dn_clear_pairs(pk_dn* d)
{
    if (!d) return PK_E_ARG;
    (void)hipStreamSynchronize(d->ctx->stream);
    for (auto& pr : d->pairs) {
        if (pr.low) (void)hipFree(pr.low);
        if (pr.high) (void)hipFree(pr.high);
    }
    d->pairs.clear();
    return PK_OK;
}

int pk_dn_train(pk_dn* d, int iters, int batch, double lr, uint64_t seed, float* losses)
{
    if (!d || iters < 0 || batch < 1 || !(lr > 0.0)) return PK_E_ARG;
    if (d->pairs.empty()) return pk_fail(d->ctx, PK_E_STATE, "pk_dn_train: no training pairs");
    if ((long)iters * batch >= (1l << 31)) return pk_fail(d->ctx, PK_E_ARG, "pk_dn_train: iters x batch too large");
    pk_ctx* c = d->ctx;
    const int npairs = (int)d->pairs.size();
    // every mini-batch's pair indices, and the pairs' buffers, on the device once
    std::vector<int> hsel((size_t)iters * batch);
    for (int it = 0; it < iters; it++)
        for (int b = 0; b < batch; b++)
            hsel[(size_t)it * batch + b] = (int)(mix64(seed ^ mix64(((uint64_t)it << 20) + b)) % npairs);
    std::vector<const float*> hptr(2 * (size_t)npairs);
    for (int k = 0; k < npairs; k++) {
        hptr[k] = d->pairs[k].low;
        hptr[npairs + k] = d->pairs[k].high;
    }
    double* dev_losses = nullptr;
    int* dsel = nullptr;
    const float** dptr = nullptr;
    int r = dalloc(c, (void**)&dev_losses, sizeof(double) * (iters > 0 ? iters : 1));
    if (!r) r = dalloc(c, (void**)&dsel, sizeof(int) * (hsel.empty() ? 1 : hsel.size()));
    if (!r) r = dalloc(c, (void**)&dptr, sizeof(float*) * hptr.size());
    if (!r && iters > 0) {
        const size_t bsel = sizeof(int) * hsel.size(), bptr = sizeof(float*) * hptr.size();
        if (hipMemcpyAsync(dsel, hsel.data(), bsel, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(dptr, hptr.data(), bptr, hipMemcpyHostToDevice, c->stream) != hipSuccess)
            r = pk_fail(c, PK_E_HIP, "pk_dn_train: upload failed");
    }
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    if (!r) {
        // k_adam clears the gradient after each step
        hipLaunchKernelGGL(k_zero_f32, dim3(nblk(NPARAM, 256)), dim3(256), 0, c->stream, (long)NPARAM, d->g);
        hipLaunchKernelGGL(k_zero_f32, dim3(nblk(2l * iters, 256) + 1), dim3(256), 0, c->stream, 2l * iters,
                           (float*)dev_losses);
    }
    for (int it = 0; it < iters && !r; it++) {
        const int* sel = hsel.data() + (size_t)it * batch;
        bool same = true;
        for (int b = 1; b < batch; b++) {
            same = same && d->pairs[sel[b]].margin == d->pairs[sel[0]].margin;
            for (int k = 0; k < 3; k++) same = same && d->pairs[sel[b]].n[k] == d->pairs[sel[0]].n[k];
        }
        // samples of one size go through every layer together (stacked); otherwise one by one
        const int groups = same ? 1 : batch, per = same ? batch : 1;
        for (int gi = 0; gi < groups && !r; gi++) {
            Vol V;
            if ((r = check_dims(d, d->pairs[sel[gi * per]].n, V)) || (r = stack_samples(d, V, per))) break;
            hipLaunchKernelGGL(k_gather, dim3(nblk(V.nv, 256), per), dim3(256), 0, c->stream, V.nv, dptr, dptr + npairs,
                               dsel + (size_t)it * batch + gi * per, d->xb, d->tgt);
            if ((r = forward(d, V))) break;
            const int m = d->pairs[sel[gi * per]].margin;
            const double counted = (double)(V.n0 - 2 * m) * (V.n1 - 2 * m) * (V.n2 - 2 * m);
            r = backward(d, V, m, 1.0 / (3.0 * counted * batch), dev_losses + it);
        }
        if (r) break;
        d->adam_t += 1;
        const float c1 = 1.0f - powf(b1, (float)d->adam_t), c2 = 1.0f - powf(b2, (float)d->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(nblk(NPARAM, 256)), dim3(256), 0, c->stream, NPARAM, d->p, d->g, d->m, d->v,
                           (float)lr, b1, b2, eps, c1, c2);
        if ((r = pack_weights(d))) break;
    }
    if (!r && losses && iters > 0) {
        std::vector<double> h(iters);
        if (hipMemcpyAsync(h.data(), dev_losses, sizeof(double) * iters, hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            r = pk_fail(c, PK_E_HIP, "pk_dn_train: loss copy failed");
        for (int it = 0; it < iters && !r; it++) losses[it] = (float)h[it];
    }
    (void)hipStreamSynchronize(c->stream);
    for (void* p : {(void*)dev_losses, (void*)dsel, (void*)dptr})
        if (p) (void)hipFree(p);
    return r;
}

int pk_dn_apply(pk_dn* d, pk_grid* g) { return pk_dn_apply_impl(d, g, true); }

}  // extern "C"

// pk_dn_apply; fill = false leaves J's ghosts as they were (for a caller that reads J only through
// the physical cells until it fills them itself); div0 != null receives D J of the input current
// (FP64, one value per physical node)
int pk_dn_apply_impl(pk_dn* d, pk_grid* g, bool fill, double* div0)
{
    if (!d || !g) return PK_E_ARG;
    if (g->ctx != d->ctx) return pk_fail(d->ctx, PK_E_ARG, "pk_dn_apply: other context");
    int n[3];
    for (int k = 0; k < 3; k++) n[k] = g->s[k] - 2 * g->ghost;
    Vol V;
    int r = check_dims(d, n, V);
    if (r) return r;
    hipStream_t s = d->ctx->stream;
    pk_ktimer tm(d->ctx, "denoise");
    // J -> network input in one pass; the output goes back through k_f32_to_grid (coalesced),
    // which measured faster than writing the grid from the last layer's epilogue
    if (g->prec == PK_F64)
        hipLaunchKernelGGL(k_grid_to_input<double>, dim3(nblk(V.nv, 256)), dim3(256), 0, s, grid_view<double>(g), V,
                           d->scale, d->xb, div0);
    else
        hipLaunchKernelGGL(k_grid_to_input<float>, dim3(nblk(V.nv, 256)), dim3(256), 0, s, grid_view<float>(g), V,
                           d->scale, d->xb, div0);
    if ((r = forward(d, V))) return r;
    if (g->prec == PK_F64)
        hipLaunchKernelGGL(k_f32_to_grid<double>, dim3(nblk(V.nv, 256)), dim3(256), 0, s, grid_view<double>(g), V,
                           1.0 / d->scale, d->out);
    else
        hipLaunchKernelGGL(k_f32_to_grid<float>, dim3(nblk(V.nv, 256)), dim3(256), 0, s, grid_view<float>(g), V,
                           1.0 / d->scale, d->out);
    PK_HIP(d->ctx, hipGetLastError());
    if (fill && g->ghost > 0) return pk_field_ghost(g, PK_FILL_J);
    return PK_OK;
}

// the device gain table of pk_step_denoised's projection (n = 0: none)
void pk_dn_charge_gain(const pk_dn* d, const double** gain, int* n, double* kmax)
{
    *gain = d->cgain;
    *n = d->cgain_n;
    *kmax = d->cgain_kmax;
}
