// nn.hip -- the kernels of the learned porosimetry surrogate: a 3-D CNN
// trained and run on BF16 operands with FP32 accumulation.
//
// Every convolution and dense layer is one GEMM on the matrix cores
// (v_mfma_f32_16x16x16_bf16): forward Y = patches(X) . W^T, weight
// gradient dW = dY^T . patches(X), input gradient dpatches = dY . W
// followed by a gather back onto the input grid.  Activations are
// channels-last [N][D][H][W][C] in BF16 (stored as raw uint16 bits);
// weights [Cout][kd][kh][kw][Cin] in BF16 with FP32 master copies.
#include "common.hpp"

namespace pmx {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---------------------------------------------------------------------------
// GEMM: C[M x N] (+)= op(A)[M x K] . op(B)[K x N]
// ---------------------------------------------------------------------------
// op(A)(m, k) = A[m*lda + k], or A[k*lda + m] when a_trans; likewise B.
// A 256-thread workgroup owns a 64 x 64 tile of C; wave w computes the
// 32 x 32 quadrant (w / 2, w % 2) as 2 x 2 MFMA tiles of 16 x 16.  K advances
// 32 at a time through LDS (zero-filled past M, N, K).
//
// MFMA 16x16x16 bf16 operand layout (lane l of 64): A rows l % 16, k =
// 4 (l / 16) .. +3; B columns l % 16, the same k; result column l % 16, rows
// 4 (l / 16) .. +3.
constexpr int kTM = 64, kTN = 64, kTK = 32;

__global__ __launch_bounds__(256) void k_gemm(int64_t M, int N, int K, const uint16_t *__restrict__ A, int64_t lda,
                                              int a_trans, const uint16_t *__restrict__ B, int64_t ldb, int b_trans,
                                              float *__restrict__ C, int64_t ldc, int accumulate, int k_chunk)
{
    __shared__ uint16_t As[kTM][kTK + 8];  // [m][k]
    __shared__ uint16_t Bs[kTN][kTK + 8];  // [n][k] (k contiguous for the operand gather)
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t m0 = (int64_t)blockIdx.x * kTM;
    const int n0 = blockIdx.y * kTN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // split K: block z covers [z k_chunk, (z+1) k_chunk) and adds atomically
    const int kb = blockIdx.z * k_chunk, ke = min(K, kb + k_chunk);
    for (int k0 = kb; k0 < ke; k0 += kTK) {
        // stage: 64 x 32 elements of each operand, 8 per thread
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = tid + 256 * e;  // 0 .. 2047
            int r, k;
            if (a_trans) {  // consecutive threads walk m (contiguous in memory)
                r = idx % kTM;
                k = idx / kTM;
            } else {
                r = idx / kTK;
                k = idx % kTK;
            }
            const int64_t gm = m0 + r;
            const int gk = k0 + k;
            uint16_t v = 0;
            if (gm < M && gk < ke) v = a_trans ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk];
            As[r][k] = v;
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int idx = tid + 256 * e;
            int c, k;
            if (b_trans) {  // B stored [n][k]: k contiguous
                c = idx / kTK;
                k = idx % kTK;
            } else {  // B stored [k][n]: n contiguous
                c = idx % kTN;
                k = idx / kTN;
            }
            const int gn = n0 + c, gk = k0 + k;
            uint16_t v = 0;
            if (gn < N && gk < ke) v = b_trans ? B[(int64_t)gn * ldb + gk] : B[(int64_t)gk * ldb + gn];
            Bs[c][k] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 16) {
            bf16x4 a[2], b[2];
            const int kq = kk + 4 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const uint16_t *pa = &As[wm + 16 * i + (lane & 15)][kq];
                a[i] = bf16x4{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
                const uint16_t *pb = &Bs[wn + 16 * i + (lane & 15)][kq];
                b[i] = bf16x4{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = n0 + wn + 16 * j + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
                if (row < M && col < N) {
                    float *p = C + (int64_t)row * ldc + col;
                    if (gridDim.z > 1)
                        atomicAdd(p, acc[i][j][v]);
                    else
                        *p = accumulate ? *p + acc[i][j][v] : acc[i][j][v];
                }
            }
        }
}

// ---------------------------------------------------------------------------
// convolution helpers: kernel 3, padding 1, stride s
// ---------------------------------------------------------------------------
struct ConvDims {
    int n, d, h, w, c;  // input
    int od, oh, ow;     // output
    int s;
    int64_t ldk;        // row length of the patch matrix (>= 27 c)
};

// patches[m][k]: m = ((b*od + z)*oh + y)*ow + x, k = ((kz*3 + ky)*3 + kx)*c + ci
__global__ __launch_bounds__(256) void k_im2col(const uint16_t *__restrict__ x, ConvDims cd,
                                                uint16_t *__restrict__ cols)
{
    const int64_t kk = 27 * (int64_t)cd.c;
    const int64_t total = (int64_t)cd.n * cd.od * cd.oh * cd.ow * cd.ldk;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t m = i / cd.ldk, k = i - m * cd.ldk;
        uint16_t v = 0;
        if (k < kk) {
            const int ci = (int)(k % cd.c), t = (int)(k / cd.c);
            const int kx = t % 3, ky = (t / 3) % 3, kz = t / 9;
            int64_t r = m;
            const int ox = (int)(r % cd.ow); r /= cd.ow;
            const int oy = (int)(r % cd.oh); r /= cd.oh;
            const int oz = (int)(r % cd.od);
            const int b = (int)(r / cd.od);
            const int ix = ox * cd.s - 1 + kx, iy = oy * cd.s - 1 + ky, iz = oz * cd.s - 1 + kz;
            if (ix >= 0 && ix < cd.w && iy >= 0 && iy < cd.h && iz >= 0 && iz < cd.d)
                v = x[((((int64_t)b * cd.d + iz) * cd.h + iy) * cd.w + ix) * cd.c + ci];
        }
        cols[i] = v;
    }
}

// dx[b][z][y][x][ci] = sum of dcols over the output positions whose patch
// covers the voxel (a gather: deterministic, no atomics)
__global__ __launch_bounds__(256) void k_col2im(const float *__restrict__ dcols, ConvDims cd, float *__restrict__ dx)
{
    const int64_t total = (int64_t)cd.n * cd.d * cd.h * cd.w * cd.c;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = i;
        const int ci = (int)(r % cd.c); r /= cd.c;
        const int ix = (int)(r % cd.w); r /= cd.w;
        const int iy = (int)(r % cd.h); r /= cd.h;
        const int iz = (int)(r % cd.d);
        const int b = (int)(r / cd.d);
        float acc = 0.f;
        for (int kz = 0; kz < 3; ++kz) {
            const int tz = iz + 1 - kz;
            if (tz < 0 || tz % cd.s) continue;
            const int oz = tz / cd.s;
            if (oz >= cd.od) continue;
            for (int ky = 0; ky < 3; ++ky) {
                const int ty = iy + 1 - ky;
                if (ty < 0 || ty % cd.s) continue;
                const int oy = ty / cd.s;
                if (oy >= cd.oh) continue;
                for (int kx = 0; kx < 3; ++kx) {
                    const int tx = ix + 1 - kx;
                    if (tx < 0 || tx % cd.s) continue;
                    const int ox = tx / cd.s;
                    if (ox >= cd.ow) continue;
                    const int64_t m = (((int64_t)b * cd.od + oz) * cd.oh + oy) * cd.ow + ox;
                    acc += dcols[m * cd.ldk + ((kz * 3 + ky) * 3 + kx) * cd.c + ci];
                }
            }
        }
        dx[i] = acc;
    }
}

// y[m][c] + bias[c], optional ReLU; writes BF16 (and FP32 when requested)
__global__ __launch_bounds__(256) void k_bias_act(const float *__restrict__ y, const float *__restrict__ bias,
                                                  int64_t m, int c, int relu, uint16_t *__restrict__ out_bf,
                                                  float *__restrict__ out_f)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m * c;
         i += (int64_t)gridDim.x * blockDim.x) {
        float v = y[i] + bias[i % c];
        if (relu) v = fmaxf(v, 0.f);
        if (out_bf) out_bf[i] = f2bf(v);
        if (out_f) out_f[i] = v;
    }
}

// dy * [activation > 0] (the BF16 activation decides), written as BF16 and FP32
__global__ __launch_bounds__(256) void k_relu_back(const float *__restrict__ dy, const uint16_t *__restrict__ act,
                                                   int64_t n, float *__restrict__ dx, uint16_t *__restrict__ dx_bf)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = (act == nullptr || bf2f(act[i]) > 0.f) ? dy[i] : 0.f;
        if (dx) dx[i] = v;
        if (dx_bf) dx_bf[i] = f2bf(v);
    }
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ a, int64_t n, uint16_t *__restrict__ o)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        o[i] = f2bf(a[i]);
}

__global__ __launch_bounds__(256) void k_u8_to_bf16(const uint8_t *__restrict__ a, int64_t n, uint16_t *__restrict__ o)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        o[i] = f2bf((float)a[i]);
}

// out[j] += sum_i a[i][j] (bias gradients; out zeroed by the caller): each
// block sums a chunk of rows with coalesced flat reads into LDS, then adds
// its n partial sums to out (float atomics: the order of the additions is
// not fixed)
__global__ __launch_bounds__(256) void k_col_sum(const float *__restrict__ a, int64_t m, int n, int64_t rows_per_block,
                                                 float *__restrict__ out)
{
    extern __shared__ float part[];
    for (int j = threadIdx.x; j < n; j += blockDim.x) part[j] = 0.f;
    __syncthreads();
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = min(m, r0 + rows_per_block);
    for (int64_t i = r0 * n + threadIdx.x; i < r1 * n; i += blockDim.x) atomicAdd(&part[i % n], a[i]);
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) atomicAdd(out + j, part[j]);
}

// global average over the spatial positions of each sample: [n][p][c] -> [n][c]
__global__ __launch_bounds__(256) void k_avgpool(const uint16_t *__restrict__ x, int n, int64_t p, int c,
                                                 float *__restrict__ out_f, uint16_t *__restrict__ out_bf)
{
    const int b = blockIdx.x / c, ch = blockIdx.x % c;
    float s = 0.f;
    for (int64_t i = threadIdx.x; i < p; i += blockDim.x) s += bf2f(x[((int64_t)b * p + i) * c + ch]);
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float v = red[0] / (float)p;
        if (out_f) out_f[blockIdx.x] = v;
        if (out_bf) out_bf[blockIdx.x] = f2bf(v);
    }
}

__global__ __launch_bounds__(256) void k_avgpool_back(const float *__restrict__ dy, int n, int64_t p, int c,
                                                      float *__restrict__ dx)
{
    const int64_t total = (int64_t)n * p * c;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int ch = (int)(i % c);
        const int b = (int)(i / (p * c));
        dx[i] = dy[b * c + ch] / (float)p;
    }
}

// Per-sample loss of pred = (curve[q], s = log variance) against the exact
// curve t: the curve learns by squared error, the variance head by the
// Gaussian negative log-likelihood of that (fixed) error,
//   L = se / 2 + (se e^-s + q s) / 2,   se = sum_j (y_j - t_j)^2,
// with s clamped to [-12, 12] in the value and passed straight through in
// the gradient.  dpred is the gradient of the batch mean (se held fixed in
// the variance term).
__global__ __launch_bounds__(64) void k_gnll(const float *__restrict__ pred, const float *__restrict__ target, int n,
                                             int q, float *__restrict__ loss, float *__restrict__ dpred)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const float *p = pred + (int64_t)b * (q + 1);
    const float s = fminf(fmaxf(p[q], -12.f), 12.f);
    const float inv = expf(-s);
    float se = 0.f;
    for (int j = 0; j < q; ++j) {
        const float e = p[j] - target[(int64_t)b * q + j];
        se += e * e;
        dpred[(int64_t)b * (q + 1) + j] = e / (float)n;
    }
    dpred[(int64_t)b * (q + 1) + q] = (-0.5f * se * inv + 0.5f * q) / (float)n;
    loss[b] = 0.5f * se + 0.5f * (se * inv + q * s);
}

// Adam on FP32 master weights, refreshing their BF16 copy
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                                              float *__restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2, uint16_t *__restrict__ w_bf)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (w_bf) w_bf[i] = f2bf(wi);
    }
}

// ---------------------------------------------------------------------------
// inference convolutions (kernel 3, padding 1, stride s; bias + ReLU + BF16
// rounding fused).  Operands are read through buffer resources: a load whose
// offset is kOob returns zero, so padding taps and tail lanes need no
// branch.  A launch covers a run of samples whose input is below 2 GiB.
// Activations may carry a one-voxel zero border ("padded"): the
// [n][d+2][h+2][w+2][c] buffer holds the tensor in its interior.
// ---------------------------------------------------------------------------

constexpr uint32_t kOob = 0x80000000u;  // >= every buffer's size
constexpr int kBufFlags = 0x00020000;   // raw buffer, 32-bit data format (CDNA)
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer(const void *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, kBufFlags);
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// two floats to a bf16 pair (round to nearest even; one v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi)
{
    return __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2_t{lo, hi}), bf16x2_t));
}

// max(v, 0) on the bits: a negative float is a negative int (no NaN canonicalisation)
__device__ __forceinline__ float relu_bits(float v) { return __int_as_float(max(__float_as_int(v), 0)); }

// First layer on the matrix cores: out^T (channels x positions) = W (channels
// x K) . P^T (K x positions) with K = 48 slots, slot 4r + kx for tap row r =
// 3kz + ky (slot 4r + 3 and rows 9..11 hold zero), so in each 16-deep MFMA
// step lane l (position l % 16, slots 4(l/16)..+3) needs the three bytes of
// one input row.  Wave = kFirstWaveRows output rows x 64 output columns, the
// loads of row r + 1 in flight while row r is multiplied; block = four waves.
// FAST (stride 2, w % 4 == 0): a row's three bytes come from two aligned
// dwords, else from three byte loads.
constexpr int kFirstWaveRows = 4, kFirstRows = 4 * kFirstWaveRows, kFirstCols = 64;

template <int MB, bool FAST>
__global__ __launch_bounds__(256) void k_conv_first(const uint8_t *__restrict__ x, uint32_t x_bytes, ConvDims cd,
                                                    int gy, int gx, int co, const uint16_t *__restrict__ w,
                                                    const float *__restrict__ bias, uint16_t *__restrict__ out,
                                                    int out_pad)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    unsigned r = (unsigned)xcd_tile(blockIdx.x, gridDim.x);
    const int bx = (int)(r % (unsigned)gx);
    r /= (unsigned)gx;
    const int by = (int)(r % (unsigned)gy);
    r /= (unsigned)gy;
    const int oz = (int)(r % (unsigned)cd.od);
    const int b = (int)(r / (unsigned)cd.od);
    const int oy0 = by * kFirstRows + wv * kFirstWaveRows;
    if (oy0 >= cd.oh) return;
    const __amdgpu_buffer_rsrc_t xb = buffer(x, x_bytes);
    // weights: A[ch = l % 16 + 16 mb][slot 16 j + 4 g + kx]
    bf16x4 wa[MB][3];
    float bs[MB][4];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int ch = 16 * mb + (lane & 15);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int row = 4 * j + g;
            short e[4] = {0, 0, 0, 0};
            if (ch < co && row < 9) {
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) e[kx] = (short)w[ch * 27 + row * 3 + kx];
            }
            wa[mb][j] = bf16x4{e[0], e[1], e[2], e[3]};
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 16 * mb + 4 * g + v;
            bs[mb][v] = c < co ? bias[c] : 0.f;
        }
    }
    // this lane's tap rows for the three MFMA steps: input plane/row offsets
    int tz[3], ty[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int row = 4 * j + g;
        tz[j] = row < 9 ? row / 3 : 1 << 20;  // rows 9..11: always out of range
        ty[j] = row - 3 * (row / 3);
    }
    const int plane = cd.h * cd.w;
    const int sample_off = b * cd.d * plane;
    const int ox_l = bx * kFirstCols + (lane & 15);
    const int x0_l = ox_l * cd.s - 1;
    const uint32_t shift = (uint32_t)(x0_l & 3);
    // raw operands of the output row in flight: 4 column blocks x 3 tap rows.  One
    // offset per tap row; the column blocks sit at constant distances (16 outputs =
    // 16 s bytes).  Loads past the row's end only feed outputs >= ow (never stored)
    // or the unused fourth byte; past the buffer they read zero.
    uint32_t rlo[4][3], rhi[4][3];
    const bool left = x0_l < 0;  // the lane of output column 0: its first byte is padding
    auto issue = [&](int rr) {
        const int oy = oy0 + rr;
        const bool row_ok = rr < kFirstWaveRows && oy < cd.oh;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int iz = oz * cd.s - 1 + tz[j], iy = oy * cd.s - 1 + ty[j];
            const bool in = row_ok && iz >= 0 && iz < cd.d && iy >= 0 && iy < cd.h;
            const int rowoff = sample_off + iz * plane + iy * cd.w;
            if (FAST) {
                // bytes 2ox-1 .. 2ox+1 lie in the dwords at a, a + 4 (a = -4 for output column 0:
                // that lane loads at 0 and moves the dword up, the padding byte comes from zero;
                // no offset may wrap, the two loads can be fused into one range-checked load)
                const uint32_t vo = in ? (uint32_t)(rowoff + (x0_l & ~3)) : kOob;  // blocks 1..3: >= 28
                const uint32_t v0 = in ? (uint32_t)(rowoff + max(x0_l & ~3, 0)) : kOob;
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    const uint32_t o = nb == 0 ? v0 : vo + 32 * nb;
                    rlo[nb][j] = __builtin_amdgcn_raw_buffer_load_b32(xb, o, 0, 0);
                    rhi[nb][j] = __builtin_amdgcn_raw_buffer_load_b32(xb, o + 4, 0, 0);
                }
                rhi[0][j] = left ? rlo[0][j] : rhi[0][j];
                rlo[0][j] = left ? 0u : rlo[0][j];
            } else {
#pragma unroll
                for (int nb = 0; nb < 4; ++nb) {
                    const int x0 = (ox_l + 16 * nb) * cd.s - 1;
                    const bool ok = in && ox_l + 16 * nb < cd.ow;
                    const uint32_t o = (uint32_t)(rowoff + x0);
                    const uint32_t b0 = __builtin_amdgcn_raw_buffer_load_b8(xb, ok && x0 >= 0 ? o : kOob, 0, 0);
                    const uint32_t b1 = __builtin_amdgcn_raw_buffer_load_b8(xb, ok && x0 + 1 < cd.w ? o + 1 : kOob, 0, 0);
                    const uint32_t b2 = __builtin_amdgcn_raw_buffer_load_b8(xb, ok && x0 + 2 < cd.w ? o + 2 : kOob, 0, 0);
                    rlo[nb][j] = b0 | (b1 << 8) | (b2 << 16);
                    rhi[nb][j] = 0u;
                }
            }
        }
    };
    const int odp = cd.od + 2 * out_pad, ohp = cd.oh + 2 * out_pad, owp = cd.ow + 2 * out_pad;
    issue(0);
    for (int rr = 0; rr < kFirstWaveRows; ++rr) {
        const int oy = oy0 + rr;
        bf16x4 tp[4][3];
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const uint32_t v = FAST ? __builtin_amdgcn_alignbyte(rhi[nb][j], rlo[nb][j], shift) : rlo[nb][j];
                // byte -> float -> bf16 (exact for 0..255); perm takes the high halves
                const float f0 = (float)(v & 0xffu), f1 = (float)((v >> 8) & 0xffu), f2 = (float)((v >> 16) & 0xffu);
                const uint32_t p01 = __builtin_amdgcn_perm(__float_as_uint(f1), __float_as_uint(f0), 0x07060302u);
                const uint32_t p2 = __float_as_uint(f2) >> 16;
                tp[nb][j] = bf16x4{(short)(p01 & 0xffffu), (short)(p01 >> 16), (short)p2, (short)0};
            }
        }
        issue(rr + 1);
        if (oy >= cd.oh) continue;  // wave-uniform
        uint16_t *orow = out + ((((int64_t)b * odp + oz + out_pad) * ohp + oy + out_pad) * owp + out_pad + ox_l) * co;
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int ox = ox_l + 16 * nb;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                floatx4 acc = {bs[mb][0], bs[mb][1], bs[mb][2], bs[mb][3]};  // the bias rides in the accumulator
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[mb][j], tp[nb][j], acc, 0, 0, 0);
                const int c0 = 16 * mb + 4 * g;
                if (ox < cd.ow && c0 < co) {
                    const uint32_t q0 = pack_bf16(relu_bits(acc[0]), relu_bits(acc[1]));
                    const uint32_t q1 = pack_bf16(relu_bits(acc[2]), relu_bits(acc[3]));
                    *reinterpret_cast<uint2 *>(orow + 16 * nb * co + c0) = make_uint2(q0, q1);
                }
            }
        }
    }
}

// The strided first layer with the block's input slab in LDS (stride 2, w a
// multiple of 16).  A block's 16 output rows x 64 output columns of one
// output plane read 3 input planes x 33 rows x 129 bytes; they are staged
// once with 16-byte loads (zeros outside the box), and each lane then takes
// its three bytes per tap row from LDS (two aligned dwords and alignbyte)
// instead of two buffer loads per tap row and column block: the direct form
// above issues 24 vector memory loads per lane per output row, mostly for
// bytes its neighbours also load.  The product and the epilogue are the
// same.
constexpr int kSlabRows = 2 * kFirstRows + 1;   // input rows per staged plane
constexpr int kSlabPitch = 16 + 2 * kFirstCols;  // 15 lead bytes, column -1, 128 columns

template <int MB>
__global__ __launch_bounds__(256) void k_conv_first_lds(const uint8_t *__restrict__ x, uint32_t x_bytes, ConvDims cd,
                                                        int gy, int gx, int co, const uint16_t *__restrict__ w,
                                                        const float *__restrict__ bias, uint16_t *__restrict__ out,
                                                        int out_pad)
{
    // + 16: the dword pair of the last column of the last row reads 4 bytes past the slab
    __shared__ __align__(16) uint8_t slab[3 * kSlabRows * kSlabPitch + 16];
    const int lane = threadIdx.x & 63, g = lane >> 4, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    unsigned r = (unsigned)xcd_tile(blockIdx.x, gridDim.x);
    const int bx = (int)(r % (unsigned)gx);
    r /= (unsigned)gx;
    const int by = (int)(r % (unsigned)gy);
    r /= (unsigned)gy;
    const int oz = (int)(r % (unsigned)cd.od);
    const int b = (int)(r / (unsigned)cd.od);
    const int oyb = by * kFirstRows;
    const __amdgpu_buffer_rsrc_t xb = buffer(x, x_bytes);
    const int plane = cd.h * cd.w;
    const int sample_off = b * cd.d * plane;
    const int ix0 = 2 * kFirstCols * bx;  // input column of slab byte 16
    // staging: per slab row, eight 16-byte chunks and the byte of column ix0 - 1
    for (int t = threadIdx.x; t < 3 * kSlabRows * 9; t += 256) {
        const int row = t / 9, part = t - 9 * row;
        const int pz = row / kSlabRows, py = row - pz * kSlabRows;
        const int iz = oz * 2 - 1 + pz, iy = oyb * 2 - 1 + py;
        const bool in = iz >= 0 && iz < cd.d && iy >= 0 && iy < cd.h;
        const uint32_t rowoff = (uint32_t)(sample_off + iz * plane + iy * cd.w);
        uint8_t *dst = slab + row * kSlabPitch;
        if (part < 8) {
            const int ix = ix0 + 16 * part;
            const uint32_t o = in && ix < cd.w ? rowoff + (uint32_t)ix : kOob;
            *reinterpret_cast<uint4 *>(dst + 16 + 16 * part) =
                __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(xb, o, 0, 0));
        } else {
            const uint32_t o = in && ix0 > 0 ? rowoff + (uint32_t)(ix0 - 1) : kOob;
            dst[15] = (uint8_t)__builtin_amdgcn_raw_buffer_load_b8(xb, o, 0, 0);
        }
    }
    // weights: A[ch = l % 16 + 16 mb][slot 16 j + 4 g + kx]
    bf16x4 wa[MB][3];
    float bs[MB][4];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb) {
        const int ch = 16 * mb + (lane & 15);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int row = 4 * j + g;
            short e[4] = {0, 0, 0, 0};
            if (ch < co && row < 9) {
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) e[kx] = (short)w[ch * 27 + row * 3 + kx];
            }
            wa[mb][j] = bf16x4{e[0], e[1], e[2], e[3]};
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int c = 16 * mb + 4 * g + v;
            bs[mb][v] = c < co ? bias[c] : 0.f;
        }
    }
    // this lane's tap rows for the three MFMA steps (rows 9..11 hold zero)
    int tz[3], ty[3];
    bool tap[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int row = 4 * j + g;
        tap[j] = row < 9;
        tz[j] = tap[j] ? row / 3 : 0;
        ty[j] = tap[j] ? row - 3 * (row / 3) : 0;
    }
    __syncthreads();
    const int oy0 = oyb + wv * kFirstWaveRows;
    if (oy0 >= cd.oh) return;  // wave-uniform, after the block's barrier
    const int cl = 2 * (lane & 15) + 15;  // slab byte of this lane's first input column (column block 0)
    const int odp = cd.od + 2 * out_pad, ohp = cd.oh + 2 * out_pad, owp = cd.ow + 2 * out_pad;
    const int ox_l = bx * kFirstCols + (lane & 15);
    for (int rr = 0; rr < kFirstWaveRows; ++rr) {
        const int oy = oy0 + rr;
        if (oy >= cd.oh) break;  // wave-uniform
        bf16x4 tp[4][3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int rowb = (tz[j] * kSlabRows + 2 * (oy - oyb) + ty[j]) * kSlabPitch;
#pragma unroll
            for (int nb = 0; nb < 4; ++nb) {
                const int o = rowb + cl + 32 * nb;
                const uint32_t *d = reinterpret_cast<const uint32_t *>(slab + (o & ~3));
                uint32_t v = __builtin_amdgcn_alignbyte(d[1], d[0], (uint32_t)(o & 3));
                v = tap[j] ? v : 0u;
                // byte -> float -> bf16 (exact for 0..255); perm takes the high halves
                const float f0 = (float)(v & 0xffu), f1 = (float)((v >> 8) & 0xffu), f2 = (float)((v >> 16) & 0xffu);
                const uint32_t p01 = __builtin_amdgcn_perm(__float_as_uint(f1), __float_as_uint(f0), 0x07060302u);
                const uint32_t p2 = __float_as_uint(f2) >> 16;
                tp[nb][j] = bf16x4{(short)(p01 & 0xffffu), (short)(p01 >> 16), (short)p2, (short)0};
            }
        }
        uint16_t *orow = out + ((((int64_t)b * odp + oz + out_pad) * ohp + oy + out_pad) * owp + out_pad + ox_l) * co;
#pragma unroll
        for (int nb = 0; nb < 4; ++nb) {
            const int ox = ox_l + 16 * nb;
#pragma unroll
            for (int mb = 0; mb < MB; ++mb) {
                floatx4 acc = {bs[mb][0], bs[mb][1], bs[mb][2], bs[mb][3]};  // the bias rides in the accumulator
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[mb][j], tp[nb][j], acc, 0, 0, 0);
                const int c0 = 16 * mb + 4 * g;
                if (ox < cd.ow && c0 < co) {
                    const uint32_t q0 = pack_bf16(relu_bits(acc[0]), relu_bits(acc[1]));
                    const uint32_t q1 = pack_bf16(relu_bits(acc[2]), relu_bits(acc[3]));
                    *reinterpret_cast<uint2 *>(orow + 16 * nb * co + c0) = make_uint2(q0, q1);
                }
            }
        }
    }
}

// The first two layers of the trunk in one pass: the uint8 image (width 128)
// -> 8 channels (stride 2) -> 16 channels (stride 2), so the 8-channel
// activations (twice the bytes of the image) never reach HBM.  A block owns 8
// output rows x 32 output columns of the second layer and walks kPairZ of its
// planes.  Per plane it stages 5 input planes x 35 rows of the image in LDS
// (the next plane's rows are loaded into registers meanwhile), computes the
// two new first-layer planes (17 rows x 64 columns, one halo row) into a ring
// of three planes, and reads the second layer's taps from the ring.  The
// first-layer products and rounding are those of k_conv_first_lds (bias in
// the accumulator, ReLU, BF16), the second layer's those of k_conv_igemm at
// c = 8 (lane group g takes tap 4 j + g of step j, bias after the sum), so the
// output equals the two-kernel path bit for bit.
constexpr int kPairZ = 16;                         // second-layer planes per block
constexpr int kPairInRows = 35;                    // image rows per staged plane
constexpr int kPairInPitch = 144;                  // 15 lead bytes, column -1 (zero), 128 columns
constexpr int kPairL0Rows = 17, kPairL0Cols = 65;  // ring plane: column 0 is the zero column -1
constexpr int kPairInBytes = 5 * kPairInRows * kPairInPitch;
constexpr int kPairL0Bytes = 3 * kPairL0Rows * kPairL0Cols * 16;
constexpr int kPairWaves = 8;  // waves per block

template <int W>
__global__ __launch_bounds__(64 * W) void k_conv_first_pair(const uint8_t *__restrict__ x, uint32_t x_bytes, int d,
                                                         int h, int nzc, const uint16_t *__restrict__ w0,
                                                         const float *__restrict__ b0,
                                                         const uint16_t *__restrict__ w1,
                                                         const float *__restrict__ b1, uint16_t *__restrict__ out,
                                                         int out_pad)
{
    // + 16: the dword pair of the last column of the last row reads 4 bytes past the image rows
    __shared__ __align__(16) uint8_t sin[kPairInBytes + 16];
    __shared__ __align__(16) uint8_t ring[kPairL0Bytes];
    constexpr int kThreads = 64 * W, kLoads = (5 * kPairInRows * 8 + kThreads - 1) / kThreads;
    constexpr int PB = 16 / W;  // second-layer blocks of 16 positions per wave
    const int lane = threadIdx.x & 63, g = lane >> 4, l16 = lane & 15;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int od = d / 4, oh = h / 4;
    unsigned r = (unsigned)xcd_tile(blockIdx.x, gridDim.x);
    const int zc = (int)(r % (unsigned)nzc);
    r /= (unsigned)nzc;
    const int by = (int)(r % (unsigned)(oh / 8));
    const int b = (int)(r / (unsigned)(oh / 8));
    const int oy0 = 8 * by, oz0 = kPairZ * zc, oz1 = min(oz0 + kPairZ, od);
    const __amdgpu_buffer_rsrc_t xb = buffer(x, x_bytes);
    const uint32_t sample_off = (uint32_t)b * (uint32_t)d * (uint32_t)h * 128u;
    // the zero column -1 of the image rows and of the ring planes (never overwritten)
    for (int t = threadIdx.x; t < 5 * kPairInRows; t += kThreads) sin[t * kPairInPitch + 15] = 0;
    for (int t = threadIdx.x; t < 3 * kPairL0Rows; t += kThreads)
        *reinterpret_cast<uint4 *>(ring + t * kPairL0Cols * 16) = make_uint4(0u, 0u, 0u, 0u);
    // first layer: A[ch = l % 16][slot 16 j + 4 g + kx], channels 8..15 zero
    bf16x4 wa[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int row = 4 * j + g;
        short e[4] = {0, 0, 0, 0};
        if (l16 < 8 && row < 9) {
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) e[kx] = (short)w0[l16 * 27 + row * 3 + kx];
        }
        wa[j] = bf16x4{e[0], e[1], e[2], e[3]};
    }
    float bs0[4], bs1[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        bs0[v] = g < 2 ? b0[4 * g + v] : 0.f;
        bs1[v] = b1[4 * g + v];
    }
    // second layer: A[ch = l % 16][k = 32 j + 8 g .. + 7] = tap 4 j + g, channels 0..7
    u32x4 wb1[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        uint32_t e[4] = {0u, 0u, 0u, 0u};
        if (4 * j + g < 27) {
            const uint16_t *src = w1 + l16 * 216 + 32 * j + 8 * g;
#pragma unroll
            for (int v = 0; v < 4; ++v) e[v] = (uint32_t)src[2 * v] | ((uint32_t)src[2 * v + 1] << 16);
        }
        wb1[j] = u32x4{e[0], e[1], e[2], e[3]};
    }
    // this lane's first-layer tap rows for the three product steps (rows 9..11 hold zero)
    int trow[3];
    bool tap[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const int row = 4 * j + g;
        tap[j] = row < 9;
        const int dz = tap[j] ? row / 3 : 0, dy = tap[j] ? row - 3 * (row / 3) : 0;
        trow[j] = dz * kPairInRows + dy;  // image row of the tap, relative to (plane 2 pl, row 2 r)
    }
    u32x4 pre[kLoads];
    // image planes zin0 .. zin0 + np - 1, rows 4 oy0 - 3 .. + 34, into registers
    auto load_image = [&](int zin0, int np) {
#pragma unroll
        for (int i = 0; i < kLoads; ++i) {
            const int t = threadIdx.x + kThreads * i;
            const int pl = t / (8 * kPairInRows), rem = t - pl * 8 * kPairInRows;
            const int iz = zin0 + pl, iy = 4 * oy0 - 3 + (rem >> 3);
            const bool ok = pl < np && iz >= 0 && iz < d && iy >= 0 && iy < h;
            const uint32_t o = ok ? sample_off + (uint32_t)((iz * h + iy) * 128 + 16 * (rem & 7)) : kOob;
            pre[i] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(xb, o, 0, 0));
        }
    };
    auto store_image = [&](int np) {
#pragma unroll
        for (int i = 0; i < kLoads; ++i) {
            const int t = threadIdx.x + kThreads * i;
            const int pl = t / (8 * kPairInRows), rem = t - pl * 8 * kPairInRows;
            if (pl < np)
                *reinterpret_cast<u32x4 *>(sin + (pl * kPairInRows + (rem >> 3)) * kPairInPitch + 16 +
                                           16 * (rem & 7)) = pre[i];
        }
    };
    // first-layer planes z0a .. z0a + np - 1 (image planes 2 z0a - 1 .. staged) into the ring;
    // positions in the zero border (plane or row -1) are written as zeros
    auto first_layer = [&](int z0a, int np) {
        const int groups = np * kPairL0Rows * 4;
#pragma unroll 2
        for (int gi = wv; gi < groups; gi += W) {
            const int pl = gi / (4 * kPairL0Rows), rem = gi - pl * 4 * kPairL0Rows;
            const int rr = rem >> 2, cb = rem & 3;
            const int z0 = z0a + pl, y0 = 2 * oy0 - 1 + rr, x0 = 16 * cb + l16;
            const int rowb = (2 * pl * kPairInRows + 2 * rr) * kPairInPitch + 15 + 2 * x0;
            floatx4 acc = {bs0[0], bs0[1], bs0[2], bs0[3]};
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const int o = rowb + trow[j] * kPairInPitch;
                const uint32_t *dw = reinterpret_cast<const uint32_t *>(sin + (o & ~3));
                uint32_t v = __builtin_amdgcn_alignbyte(dw[1], dw[0], (uint32_t)(o & 3));
                v = tap[j] ? v : 0u;
                const float f0 = (float)(v & 0xffu), f1 = (float)((v >> 8) & 0xffu), f2 = (float)((v >> 16) & 0xffu);
                const uint32_t p01 = __builtin_amdgcn_perm(__float_as_uint(f1), __float_as_uint(f0), 0x07060302u);
                const uint32_t p2 = __float_as_uint(f2) >> 16;
                const bf16x4 tp = bf16x4{(short)(p01 & 0xffffu), (short)(p01 >> 16), (short)p2, (short)0};
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[j], tp, acc, 0, 0, 0);
            }
            if (g < 2) {
                const bool zero = z0 < 0 || y0 < 0;
                const uint32_t q0 = zero ? 0u : pack_bf16(relu_bits(acc[0]), relu_bits(acc[1]));
                const uint32_t q1 = zero ? 0u : pack_bf16(relu_bits(acc[2]), relu_bits(acc[3]));
                const int slot = (z0 + 3) % 3;
                *reinterpret_cast<uint2 *>(ring + ((slot * kPairL0Rows + rr) * kPairL0Cols + x0 + 1) * 16 + 8 * g) =
                    make_uint2(q0, q1);
            }
        }
    };
    const int odp = od + 2 * out_pad, ohp = oh + 2 * out_pad, owp = 32 + 2 * out_pad;
    // prologue: first-layer plane 2 oz0 - 1 (zeros for oz0 = 0)
    load_image(4 * oz0 - 3, 3);
    store_image(3);
    load_image(4 * oz0 - 1, 5);
    __syncthreads();
    first_layer(2 * oz0 - 1, 1);
    for (int oz = oz0; oz < oz1; ++oz) {
        __syncthreads();  // the staged image and the ring's oldest planes are free
        store_image(5);
        if (oz + 1 < oz1) load_image(4 * oz + 3, 5);
        __syncthreads();
        first_layer(2 * oz, 2);
        __syncthreads();
        // second layer: block q = PB wv + pb of 16 positions is row q / 2, columns 16 (q % 2) ..
        floatx4 acc[PB];
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) acc[pb] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const int t = 4 * j + g;
            const bool kin = t < 27;
            const int tt = kin ? t : 0;
            const int dz = (tt * 29) >> 8, rem = tt - 9 * dz;
            const int dy = (rem * 11) >> 5, dx = rem - 3 * dy;
            const int slot = (2 * oz - 1 + dz + 3) % 3;
#pragma unroll
            for (int pb = 0; pb < PB; ++pb) {
                const int q = PB * wv + pb, oyl = q >> 1, ox = 16 * (q & 1) + l16;
                u32x4 v = *reinterpret_cast<const u32x4 *>(
                    ring + ((slot * kPairL0Rows + 2 * oyl + dy) * kPairL0Cols + 2 * ox + dx) * 16);
                if (!kin) v = u32x4{0u, 0u, 0u, 0u};
                acc[pb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, wb1[j]),
                                                                  __builtin_bit_cast(bf16x8, v), acc[pb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) {
            const int q = PB * wv + pb, oy = oy0 + (q >> 1), ox = 16 * (q & 1) + l16;
            float y[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) y[v] = relu_bits(acc[pb][v] + bs1[v]);
            const int64_t pos = (((int64_t)b * odp + oz + out_pad) * ohp + oy + out_pad) * owp + ox + out_pad;
            *reinterpret_cast<uint2 *>(out + pos * 16 + 4 * g) = make_uint2(pack_bf16(y[0], y[1]), pack_bf16(y[2], y[3]));
        }
    }
}

// Implicit-GEMM convolution for c % 8 == 0 on the matrix cores, computed as
// out^T (channels x positions) = W (channels x K) . P^T (K x positions) with
// k = tap * c + ci.  Wave = 16 P output positions x 16 CB output channels;
// no LDS, no barrier: in each 32-deep step lane l gathers the eight channels
// k = 32 j + 8 (l / 16) .. + 7 of its position (one 16-byte load: with
// c % 8 == 0 a run never crosses a tap) and the same k of its weight row,
// for one 16x16x32 product per (position block, channel block); after the
// loop it holds four consecutive channels of one position, stored as one
// 8-byte write.  Each lane follows the (tap, channel) of its own run: the
// four lane groups of a step can sit in up to four taps (c = 8).  PAD: the
// input is padded, every tap in range, no masks.
template <int P, int CB, bool PAD>
__global__ __launch_bounds__(256) void k_conv_igemm(const uint16_t *__restrict__ x, uint32_t x_bytes, ConvDims cd,
                                                    int co, const uint16_t *__restrict__ w,
                                                    const float *__restrict__ bias, int relu,
                                                    uint16_t *__restrict__ out, int out_pad)
{
    const int lane = threadIdx.x & 63, g = lane >> 4;
    const int M = cd.n * cd.od * cd.oh * cd.ow;  // < 2^31 (checked on the host)
    const int K = 27 * cd.c;
    const int p0 = ((int)xcd_tile(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6)) * 16 * P;
    if (p0 >= M) return;
    const int n0 = blockIdx.y * 16 * CB;
    const __amdgpu_buffer_rsrc_t xb = buffer(x, x_bytes);
    const __amdgpu_buffer_rsrc_t wb = buffer(w, (uint32_t)co * (uint32_t)K * 2u);
    const int dp = cd.d + (PAD ? 2 : 0), hp = cd.h + (PAD ? 2 : 0), wp = cd.w + (PAD ? 2 : 0);
    const int sz = hp * wp * cd.c, sy = wp * cd.c;  // element strides of z and y
    // per position block: the byte offset of tap (0,0,0), channel 0, and
    // (unpadded input) the in-range taps per axis
    uint32_t base[P];
    uint32_t msk[P];  // bits 0-2: z taps, 3-5: y taps, 6-8: x taps
#pragma unroll
    for (int pb = 0; pb < P; ++pb) {
        const int m = p0 + 16 * pb + (lane & 15);
        int rq = m < M ? m : 0;
        const int ox = rq % cd.ow; rq /= cd.ow;
        const int oy = rq % cd.oh; rq /= cd.oh;
        const int oz = rq % cd.od;
        const int b = rq / cd.od;
        const int iz = oz * cd.s - 1, iy = oy * cd.s - 1, ix = ox * cd.s - 1;
        uint32_t mk = 0u;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
            mk |= (uint32_t)(iz + t >= 0 && iz + t < cd.d) << t;
            mk |= (uint32_t)(iy + t >= 0 && iy + t < cd.h) << (3 + t);
            mk |= (uint32_t)(ix + t >= 0 && ix + t < cd.w) << (6 + t);
        }
        msk[pb] = PAD ? 0x1ffu : mk;
        const int o = PAD ? (((b * dp + iz + 1) * hp + iy + 1) * wp + ix + 1) * cd.c
                          : (((b * cd.d + iz) * cd.h + iy) * cd.w + ix) * cd.c;
        base[pb] = m < M ? (uint32_t)(2 * o) : kOob;
    }
    uint32_t wbase[CB];
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        const int n = n0 + 16 * cb + (lane & 15);
        wbase[cb] = n < co ? (uint32_t)(2 * (n * K + 8 * g)) : kOob;
    }
    floatx4 acc[P][CB];
#pragma unroll
    for (int pb = 0; pb < P; ++pb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) acc[pb][cb] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int steps = (K + 31) / 32;
    // software pipeline over S operand slots: slot q holds step j + q; its
    // products are issued, then the slot is reloaded with step j + S + q, so
    // the loads of the other S - 1 slots stay in flight under the products
    constexpr int S = 3;
    u32x4 rp[S][P], rw[S][CB];
    auto tap_off = [&](int tap) {  // element offset of tap (dz, dy, dx); tap < 27
        const int dz = (tap * 29) >> 8, rem = tap - 9 * dz;
        const int dy = (rem * 11) >> 5, dx = rem - 3 * dy;
        return dz * sz + dy * sy + dx * cd.c;
    };
    auto tap_bits = [&](int tap) {
        const int dz = (tap * 29) >> 8, rem = tap - 9 * dz;
        const int dy = (rem * 11) >> 5, dx = rem - 3 * dy;
        return (1u << dz) | (8u << dy) | (64u << dx);
    };
    // (tap, first channel) of this lane's run in the next step to issue
    int t_l = 0, ci_l = 8 * g;
    while (ci_l >= cd.c) {
        ci_l -= cd.c;
        ++t_l;
    }
    auto issue = [&](int q, int j) {  // step j into slot q
        const int ku = 32 * j;          // wave-uniform
        const bool kin = t_l < 27;      // k = ku + 8 g < K
        const int tap = kin ? t_l : 0;
        const uint32_t off = (uint32_t)(2 * (tap_off(tap) + ci_l)), need = tap_bits(tap);
#pragma unroll
        for (int pb = 0; pb < P; ++pb) {
            // without a border the tap (0,0,0) offset of an edge position is negative:
            // only the sum of base and off is a byte offset
            const bool ok = kin && (PAD || (msk[pb] & need) == need);
            rp[q][pb] = __builtin_bit_cast(
                u32x4, __builtin_amdgcn_raw_buffer_load_b128(xb, ok ? base[pb] + off : kOob, 0, 0));
        }
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
            rw[q][cb] = __builtin_bit_cast(
                u32x4, __builtin_amdgcn_raw_buffer_load_b128(wb, kin ? wbase[cb] : kOob, 2 * ku, 0));
        ci_l += 32;
        while (ci_l >= cd.c) {
            ci_l -= cd.c;
            ++t_l;
        }
    };
#pragma unroll
    for (int q = 0; q < S; ++q) issue(q, q);
    for (int j = 0; j < steps; j += S) {
#pragma unroll
        for (int q = 0; q < S; ++q) {
#pragma unroll
            for (int pb = 0; pb < P; ++pb)
#pragma unroll
                for (int cb = 0; cb < CB; ++cb)
                    acc[pb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, rw[q][cb]),
                                                                          __builtin_bit_cast(bf16x8, rp[q][pb]),
                                                                          acc[pb][cb], 0, 0, 0);
            issue(q, j + S + q);
        }
    }
    const int odp = cd.od + 2 * out_pad, ohp = cd.oh + 2 * out_pad, owp = cd.ow + 2 * out_pad;
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        const int c0 = n0 + 16 * cb + 4 * g;
        if (c0 >= co) continue;
        float bb[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) bb[v] = c0 + v < co ? bias[c0 + v] : 0.f;
#pragma unroll
        for (int pb = 0; pb < P; ++pb) {
            const int m = p0 + 16 * pb + (lane & 15);
            if (m >= M) continue;
            int rq = m;
            const int ox = rq % cd.ow; rq /= cd.ow;
            const int oy = rq % cd.oh; rq /= cd.oh;
            const int oz = rq % cd.od;
            const int b = rq / cd.od;
            const int64_t pos = (((int64_t)b * odp + oz + out_pad) * ohp + oy + out_pad) * owp + ox + out_pad;
            float y[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                y[v] = acc[pb][cb][v] + bb[v];
                if (relu) y[v] = relu_bits(y[v]);
            }
            uint16_t *o = out + pos * co + c0;
            if ((co & 3) == 0) {
                *reinterpret_cast<uint2 *>(o) = make_uint2(pack_bf16(y[0], y[1]), pack_bf16(y[2], y[3]));
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (c0 + v < co) o[v] = f2bf(y[v]);
            }
        }
    }
}

unsigned grid1(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

int conv_dims(const int32_t *in_shape, int stride, int64_t ldk, ConvDims *cd)
{
    for (int i = 0; i < 5; ++i)
        if (in_shape[i] < 1) return fail(PMX_ERR_INVALID, "conv input shape must be positive");
    if (stride != 1 && stride != 2) return fail(PMX_ERR_INVALID, "stride must be 1 or 2");
    if (ldk < 27 * (int64_t)in_shape[4]) return fail(PMX_ERR_INVALID, "patch row length below 27 * channels");
    cd->n = in_shape[0];
    cd->d = in_shape[1];
    cd->h = in_shape[2];
    cd->w = in_shape[3];
    cd->c = in_shape[4];
    cd->s = stride;
    cd->od = (cd->d - 1) / stride + 1;
    cd->oh = (cd->h - 1) / stride + 1;
    cd->ow = (cd->w - 1) / stride + 1;
    cd->ldk = ldk;
    return PMX_OK;
}

}  // namespace
}  // namespace pmx

extern "C" {

int pmx_gemm_bf16(int64_t M, int N, int K, const uint16_t *A, int64_t lda, int a_trans, const uint16_t *B, int64_t ldb,
                  int b_trans, float *C, int64_t ldc, int accumulate, int k_splits, hipStream_t s)
{
    pmx::clear_error();
    if (M < 0 || N < 0 || K < 0) return pmx::fail(PMX_ERR_INVALID, "negative GEMM size");
    if (!A || !B || !C) return pmx::fail(PMX_ERR_INVALID, "NULL operand");
    if ((a_trans ? lda < M : lda < K) || (b_trans ? ldb < K : ldb < N) || ldc < N)
        return pmx::fail(PMX_ERR_INVALID, "leading dimension too small");
    if (M == 0 || N == 0) return PMX_OK;
    const int64_t mt = (M + pmx::kTM - 1) / pmx::kTM;
    if (mt > INT32_MAX || (N + pmx::kTN - 1) / pmx::kTN > 65535) return pmx::fail(PMX_ERR_INVALID, "GEMM too large");
    if (k_splits < 1 || k_splits > 65535) return pmx::fail(PMX_ERR_INVALID, "k_splits must be in [1, 65535]");
    // K chunks are whole 32-wide steps
    const int k_chunk = (int)(((int64_t)K + k_splits - 1) / k_splits + pmx::kTK - 1) / pmx::kTK * pmx::kTK;
    const int splits = k_splits > 1 ? (K + k_chunk - 1) / k_chunk : 1;
    if (splits > 1 && !accumulate) return pmx::fail(PMX_ERR_INVALID, "split K adds into C: pass accumulate=1");
    const dim3 grid((unsigned)mt, (unsigned)((N + pmx::kTN - 1) / pmx::kTN), (unsigned)splits);
    hipLaunchKernelGGL(pmx::k_gemm, grid, dim3(256), 0, s, M, N, K, A, lda, a_trans, B, ldb, b_trans, C, ldc,
                       accumulate, splits > 1 ? k_chunk : K);
    return pmx::check_launch("gemm");
}

int pmx_im2col3d(const uint16_t *x, const int32_t *in_shape, int stride, int64_t ldk, uint16_t *cols, hipStream_t s)
{
    pmx::clear_error();
    pmx::ConvDims cd;
    if (int e = pmx::conv_dims(in_shape, stride, ldk, &cd)) return e;
    if (!x || !cols) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    const int64_t total = (int64_t)cd.n * cd.od * cd.oh * cd.ow * cd.ldk;
    hipLaunchKernelGGL(pmx::k_im2col, dim3(pmx::grid1(total)), dim3(256), 0, s, x, cd, cols);
    return pmx::check_launch("im2col");
}

int pmx_col2im3d(const float *dcols, const int32_t *in_shape, int stride, int64_t ldk, float *dx, hipStream_t s)
{
    pmx::clear_error();
    pmx::ConvDims cd;
    if (int e = pmx::conv_dims(in_shape, stride, ldk, &cd)) return e;
    if (!dcols || !dx) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    const int64_t total = (int64_t)cd.n * cd.d * cd.h * cd.w * cd.c;
    hipLaunchKernelGGL(pmx::k_col2im, dim3(pmx::grid1(total)), dim3(256), 0, s, dcols, cd, dx);
    return pmx::check_launch("col2im");
}

int pmx_bias_act(const float *y, const float *bias, int64_t m, int c, int relu, uint16_t *out_bf16, float *out_f32,
                 hipStream_t s)
{
    pmx::clear_error();
    if (!y || !bias || (!out_bf16 && !out_f32)) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    hipLaunchKernelGGL(pmx::k_bias_act, dim3(pmx::grid1(m * c)), dim3(256), 0, s, y, bias, m, c, relu, out_bf16,
                       out_f32);
    return pmx::check_launch("bias_act");
}

int pmx_relu_backward(const float *dy, const uint16_t *act, int64_t n, float *dx, uint16_t *dx_bf16, hipStream_t s)
{
    pmx::clear_error();
    if (!dy || (!dx && !dx_bf16)) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    hipLaunchKernelGGL(pmx::k_relu_back, dim3(pmx::grid1(n)), dim3(256), 0, s, dy, act, n, dx, dx_bf16);
    return pmx::check_launch("relu_backward");
}

int pmx_to_bf16(const float *a, int64_t n, uint16_t *out, hipStream_t s)
{
    pmx::clear_error();
    if (!a || !out) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    hipLaunchKernelGGL(pmx::k_to_bf16, dim3(pmx::grid1(n)), dim3(256), 0, s, a, n, out);
    return pmx::check_launch("to_bf16");
}

int pmx_u8_to_bf16(const uint8_t *a, int64_t n, uint16_t *out, hipStream_t s)
{
    pmx::clear_error();
    if (!a || !out) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    hipLaunchKernelGGL(pmx::k_u8_to_bf16, dim3(pmx::grid1(n)), dim3(256), 0, s, a, n, out);
    return pmx::check_launch("u8_to_bf16");
}

int pmx_col_sum(const float *a, int64_t m, int n, float *out, hipStream_t s)
{
    pmx::clear_error();
    if (!a || !out || n < 1 || n > 8192) return pmx::fail(PMX_ERR_INVALID, "bad column sum");
    if (m < 1) return PMX_OK;
    const int64_t blocks = std::min<int64_t>(1024, (m * n + 65535) / 65536);
    const int64_t rows = (m + blocks - 1) / blocks;
    hipLaunchKernelGGL(pmx::k_col_sum, dim3((unsigned)((m + rows - 1) / rows)), dim3(256), sizeof(float) * n, s, a, m, n,
                       rows, out);
    return pmx::check_launch("col_sum");
}

int pmx_avgpool(const uint16_t *x, int n, int64_t p, int c, float *out_f32, uint16_t *out_bf16, hipStream_t s)
{
    pmx::clear_error();
    if (!x || n < 1 || p < 1 || c < 1 || (!out_f32 && !out_bf16)) return pmx::fail(PMX_ERR_INVALID, "bad pooling");
    hipLaunchKernelGGL(pmx::k_avgpool, dim3((unsigned)(n * c)), dim3(256), 0, s, x, n, p, c, out_f32, out_bf16);
    return pmx::check_launch("avgpool");
}

int pmx_avgpool_backward(const float *dy, int n, int64_t p, int c, float *dx, hipStream_t s)
{
    pmx::clear_error();
    if (!dy || !dx) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    hipLaunchKernelGGL(pmx::k_avgpool_back, dim3(pmx::grid1((int64_t)n * p * c)), dim3(256), 0, s, dy, n, p, c, dx);
    return pmx::check_launch("avgpool_backward");
}

int pmx_gnll_loss(const float *pred, const float *target, int n, int q, float *loss, float *dpred, hipStream_t s)
{
    pmx::clear_error();
    if (!pred || !target || !loss || !dpred || n < 1 || q < 1) return pmx::fail(PMX_ERR_INVALID, "bad loss call");
    hipLaunchKernelGGL(pmx::k_gnll, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, pred, target, n, q, loss, dpred);
    return pmx::check_launch("gnll");
}

int pmx_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps, int step,
             uint16_t *w_bf16, hipStream_t s)
{
    pmx::clear_error();
    if (!w || !g || !m || !v || step < 1) return pmx::fail(PMX_ERR_INVALID, "bad Adam call");
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(pmx::k_adam, dim3(pmx::grid1(n)), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, c1, c2,
                       w_bf16);
    return pmx::check_launch("adam");
}

int pmx_conv3d_first(const uint8_t *x, const int32_t *in_shape, int stride, int co, const uint16_t *w_bf16,
                     const float *bias, uint16_t *out, int out_pad, hipStream_t s)
{
    pmx::clear_error();
    pmx::ConvDims cd;
    if (int e = pmx::conv_dims(in_shape, stride, 27 * (int64_t)in_shape[4], &cd)) return e;
    if (cd.c != 1) return pmx::fail(PMX_ERR_INVALID, "the first-layer kernel takes one input channel");
    if (co < 4 || co > 32 || co % 4) return pmx::fail(PMX_ERR_INVALID, "first-layer co must be a multiple of 4 in [4, 32]");
    if (out_pad != 0 && out_pad != 1) return pmx::fail(PMX_ERR_INVALID, "out_pad must be 0 or 1");
    if (!x || !w_bf16 || !bias || !out) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    if ((uintptr_t)out % 8) return pmx::fail(PMX_ERR_INVALID, "out must be 8-byte aligned");
    const int64_t sample = (int64_t)cd.d * cd.h * cd.w;
    if (sample >= pmx::kOob) return pmx::fail(PMX_ERR_INVALID, "one sample exceeds 2 GiB");
    const int64_t chunk = std::max<int64_t>(1, (int64_t)(pmx::kOob - 1) / sample);
    const bool fast = stride == 2 && cd.w % 4 == 0;
    const bool slab = stride == 2 && cd.w % 16 == 0;  // the LDS-staged form
    const int gy = (cd.oh + pmx::kFirstRows - 1) / pmx::kFirstRows, gx = (cd.ow + pmx::kFirstCols - 1) / pmx::kFirstCols;
    const int64_t out_sample = (int64_t)(cd.od + 2 * out_pad) * (cd.oh + 2 * out_pad) * (cd.ow + 2 * out_pad) * co;
    for (int64_t n0 = 0; n0 < cd.n; n0 += chunk) {
        pmx::ConvDims c = cd;
        c.n = (int)std::min<int64_t>(chunk, cd.n - n0);
        const int64_t blocks = (int64_t)c.n * c.od * gy * gx;
        if (blocks > INT32_MAX) return pmx::fail(PMX_ERR_INVALID, "input too large for one launch");
        const uint8_t *xc = x + n0 * sample;
        uint16_t *oc = out + n0 * out_sample;
        const uint32_t xbytes = (uint32_t)(c.n * sample);
        const dim3 grid((unsigned)blocks);
#define PMX_FIRST(MB, F) \
    hipLaunchKernelGGL((pmx::k_conv_first<MB, F>), grid, dim3(256), 0, s, xc, xbytes, c, gy, gx, co, w_bf16, bias, oc, out_pad)
#define PMX_FIRST_LDS(MB)                                                                                  \
    hipLaunchKernelGGL((pmx::k_conv_first_lds<MB>), grid, dim3(256), 0, s, xc, xbytes, c, gy, gx, co, w_bf16, bias, \
                       oc, out_pad)
        if (slab) {
            if (co <= 16) PMX_FIRST_LDS(1); else PMX_FIRST_LDS(2);
        } else if (co <= 16) {
            if (fast) PMX_FIRST(1, true); else PMX_FIRST(1, false);
        } else {
            if (fast) PMX_FIRST(2, true); else PMX_FIRST(2, false);
        }
#undef PMX_FIRST
#undef PMX_FIRST_LDS
        if (int e = pmx::check_launch("conv3d_first")) return e;
    }
    return PMX_OK;
}

int pmx_conv3d_first_pair(const uint8_t *x, const int32_t *in_shape, const uint16_t *w0_bf16, const float *b0,
                          const uint16_t *w1_bf16, const float *b1, uint16_t *out, int out_pad, hipStream_t s)
{
    pmx::clear_error();
    pmx::ConvDims cd;
    if (int e = pmx::conv_dims(in_shape, 2, 27, &cd)) return e;
    if (cd.c != 1) return pmx::fail(PMX_ERR_INVALID, "the fused first layers take one input channel");
    if (cd.w != 128 || cd.h % 32 || cd.d % 4)
        return pmx::fail(PMX_ERR_INVALID, "the fused first layers take width 128, height a multiple of 32 and "
                                          "depth a multiple of 4");
    if (out_pad != 0 && out_pad != 1) return pmx::fail(PMX_ERR_INVALID, "out_pad must be 0 or 1");
    if (!x || !w0_bf16 || !b0 || !w1_bf16 || !b1 || !out) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    if ((uintptr_t)out % 8) return pmx::fail(PMX_ERR_INVALID, "out must be 8-byte aligned");
    const int64_t sample = (int64_t)cd.d * cd.h * cd.w;
    if (sample >= pmx::kOob) return pmx::fail(PMX_ERR_INVALID, "one sample exceeds 2 GiB");
    const int64_t chunk = std::max<int64_t>(1, (int64_t)(pmx::kOob - 1) / sample);
    const int od = cd.d / 4, oh = cd.h / 4, nzc = (od + pmx::kPairZ - 1) / pmx::kPairZ;
    const int64_t out_sample = (int64_t)(od + 2 * out_pad) * (oh + 2 * out_pad) * (32 + 2 * out_pad) * 16;
    for (int64_t n0 = 0; n0 < cd.n; n0 += chunk) {
        const int64_t nc = std::min<int64_t>(chunk, cd.n - n0);
        const int64_t blocks = nc * (oh / 8) * nzc;
        if (blocks > INT32_MAX) return pmx::fail(PMX_ERR_INVALID, "input too large for one launch");
        hipLaunchKernelGGL(pmx::k_conv_first_pair<pmx::kPairWaves>, dim3((unsigned)blocks),
                           dim3(64 * pmx::kPairWaves), 0, s, x + n0 * sample,
                           (uint32_t)(nc * sample), cd.d, cd.h, nzc, w0_bf16, b0, w1_bf16, b1, out + n0 * out_sample,
                           out_pad);
        if (int e = pmx::check_launch("conv3d_first_pair")) return e;
    }
    return PMX_OK;
}

int pmx_conv3d_bf16(const uint16_t *x, const int32_t *in_shape, int in_pad, int stride, int co,
                    const uint16_t *w_bf16, const float *bias, int relu, uint16_t *out, int out_pad, hipStream_t s)
{
    pmx::clear_error();
    pmx::ConvDims cd;
    if (int e = pmx::conv_dims(in_shape, stride, 27 * (int64_t)in_shape[4], &cd)) return e;
    if (cd.c % 8) return pmx::fail(PMX_ERR_INVALID, "input channels must be a multiple of 8");
    if (co < 1) return pmx::fail(PMX_ERR_INVALID, "co must be positive");
    if ((in_pad != 0 && in_pad != 1) || (out_pad != 0 && out_pad != 1))
        return pmx::fail(PMX_ERR_INVALID, "in_pad and out_pad must be 0 or 1");
    if (!x || !w_bf16 || !bias || !out) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    if ((uintptr_t)x % 8 || (uintptr_t)w_bf16 % 8 || (uintptr_t)out % 8)
        return pmx::fail(PMX_ERR_INVALID, "x, w and out must be 8-byte aligned");
    const int64_t K = 27 * (int64_t)cd.c;
    if (2 * K * co >= pmx::kOob) return pmx::fail(PMX_ERR_INVALID, "weights exceed 2 GiB");
    const int64_t sample = 2 * (int64_t)(cd.d + 2 * in_pad) * (cd.h + 2 * in_pad) * (cd.w + 2 * in_pad) * cd.c;
    if (sample >= pmx::kOob) return pmx::fail(PMX_ERR_INVALID, "one sample exceeds 2 GiB");
    const int64_t chunk = std::max<int64_t>(1, (int64_t)(pmx::kOob - 1) / sample);
    const int64_t per = (int64_t)cd.od * cd.oh * cd.ow;
    const int64_t out_sample = (int64_t)(cd.od + 2 * out_pad) * (cd.oh + 2 * out_pad) * (cd.ow + 2 * out_pad) * co;
    // channel blocks per wave: the largest of 4, 3, 2, 1 that tiles ceil(co / 16) exactly, else 4
    const int c16 = (co + 15) / 16;
    const int cb = c16 % 4 == 0 ? 4 : c16 % 3 == 0 ? 3 : c16 % 2 == 0 ? 2 : c16 == 1 ? 1 : 4;
    const int64_t ny = (c16 + cb - 1) / cb;
    for (int64_t n0 = 0; n0 < cd.n; n0 += chunk) {
        pmx::ConvDims c = cd;
        c.n = (int)std::min<int64_t>(chunk, cd.n - n0);
        const int64_t M = c.n * per;
        if (M > INT32_MAX / 2) return pmx::fail(PMX_ERR_INVALID, "too many output positions");
        // positions per wave: 64, fewer when that leaves under 4096 waves
        int pos = 4;
        while (pos > 1 && (M + 16 * pos - 1) / (16 * pos) * ny < 4096) pos /= 2;
        const int64_t blocks = (M + 4 * 16 * pos - 1) / (4 * 16 * pos);
        const dim3 grid((unsigned)blocks, (unsigned)ny);
        const uint16_t *xc = x + n0 * (sample / 2);
        uint16_t *oc = out + n0 * out_sample;
        const uint32_t xbytes = (uint32_t)(c.n * sample);
#define PMX_IG(P, CB, PD)                                                                                          \
    hipLaunchKernelGGL((pmx::k_conv_igemm<P, CB, PD>), grid, dim3(256), 0, s, xc, xbytes, c, co, w_bf16, bias, relu, \
                       oc, out_pad)
#define PMX_IGP(CB, PD)       \
    do {                      \
        if (pos == 4)         \
            PMX_IG(4, CB, PD); \
        else if (pos == 2)    \
            PMX_IG(2, CB, PD); \
        else                  \
            PMX_IG(1, CB, PD); \
    } while (0)
#define PMX_IGC(PD)                    \
    switch (cb) {                      \
        case 1: PMX_IGP(1, PD); break; \
        case 2: PMX_IGP(2, PD); break; \
        case 3: PMX_IGP(3, PD); break; \
        default: PMX_IGP(4, PD);       \
    }
        if (in_pad) {
            PMX_IGC(true);
        } else {
            PMX_IGC(false);
        }
#undef PMX_IGC
#undef PMX_IGP
#undef PMX_IG
        if (int e = pmx::check_launch("conv3d_bf16")) return e;
    }
    return PMX_OK;
}

}  // extern "C"
