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
// inference convolutions (kernel 3, padding 1, stride s; bias + ReLU fused)
// ---------------------------------------------------------------------------

// First layer, one input channel: a thread per output voxel reads its 27
// input voxels (the u8 image) and writes co (<= 32) BF16 channels.

// First layer on the matrix cores: out^T (channels x positions) = W (channels
// x K) . P^T (K x positions) with K = 48 slots, slot 4r + kx for tap row r =
// 3kz + ky (slot 4r + 3 and rows 9..11 hold zero), so in each 16-deep MFMA
// step lane l (position l % 16, slots 4(l/16)..+3) needs the three bytes of
// one input row.  Wave = 64 output columns of one output row (four MFMA
// column blocks); block = four output rows.  FAST (stride 2, w % 4 == 0,
// 4-byte aligned image): a row's three bytes come from two aligned dwords.
constexpr int kFirstRows = 4, kFirstCols = 64;

__device__ __forceinline__ uint32_t bf16_pair(uint32_t lo, uint32_t hi)
{
    return (__float_as_uint((float)lo) >> 16) | (__float_as_uint((float)hi) & 0xffff0000u);
}

template <int MB, bool FAST>
__global__ __launch_bounds__(256) void k_conv_first(const uint8_t *__restrict__ x, ConvDims cd, int gy, int gx, int co,
                                                    const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                                    uint16_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    unsigned r = blockIdx.x;
    const int bx = (int)(r % (unsigned)gx);
    r /= (unsigned)gx;
    const int by = (int)(r % (unsigned)gy);
    r /= (unsigned)gy;
    const int oz = (int)(r % (unsigned)cd.od);
    const int b = (int)(r / (unsigned)cd.od);
    const int oy = by * kFirstRows + wv;
    if (oy >= cd.oh) return;
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
    const int64_t out_row = (((int64_t)b * cd.od + oz) * cd.oh + oy) * cd.ow;
#pragma unroll
    for (int nb = 0; nb < 4; ++nb) {
        const int ox = bx * kFirstCols + 16 * nb + (lane & 15);
        bf16x4 tp[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int row = 4 * j + g;
            const int kz = row / 3, ky = row - 3 * kz;
            const int iz = oz * cd.s - 1 + kz, iy = oy * cd.s - 1 + ky;
            const bool in = row < 9 && ox < cd.ow && iz >= 0 && iz < cd.d && iy >= 0 && iy < cd.h;
            const uint8_t *src = x + (((int64_t)b * cd.d + (in ? iz : 0)) * cd.h + (in ? iy : 0)) * cd.w;
            uint32_t b0 = 0u, b1 = 0u, b2 = 0u;
            if (FAST) {
                // bytes 2ox-1 .. 2ox+1 lie in the dwords at a and a + 4
                const int x0 = 2 * ox - 1;
                const int a = x0 & ~3;
                uint32_t lo = 0u, hi = 0u;
                if (in && a >= 0) lo = *reinterpret_cast<const uint32_t *>(src + a);
                if (in && a + 4 < cd.w) hi = *reinterpret_cast<const uint32_t *>(src + a + 4);
                const uint32_t v = __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)(x0 & 3));
                b0 = v & 0xffu;
                b1 = (v >> 8) & 0xffu;
                b2 = (v >> 16) & 0xffu;
            } else if (in) {
                const int x0 = ox * cd.s - 1;
                if (x0 >= 0) b0 = src[x0];
                if (x0 + 1 < cd.w) b1 = src[x0 + 1];
                if (x0 + 2 < cd.w) b2 = src[x0 + 2];
            }
            const uint32_t p01 = bf16_pair(b0, b1), p2 = __float_as_uint((float)b2) >> 16;
            tp[j] = bf16x4{(short)(p01 & 0xffffu), (short)(p01 >> 16), (short)p2, (short)0};
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int j = 0; j < 3; ++j) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wa[mb][j], tp[j], acc, 0, 0, 0);
            const int c0 = 16 * mb + 4 * g;
            if (ox < cd.ow && c0 < co) {
                const uint32_t q0 = f2bf(fmaxf(acc[0] + bs[mb][0], 0.f)) | ((uint32_t)f2bf(fmaxf(acc[1] + bs[mb][1], 0.f)) << 16);
                const uint32_t q1 = f2bf(fmaxf(acc[2] + bs[mb][2], 0.f)) | ((uint32_t)f2bf(fmaxf(acc[3] + bs[mb][3], 0.f)) << 16);
                *reinterpret_cast<uint2 *>(out + (out_row + ox) * co + c0) = make_uint2(q0, q1);
            }
        }
    }
}

// Implicit-GEMM convolution for c % 8 == 0: out[m][co] = ReLU(bias + sum_k
// patch(m, k) w[co][k]).  A workgroup owns 64 output positions x NT output
// channels; each K step of 32 gathers one 16-byte run of 8 channels per
// thread straight from the input (k = tap * c + ci, a run never crosses a
// tap), so no patch matrix is written.  Wave w computes rows 16w .. 16w+15.
template <int NT>
__global__ __launch_bounds__(256) void k_conv_igemm(const uint16_t *__restrict__ x, ConvDims cd, int co,
                                                    const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                                    int relu, uint16_t *__restrict__ out)
{
    __shared__ __align__(16) uint16_t As[64][kTK + 8];
    __shared__ __align__(16) uint16_t Bs[NT][kTK + 8];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t M = (int64_t)cd.n * cd.od * cd.oh * cd.ow;
    const int K = 27 * cd.c;
    const int64_t m0 = (int64_t)blockIdx.x * 64;
    const int n0 = blockIdx.y * NT;
    // this thread's staged row and 8-channel run within each K step
    const int ar = tid >> 2, ac = (tid & 3) * 8;
    const int64_t am = m0 + ar;
    int ab = 0, az = 0, ay = 0, ax = 0;
    {
        int64_t r = am < M ? am : 0;
        ax = (int)(r % cd.ow) * cd.s - 1; r /= cd.ow;
        ay = (int)(r % cd.oh) * cd.s - 1; r /= cd.oh;
        az = (int)(r % cd.od) * cd.s - 1;
        ab = (int)(r / cd.od);
    }
    floatx4 acc[NT / 16];
#pragma unroll
    for (int j = 0; j < NT / 16; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += kTK) {
        {
            const int k = k0 + ac;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (am < M && k < K) {
                const int tap = k / cd.c, ci = k - tap * cd.c;
                const int iz = az + tap / 9, iy = ay + (tap / 3) % 3, ix = ax + tap % 3;
                if (iz >= 0 && iz < cd.d && iy >= 0 && iy < cd.h && ix >= 0 && ix < cd.w)
                    v = *reinterpret_cast<const uint4 *>(
                        x + ((((int64_t)ab * cd.d + iz) * cd.h + iy) * cd.w + ix) * cd.c + ci);
            }
            *reinterpret_cast<uint4 *>(&As[ar][ac]) = v;
        }
        for (int e = tid; e < NT * (kTK / 8); e += 256) {
            const int c = e / (kTK / 8), kc = (e % (kTK / 8)) * 8;
            const int gn = n0 + c, gk = k0 + kc;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (gn < co && gk < K) v = *reinterpret_cast<const uint4 *>(w + (int64_t)gn * K + gk);
            *reinterpret_cast<uint4 *>(&Bs[c][kc]) = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 16) {
            const int kq = kk + 4 * (lane >> 4);
            const uint16_t *pa = &As[16 * wv + (lane & 15)][kq];
            const bf16x4 a = bf16x4{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
#pragma unroll
            for (int j = 0; j < NT / 16; ++j) {
                const uint16_t *pb = &Bs[16 * j + (lane & 15)][kq];
                const bf16x4 b = bf16x4{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NT / 16; ++j) {
        const int col = n0 + 16 * j + (lane & 15);
        if (col >= co) continue;
        const float bb = bias[col];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t row = m0 + 16 * wv + 4 * (lane >> 4) + v;
            if (row < M) {
                float y = acc[j][v] + bb;
                if (relu) y = fmaxf(y, 0.f);
                out[row * co + col] = f2bf(y);
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
                     const float *bias, uint16_t *out, hipStream_t s)
{
    pmx::clear_error();
    pmx::ConvDims cd;
    if (int e = pmx::conv_dims(in_shape, stride, 27 * (int64_t)in_shape[4], &cd)) return e;
    if (cd.c != 1) return pmx::fail(PMX_ERR_INVALID, "the first-layer kernel takes one input channel");
    if (co < 4 || co > 32 || co % 4) return pmx::fail(PMX_ERR_INVALID, "first-layer co must be a multiple of 4 in [4, 32]");
    if (!x || !w_bf16 || !bias || !out) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    if ((uintptr_t)out % 8) return pmx::fail(PMX_ERR_INVALID, "out must be 8-byte aligned");
    const int gy = (cd.oh + pmx::kFirstRows - 1) / pmx::kFirstRows, gx = (cd.ow + pmx::kFirstCols - 1) / pmx::kFirstCols;
    const int64_t blocks = (int64_t)cd.n * cd.od * gy * gx;
    if (blocks > UINT32_MAX / 2) return pmx::fail(PMX_ERR_INVALID, "input too large for one launch");
    const bool fast = stride == 2 && cd.w % 4 == 0 && (uintptr_t)x % 4 == 0;
    const dim3 grid((unsigned)blocks);
#define PMX_FIRST(MB)                                                                                              \
    do {                                                                                                           \
        if (fast)                                                                                                  \
            hipLaunchKernelGGL((pmx::k_conv_first<MB, true>), grid, dim3(256), 0, s, x, cd, gy, gx, co, w_bf16, bias, \
                               out);                                                                               \
        else                                                                                                       \
            hipLaunchKernelGGL((pmx::k_conv_first<MB, false>), grid, dim3(256), 0, s, x, cd, gy, gx, co, w_bf16,    \
                               bias, out);                                                                         \
    } while (0)
    if (co <= 16)
        PMX_FIRST(1);
    else
        PMX_FIRST(2);
#undef PMX_FIRST
    return pmx::check_launch("conv3d_first");
}

int pmx_conv3d_bf16(const uint16_t *x, const int32_t *in_shape, int stride, int co, const uint16_t *w_bf16,
                    const float *bias, int relu, uint16_t *out, hipStream_t s)
{
    pmx::clear_error();
    pmx::ConvDims cd;
    if (int e = pmx::conv_dims(in_shape, stride, 27 * (int64_t)in_shape[4], &cd)) return e;
    if (cd.c % 8) return pmx::fail(PMX_ERR_INVALID, "input channels must be a multiple of 8");
    if (co < 1) return pmx::fail(PMX_ERR_INVALID, "co must be positive");
    if (!x || !w_bf16 || !bias || !out) return pmx::fail(PMX_ERR_INVALID, "NULL buffer");
    if ((uintptr_t)x % 16 || (uintptr_t)w_bf16 % 16) return pmx::fail(PMX_ERR_INVALID, "x and w must be 16-byte aligned");
    const int64_t M = (int64_t)cd.n * cd.od * cd.oh * cd.ow;
    const int64_t mt = (M + 63) / 64;
    if (mt > INT32_MAX) return pmx::fail(PMX_ERR_INVALID, "too many output positions");
    if (co <= 16)
        hipLaunchKernelGGL(pmx::k_conv_igemm<16>, dim3((unsigned)mt, 1), dim3(256), 0, s, x, cd, co, w_bf16, bias, relu,
                           out);
    else if (co <= 32)
        hipLaunchKernelGGL(pmx::k_conv_igemm<32>, dim3((unsigned)mt, 1), dim3(256), 0, s, x, cd, co, w_bf16, bias, relu,
                           out);
    else if (co <= 48)
        hipLaunchKernelGGL(pmx::k_conv_igemm<48>, dim3((unsigned)mt, 1), dim3(256), 0, s, x, cd, co, w_bf16, bias, relu,
                           out);
    else
        hipLaunchKernelGGL(pmx::k_conv_igemm<64>, dim3((unsigned)mt, (unsigned)((co + 63) / 64)), dim3(256), 0, s, x, cd,
                           co, w_bf16, bias, relu, out);
    return pmx::check_launch("conv3d_bf16");
}

}  // extern "C"
