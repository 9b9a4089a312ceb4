// dme_learned.hip -- the learned proposal network: a 3x3 convolution layer
// on the block grid with 32 channels, BF16 operands on MFMA with FP32
// accumulation, and the kernels that train it (backward data, weight and
// bias gradients, ReLU backward, the winner-takes-all proposal loss, Adam).
// It has no counterpart in the reference encoder; the stage that turns its
// proposals into level-0 pel vectors is in dme_kernels.hip.
//
// Tensors are NHWC over the block grid with C = 32 channels: element
// (n, y, x, c) at ((n * H + y) * W + x) * 32 + c.  Weights of a layer are f32
// w[tap][ci][co] (tap = (dy + 1) * 3 + dx + 1) and bias[co]; the MFMA kernels
// read them as BF16 B-fragments made by k_lp_frags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "dme_device.h"

namespace dme {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

#define LP_C 32

// float -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t lp_f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float lp_bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// B-fragments of v_mfma_f32_16x16x32_bf16: lane l holds B[k = (l/16)*8 + j][n = l%16], j = 0..7.
// Forward:  frag[(tap*2 + nt)*64 + l][j] = w[tap][(l/16)*8 + j][nt*16 + l%16]
// Backward data is the same convolution over the output gradient with the
// taps flipped and the channel roles swapped:
//           frag[(tap*2 + nt)*64 + l][j] = w[8 - tap][nt*16 + l%16][(l/16)*8 + j]
__global__ __launch_bounds__(256) void k_lp_frags(const float *__restrict__ w, uint16_t *__restrict__ fwd,
                                                  uint16_t *__restrict__ bwd) {
    const int i = blockIdx.x * 256 + threadIdx.x;  // (tap, nt, lane)
    if (i >= 9 * 2 * 64) return;
    const int l = i & 63, nt = (i >> 6) & 1, tap = i >> 7;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int kk = (l >> 4) * 8 + j, nn = nt * 16 + (l & 15);
        fwd[i * 8 + j] = lp_f2bf(w[(tap * LP_C + kk) * LP_C + nn]);
        bwd[i * 8 + j] = lp_f2bf(w[((8 - tap) * LP_C + nn) * LP_C + kk]);
    }
}

// out[p][co] = bias[co] + sum_{tap, ci} in[p + off(tap)][ci] * w[tap][ci][co],
// zero outside the grid.  One wave makes 16 positions x 32 channels: a GEMM
// with K = 9 taps x 32 channels in nine 16x16x32 MFMA steps per 16-channel
// tile.  The A operand of lane l is the 8 channels (l/16)*8.. of position
// p0 + l%16 at the tap's neighbour: one 16-byte load.
__global__ __launch_bounds__(256) void k_lp_conv(const uint16_t *__restrict__ in, int npos, int H, int W,
                                                 const bf16x8 *__restrict__ frag, const float *__restrict__ bias,
                                                 int relu, uint16_t *__restrict__ out, float *__restrict__ out_f32) {
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), l = threadIdx.x & 63;
    const int p0 = wv * 16;
    if (p0 >= npos) return;
    const int p = p0 + (l & 15), kg = l >> 4;
    const bool pv = p < npos;
    const int hw = H * W;
    const int n = pv ? p / hw : 0, rem = pv ? p - n * hw : 0, y = rem / W, x = rem - y * W;
    floatx4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = acc0;
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
        bf16x8 a;
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = (__bf16)0.0f;
        if (pv && yy >= 0 && yy < H && xx >= 0 && xx < W)
            a = *reinterpret_cast<const bf16x8 *>(in + ((size_t)(n * H + yy) * W + xx) * LP_C + kg * 8);
        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, frag[(tap * 2 + 0) * 64 + l], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, frag[(tap * 2 + 1) * 64 + l], acc1, 0, 0, 0);
    }
    // D[m = (l/16)*4 + j][n = l%16]: position p0 + m, channels l%16 and 16 + l%16
    const int c0 = l & 15, c1 = 16 + (l & 15);
    const float b0 = bias ? bias[c0] : 0.0f, b1 = bias ? bias[c1] : 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = p0 + kg * 4 + j;
        if (q >= npos) continue;
        float v0 = acc0[j] + b0, v1 = acc1[j] + b1;
        if (relu) {
            v0 = fmaxf(v0, 0.0f);
            v1 = fmaxf(v1, 0.0f);
        }
        if (out) {
            out[(size_t)q * LP_C + c0] = lp_f2bf(v0);
            out[(size_t)q * LP_C + c1] = lp_f2bf(v1);
        }
        if (out_f32) {
            out_f32[(size_t)q * LP_C + c0] = v0;
            out_f32[(size_t)q * LP_C + c1] = v1;
        }
    }
}

// dw[tap][ci][co] += sum_p in[p + off(tap)][ci] * g[p][co].  One wave per
// (chunk of 256 positions, tap): per tap a GEMM with M = ci, N = co and
// K = positions, 16x16x32 MFMA tiles (2 x 2), added into dw atomically.
// Lane l of the A operand holds channel l%16 of 8 consecutive positions.
__global__ __launch_bounds__(256) void k_lp_wgrad(const uint16_t *__restrict__ in, const uint16_t *__restrict__ g,
                                                  int npos, int H, int W, float *__restrict__ dw) {
    const int wv = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), l = threadIdx.x & 63;
    const int chunk = wv / 9, tap = wv - chunk * 9;
    const int pc = chunk * 256;
    if (pc >= npos) return;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1, hw = H * W;
    floatx4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = (floatx4){0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < 8; ++ks) {
        bf16x8 A0, A1, B0, B1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int p = pc + ks * 32 + (l >> 4) * 8 + j;
            uint16_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
            if (p < npos) {
                const int n = p / hw, rem = p - n * hw, y = rem / W, x = rem - y * W;
                const int yy = y + dy, xx = x + dx;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                    const uint16_t *s = in + ((size_t)(n * H + yy) * W + xx) * LP_C;
                    a0 = s[l & 15];
                    a1 = s[16 + (l & 15)];
                }
                b0 = g[(size_t)p * LP_C + (l & 15)];
                b1 = g[(size_t)p * LP_C + 16 + (l & 15)];
            }
            A0[j] = __builtin_bit_cast(__bf16, a0);
            A1[j] = __builtin_bit_cast(__bf16, a1);
            B0[j] = __builtin_bit_cast(__bf16, b0);
            B1[j] = __builtin_bit_cast(__bf16, b1);
        }
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0, B1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1, B1, acc[1][1], 0, 0, 0);
    }
    // D[m = (l/16)*4 + j][n = l%16] of tile (mt, nt): ci = mt*16 + m, co = nt*16 + n
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int ci = mt * 16 + (l >> 4) * 4 + j, co = nt * 16 + (l & 15);
                atomicAdd(&dw[(tap * LP_C + ci) * LP_C + co], acc[mt][nt][j]);
            }
}

// db[co] += sum_p g[p][co]
__global__ __launch_bounds__(256) void k_lp_bgrad(const uint16_t *__restrict__ g, int npos, float *__restrict__ db) {
    const int c = threadIdx.x & 31, row = threadIdx.x >> 5;  // 8 rows of 32 channels
    float s = 0.0f;
    for (int p = blockIdx.x * 8 + row; p < npos; p += gridDim.x * 8) s += lp_bf2f(g[(size_t)p * LP_C + c]);
    __shared__ float red[8][32];
    red[row][c] = s;
    __syncthreads();
    if (row == 0) {
        float t = 0.0f;
        for (int r = 0; r < 8; ++r) t += red[r][c];
        atomicAdd(&db[c], t);
    }
}

// out = act > 0 ? bf16(g) : 0 (act = the layer's ReLU output); without act,
// a plain conversion.
__global__ __launch_bounds__(256) void k_lp_zero(uint32_t *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) p[i] = 0u;
}

__global__ __launch_bounds__(256) void k_lp_relu_bwd(const uint16_t *__restrict__ act, const float *__restrict__ g,
                                                     size_t count, uint16_t *__restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256)
        out[i] = (!act || lp_bf2f(act[i]) > 0.0f) ? lp_f2bf(g[i]) : (uint16_t)0;
}

// Adam (Kingma & Ba) on f32 parameters; bc1 = 1 - b1^t, bc2 = 1 - b2^t.
__global__ __launch_bounds__(256) void k_lp_adam(float *__restrict__ p, const float *__restrict__ g,
                                                 float *__restrict__ m, float *__restrict__ v, int count, float lr,
                                                 float b1, float b2, float eps, float bc1, float bc2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// Proposal loss per position and reference r (channels of the last layer:
// 8r + 2h + {0,1} = offset of hypothesis h in units of LP_SCALE pel from the
// guide, 16 + r = predicted log(1 + SAD)); labels[p][4r + {0..3}] = target
// offset in pel (x, y), target log(1 + SAD), valid (0/1).
//   L = e_best + alpha * sum_h e_h + beta * q(c - t),  e_h = |S o_h - t|_1
// (winner takes all over the hypotheses); q is the pinball loss of the
// LP_TAU quantile, so c predicts a high quantile of the log SAD and the
// fallback search (taken above 1.2 x that) stays rare.  dout gets
// dL/dout * scale; the scaled loss is added to *loss.
#define LP_SCALE 8.0f
#define LP_TAU 0.8f
__global__ __launch_bounds__(256) void k_lp_wta(const float *__restrict__ out, const float *__restrict__ lab,
                                                int npos, float alpha, float beta, float scale,
                                                float *__restrict__ dout, float *__restrict__ loss) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    float L = 0.0f;
    if (p < npos) {
        const float *o = out + (size_t)p * LP_C;
        float *d = dout + (size_t)p * LP_C;
        for (int c = 0; c < LP_C; ++c) d[c] = 0.0f;
        for (int r = 0; r < 2; ++r) {
            const float *t = lab + (size_t)p * 8 + 4 * r;
            if (t[3] == 0.0f) continue;
            float e[4];
            int hb = 0;
            for (int h = 0; h < 4; ++h) {
                e[h] = fabsf(LP_SCALE * o[8 * r + 2 * h] - t[0]) + fabsf(LP_SCALE * o[8 * r + 2 * h + 1] - t[1]);
                if (e[h] < e[hb]) hb = h;
            }
            L += e[hb];
            for (int h = 0; h < 4; ++h) {
                L += alpha * e[h];
                const float wgt = (h == hb ? 1.0f : 0.0f) + alpha;
                const float sx = LP_SCALE * o[8 * r + 2 * h] - t[0], sy = LP_SCALE * o[8 * r + 2 * h + 1] - t[1];
                d[8 * r + 2 * h] = scale * wgt * LP_SCALE * (sx > 0.0f ? 1.0f : (sx < 0.0f ? -1.0f : 0.0f));
                d[8 * r + 2 * h + 1] = scale * wgt * LP_SCALE * (sy > 0.0f ? 1.0f : (sy < 0.0f ? -1.0f : 0.0f));
            }
            const float dc = o[16 + r] - t[2];
            L += beta * (dc > 0.0f ? (1.0f - LP_TAU) * dc : -LP_TAU * dc);
            d[16 + r] = scale * beta * (dc > 0.0f ? (1.0f - LP_TAU) : (dc < 0.0f ? -LP_TAU : 0.0f));
        }
    }
    // one atomic per wave
    for (int m = 32; m >= 1; m >>= 1) L += __shfl_xor(L, m, 64);
    if ((threadIdx.x & 63) == 0 && L != 0.0f) atomicAdd(loss, L * scale);
}

hipError_t launch_lp_frags(const float *w, uint16_t *fwd, uint16_t *bwd, hipStream_t st) {
    hipLaunchKernelGGL(k_lp_frags, dim3((9 * 2 * 64 + 255) / 256), dim3(256), 0, st, w, fwd, bwd);
    return hipGetLastError();
}

hipError_t launch_lp_conv(const uint16_t *in, int npos, int H, int W, const uint16_t *frag, const float *bias, int relu,
                          uint16_t *out, float *out_f32, hipStream_t st) {
    if (npos <= 0) return hipSuccess;
    const int waves = (npos + 15) / 16;
    hipLaunchKernelGGL(k_lp_conv, dim3((waves + 3) / 4), dim3(256), 0, st, in, npos, H, W,
                       reinterpret_cast<const bf16x8 *>(frag), bias, relu, out, out_f32);
    return hipGetLastError();
}

hipError_t launch_lp_wgrad(const uint16_t *in, const uint16_t *g, int npos, int H, int W, float *dw, float *db,
                           hipStream_t st) {
    if (npos <= 0) return hipSuccess;
    const int waves = ((npos + 255) / 256) * 9;
    hipLaunchKernelGGL(k_lp_wgrad, dim3((waves + 3) / 4), dim3(256), 0, st, in, g, npos, H, W, dw);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_lp_bgrad, dim3(std::min(1024, (npos + 7) / 8)), dim3(256), 0, st, g, npos, db);
    return hipGetLastError();
}

hipError_t launch_lp_zero(void *p, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 4;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lp_zero, dim3((int)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0, st,
                       reinterpret_cast<uint32_t *>(p), n);
    return hipGetLastError();
}

hipError_t launch_lp_relu_bwd(const uint16_t *act, const float *g, size_t count, uint16_t *out, hipStream_t st) {
    if (count == 0) return hipSuccess;
    const int blocks = (int)std::min<size_t>((count + 255) / 256, 8192);
    hipLaunchKernelGGL(k_lp_relu_bwd, dim3(blocks), dim3(256), 0, st, act, g, count, out);
    return hipGetLastError();
}

hipError_t launch_lp_adam(float *p, const float *g, float *m, float *v, int count, float lr, float b1, float b2,
                          float eps, float bc1, float bc2, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lp_adam, dim3((count + 255) / 256), dim3(256), 0, st, p, g, m, v, count, lr, b1, b2, eps,
                       bc1, bc2);
    return hipGetLastError();
}

hipError_t launch_lp_wta(const float *out, const float *lab, int npos, float alpha, float beta, float scale,
                         float *dout, float *loss, hipStream_t st) {
    if (npos <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lp_wta, dim3((npos + 255) / 256), dim3(256), 0, st, out, lab, npos, alpha, beta, scale, dout,
                       loss);
    return hipGetLastError();
}

}  // namespace dme
