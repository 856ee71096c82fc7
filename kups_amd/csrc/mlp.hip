// Learned GCMC insertion proposal: voxel descriptors, MLP inference (BF16
// MFMA), MLP training (FP32 MFMA forward/backward with BF16-rounded weights,
// Adam) and the per-chain proposal table.  No counterpart in the reference: the
// proposal enters kUPS's MCMC through the ChangesFn log ratio
// (core/propagator.py:230-235, added at 332-334).
//
// Network: 32 -> 128 -> 128 -> 128 -> 1, SiLU.  Flat parameter vector
// (FP32 master, BF16 copy for the MFMA kernel):
//   W1[32][128] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] w4[128] b4
#include "common.hpp"

namespace kupsmi {

constexpr int ML_IN = 32, ML_H = 128;
constexpr int ML_OFF_W1 = 0;
constexpr int ML_OFF_B1 = ML_OFF_W1 + ML_IN * ML_H;
constexpr int ML_OFF_W2 = ML_OFF_B1 + ML_H;
constexpr int ML_OFF_B2 = ML_OFF_W2 + ML_H * ML_H;
constexpr int ML_OFF_W3 = ML_OFF_B2 + ML_H;
constexpr int ML_OFF_B3 = ML_OFF_W3 + ML_H * ML_H;
constexpr int ML_OFF_W4 = ML_OFF_B3 + ML_H;
constexpr int ML_OFF_B4 = ML_OFF_W4 + ML_H;
constexpr int ML_NP = ML_OFF_B4 + 1;  // 37377

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short f2bf(float f) {
  // round to nearest even (finite inputs)
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf2f(unsigned short h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }

// ------------------------------------------------------- descriptors ----
// 32-bin radial histogram of host atoms around every voxel centre of every
// chain (min image, r < rmax), normalised by the shell volume times a
// reference density: d_b = n_b / (4/3 pi (r_{b+1}^3 - r_b^3) rho0).
__global__ __launch_bounds__(256) void ml_desc_kernel(int n_chain, int nv, const int32_t* host_off,
                                                      const float4* host, const double* tril,
                                                      float rmax, float rho0, unsigned short* desc) {
  extern __shared__ float4 hs[];
  const int ch = blockIdx.y;
  const int h0 = host_off[ch], nh = host_off[ch + 1] - h0;
  for (int j = threadIdx.x; j < nh; j += blockDim.x) hs[j] = host[h0 + j];
  __syncthreads();
  const int nvox = nv * nv * nv;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nvox) return;
  const double* t = tril + 6 * ch;
  double li[6];
  tril_inverse(t, li);
  float L[6], Li[6];
  for (int k = 0; k < 6; ++k) { L[k] = (float)t[k]; Li[k] = (float)li[k]; }
  const int i0 = v / (nv * nv), i1 = (v / nv) % nv, i2 = v % nv;
  float cx, cy, cz;
  // voxel (i0, i1, i2) covers fractional [i/nv, (i+1)/nv) of the insertion offset
  tmul(L, (i0 + 0.5f) / nv, (i1 + 0.5f) / nv, (i2 + 0.5f) / nv, cx, cy, cz);
  int cnt[ML_IN];
#pragma unroll
  for (int b = 0; b < ML_IN; ++b) cnt[b] = 0;
  const float rm2 = rmax * rmax, inv_dr = ML_IN / rmax;
  for (int j = 0; j < nh; ++j) {
    const float4 h = hs[j];
    float f0, f1, f2;
    tmul(Li, h.x - cx, h.y - cy, h.z - cz, f0, f1, f2);
    f0 -= rintf(f0); f1 -= rintf(f1); f2 -= rintf(f2);
    float d0, d1, d2;
    tmul(L, f0, f1, f2, d0, d1, d2);
    const float r2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (r2 < rm2) {
      int b = (int)(sqrtf(r2) * inv_dr);
      b = b > ML_IN - 1 ? ML_IN - 1 : b;
#pragma unroll
      for (int c = 0; c < ML_IN; ++c) cnt[c] += (c == b);
    }
  }
  unsigned short* o = desc + ((size_t)ch * nvox + v) * ML_IN;
  const float dr = rmax / ML_IN;
#pragma unroll
  for (int b = 0; b < ML_IN; ++b) {
    const float r0 = b * dr, r1 = (b + 1) * dr;
    const float shell = 4.18879020478639f * (r1 * r1 * r1 - r0 * r0 * r0) * rho0;
    o[b] = f2bf(cnt[b] / shell);
  }
}

// ---------------------------------------------------- MFMA inference ----
// One wave per 16-row tile: three 16 x 128 layers as gfx950's
// v_mfma_f32_16x16x32_bf16 (A: 16 x 32 activations, B: 32 x 16 weights, FP32
// accumulators; lane l holds row / column l % 16 and eight consecutive k from
// 8 (l / 16), the same k set on both sides).  The eight 16-column output
// tiles of a layer are accumulated together, so eight independent MFMA chains
// share each A fragment; bias and SiLU are applied on the accumulators
// (reciprocal by v_rcp_f32: the result is rounded to BF16), activations are
// re-staged in LDS as BF16.  Weights are held transposed ([out][in]) in LDS
// so that every B fragment is one 16-byte load; rows are padded by 8
// elements (16 bytes) so the 16 rows of a fragment fall in different banks.
constexpr int ML_WAVES = 8;
constexpr int ML_LD_IN = ML_IN + 8;  // padded row lengths (BF16 elements)
constexpr int ML_LD_H = ML_H + 8;

__device__ inline float silu_bf(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

template <int KIN, int LDK>
__device__ inline void ml_layer(const unsigned short* act, const unsigned short* wT,
                                const float* bias, unsigned short* out, int lane) {
  const int row = lane & 15, kq = (lane >> 4) * 8;
  f32x4 acc[ML_H / 16];
#pragma unroll
  for (int nt = 0; nt < ML_H / 16; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KIN / 32; ++ks) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(act + row * LDK + ks * 32 + kq);
#pragma unroll
    for (int nt = 0; nt < ML_H / 16; ++nt) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(wT + (nt * 16 + row) * LDK + ks * 32 + kq);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[nt], 0, 0, 0);
    }
  }
  // D[i][j]: i = (lane / 16) * 4 + r, j = lane % 16 (+ 16 nt)
#pragma unroll
  for (int nt = 0; nt < ML_H / 16; ++nt) {
    const int j = nt * 16 + (lane & 15);
    const float bj = bias[j];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = (lane >> 4) * 4 + r;
      out[i * ML_LD_H + j] = f2bf(silu_bf(acc[nt][r] + bj));
    }
  }
}

__global__ __launch_bounds__(64 * ML_WAVES) void ml_infer_kernel(int n_rows, const unsigned short* desc,
                                                                const unsigned short* wbf,
                                                                const float* wf, float* logits) {
  __shared__ __attribute__((aligned(16))) unsigned short w1t[ML_H * ML_LD_IN];
  __shared__ __attribute__((aligned(16))) unsigned short w2t[ML_H * ML_LD_H], w3t[ML_H * ML_LD_H];
  __shared__ float b1[ML_H], b2[ML_H], b3[ML_H], w4[ML_H];
  __shared__ __attribute__((aligned(16))) unsigned short bufA[ML_WAVES][16 * ML_LD_H];
  __shared__ __attribute__((aligned(16))) unsigned short bufB[ML_WAVES][16 * ML_LD_H];
  const int tid = threadIdx.x;
  for (int e = tid; e < ML_IN * ML_H; e += blockDim.x) {
    const int k = e / ML_H, n = e % ML_H;  // W1[k][n]
    w1t[n * ML_LD_IN + k] = wbf[ML_OFF_W1 + e];
  }
  for (int e = tid; e < ML_H * ML_H; e += blockDim.x) {
    const int k = e / ML_H, n = e % ML_H;
    w2t[n * ML_LD_H + k] = wbf[ML_OFF_W2 + e];
    w3t[n * ML_LD_H + k] = wbf[ML_OFF_W3 + e];
  }
  for (int e = tid; e < ML_H; e += blockDim.x) {
    b1[e] = wf[ML_OFF_B1 + e];
    b2[e] = wf[ML_OFF_B2 + e];
    b3[e] = wf[ML_OFF_B3 + e];
    w4[e] = bf2f(wbf[ML_OFF_W4 + e]);
  }
  __syncthreads();
  const float b4 = wf[ML_OFF_B4];
  const int wave = tid >> 6, lane = tid & 63;
  unsigned short* A = bufA[wave];
  unsigned short* B = bufB[wave];
  const int ntiles = (n_rows + 15) / 16;
  for (int tile = blockIdx.x * ML_WAVES + wave; tile < ntiles; tile += gridDim.x * ML_WAVES) {
    const int r0 = tile * 16;
    // input tile 16 x 32: one 16-byte load per lane (rows past the end are zero)
    {
      const int r = lane >> 2, k8 = (lane & 3) * 8;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (r0 + r < n_rows) v = *reinterpret_cast<const uint4*>(desc + (size_t)(r0 + r) * ML_IN + k8);
      *reinterpret_cast<uint4*>(A + r * ML_LD_IN + k8) = v;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    ml_layer<ML_IN, ML_LD_IN>(A, w1t, b1, B, lane);
    __builtin_amdgcn_wave_barrier();
    ml_layer<ML_H, ML_LD_H>(B, w2t, b2, A, lane);
    __builtin_amdgcn_wave_barrier();
    ml_layer<ML_H, ML_LD_H>(A, w3t, b3, B, lane);
    __builtin_amdgcn_wave_barrier();
    // output: 4 lanes per row, 32 products each, shuffle reduction
    const int row = lane >> 2, part = lane & 3;
    float s = 0.f;
    for (int j = part * 32; j < part * 32 + 32; ++j) s += bf2f(B[row * ML_LD_H + j]) * w4[j];
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    if (part == 0 && r0 + row < n_rows) logits[r0 + row] = s + b4;
    __builtin_amdgcn_wave_barrier();
  }
}

// ------------------------------------------------------------ training ----
// Forward and backward on the matrix cores (v_mfma_f32_16x16x4f32: FP32
// operands and accumulation) with the BF16-rounded weights; loss = 1/2 mean
// (f(x) - y)^2.  One 256-thread workgroup per slice of ML_SLICE samples, taken
// as 16-sample tiles.  Wave w owns hidden columns [32w, 32w + 32) of every
// layer: it computes those columns of Z1..Z3 and of the back-propagated
// errors, and the same columns of the weight gradients, which stay in
// accumulator registers over the slice and are written once as the
// workgroup's partial.  Weights sit in LDS as BF16 (the training copy is
// BF16-rounded, so this is exact) in [in][out] order; activations and errors
// as FP32.  A pass takes ML_RT 16-row tiles of samples, so every weight
// fragment read from LDS feeds ML_RT independent accumulators.  Per 16
// samples and wave: 16 + 64 + 64 forward MFMAs, 64 + 64 error and
// 64 + 64 + 16 weight-gradient MFMAs.
//
// 16x16x4 operands: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k; result
// D[i][j] in lane j + 16 (i / 4), element i % 4.
constexpr int ML_SLICE = 32;
constexpr int ML_RT = 2;         // 16-row sample tiles per pass
constexpr int ML_ROWS = 16 * ML_RT;
constexpr int ML_TW = 256;       // threads of a training workgroup (4 waves)
constexpr int ML_LDW = ML_H + 8;    // BF16 weight rows
constexpr int ML_LDA = ML_H + 20;   // FP32 activation / error rows
constexpr int ML_LDX = ML_IN + 20;  // FP32 input rows

__device__ inline f32x4 mfma4(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline float dsilu(float z) {
  const float sg = 1.0f / (1.0f + __expf(-z));
  return sg * (1.0f + z * (1.0f - sg));
}

// z[rt][t] = H (row tile rt, 16 x K) . W[:, columns of tile nt0 + t], W in
// [in][out] order
template <int K, int LDH>
__device__ inline void ml_fwd_cols(const float* H, const unsigned short* W, int nt0, int lane,
                                   f32x4 (&z)[ML_RT][2]) {
  const int i = lane & 15, q = lane >> 4;
#pragma unroll
  for (int rt = 0; rt < ML_RT; ++rt) z[rt][0] = z[rt][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
  for (int k0 = 0; k0 < K; k0 += 4) {
    float a[ML_RT];
#pragma unroll
    for (int rt = 0; rt < ML_RT; ++rt) a[rt] = H[(rt * 16 + i) * LDH + k0 + q];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const float b = bf2f(W[(k0 + q) * ML_LDW + (nt0 + t) * 16 + i]);
#pragma unroll
      for (int rt = 0; rt < ML_RT; ++rt) z[rt][t] = mfma4(a[rt], b, z[rt][t]);
    }
  }
}

// e[t][s][j] = sum_u D[s][u] W[j][u] for j in tile nt0 + t (the error of the
// layer below, before its SiLU derivative)
__device__ inline void ml_bwd_cols(const float* D, const unsigned short* W, int nt0, int lane,
                                   f32x4 (&e)[ML_RT][2]) {
  const int i = lane & 15, q = lane >> 4;
#pragma unroll
  for (int rt = 0; rt < ML_RT; ++rt) e[rt][0] = e[rt][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
  for (int k0 = 0; k0 < ML_H; k0 += 4) {
    float a[ML_RT];
#pragma unroll
    for (int rt = 0; rt < ML_RT; ++rt) a[rt] = D[(rt * 16 + i) * ML_LDA + k0 + q];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const float b = bf2f(W[((nt0 + t) * 16 + i) * ML_LDW + k0 + q]);
#pragma unroll
      for (int rt = 0; rt < ML_RT; ++rt) e[rt][t] = mfma4(a[rt], b, e[rt][t]);
    }
  }
}

// g[mt][t][j][u] += sum_s H[s][16 mt + j] D[s][16 (nt0 + t) + u] over the pass's samples
template <int M, int LDH>
__device__ inline void ml_wgrad(const float* H, const float* D, int nt0, int lane,
                                f32x4 (&g)[M / 16][2]) {
  const int i = lane & 15, q = lane >> 4;
#pragma unroll
  for (int k0 = 0; k0 < ML_ROWS; k0 += 4) {
    const float b0 = D[(k0 + q) * ML_LDA + nt0 * 16 + i];
    const float b1 = D[(k0 + q) * ML_LDA + nt0 * 16 + 16 + i];
#pragma unroll
    for (int mt = 0; mt < M / 16; ++mt) {
      const float a = H[(k0 + q) * LDH + mt * 16 + i];
      g[mt][0] = mfma4(a, b0, g[mt][0]);
      g[mt][1] = mfma4(a, b1, g[mt][1]);
    }
  }
}

__device__ inline void ml_load_weights(unsigned short* dst, const unsigned short* src, int rows,
                                       int tid) {
  // src rows of 128 BF16 values, eight per 16-byte load (LDS rows are 272 bytes)
  const uint4* s = reinterpret_cast<const uint4*>(src);
  for (int e = tid; e < rows * (ML_H / 8); e += ML_TW) {
    const int r = e / (ML_H / 8), c = 8 * (e % (ML_H / 8));
    *reinterpret_cast<uint4*>(dst + r * ML_LDW + c) = s[e];
  }
}

__global__ __launch_bounds__(ML_TW) void ml_train_kernel(int n_samples, const int32_t* rows,
                                                         const float* targets,
                                                         const unsigned short* desc,
                                                         const unsigned short* wbf, const float* wq,
                                                         float* partial, float* loss_partial,
                                                         float inv_batch) {
  __shared__ __attribute__((aligned(16))) unsigned short sW1[ML_IN * ML_LDW];
  __shared__ __attribute__((aligned(16))) unsigned short sW2[ML_H * ML_LDW], sW3[ML_H * ML_LDW];
  __shared__ float sX[ML_ROWS * ML_LDX], sH1[ML_ROWS * ML_LDA], sH2[ML_ROWS * ML_LDA],
      sD[ML_ROWS * ML_LDA];
  __shared__ float sB1[ML_H], sB2[ML_H], sB3[ML_H], sW4[ML_H], sRed[4][ML_ROWS], sY[ML_ROWS];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, i = lane & 15, q = lane >> 4;
  const int nt0 = 2 * wave;
  ml_load_weights(sW1, wbf + ML_OFF_W1, ML_IN, tid);
  ml_load_weights(sW2, wbf + ML_OFF_W2, ML_H, tid);
  ml_load_weights(sW3, wbf + ML_OFF_W3, ML_H, tid);
  if (tid < ML_H) {
    sB1[tid] = wq[ML_OFF_B1 + tid];
    sB2[tid] = wq[ML_OFF_B2 + tid];
    sB3[tid] = wq[ML_OFF_B3 + tid];
    sW4[tid] = wq[ML_OFF_W4 + tid];
  }
  const float b4 = wq[ML_OFF_B4];
  f32x4 g1[ML_IN / 16][2], g2[ML_H / 16][2], g3[ML_H / 16][2];
#pragma unroll
  for (int mt = 0; mt < ML_H / 16; ++mt)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      g2[mt][t] = g3[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
      if (mt < ML_IN / 16) g1[mt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  float gb1[2] = {0.f, 0.f}, gb2[2] = {0.f, 0.f}, gb3[2] = {0.f, 0.f}, gw4[2] = {0.f, 0.f};
  float gb4 = 0.f, lsum = 0.f;
  const int s0 = blockIdx.x * ML_SLICE;
  const int s1 = s0 + ML_SLICE < n_samples ? s0 + ML_SLICE : n_samples;
  for (int t0 = s0; t0 < s1; t0 += ML_ROWS) {
    __syncthreads();  // the previous pass is done with sX, sH1, sH2, sD
    for (int e = tid; e < ML_ROWS * ML_IN; e += ML_TW) {
      const int r = e / ML_IN, c = e % ML_IN;
      sX[r * ML_LDX + c] = t0 + r < s1 ? bf2f(desc[(size_t)rows[t0 + r] * ML_IN + c]) : 0.f;
    }
    if (tid < ML_ROWS) sY[tid] = t0 + tid < s1 ? targets[t0 + tid] : 0.f;
    __syncthreads();
    // forward; z1, z2, z3 keep the pre-activations of the wave's columns
    f32x4 z1[ML_RT][2], z2[ML_RT][2], z3[ML_RT][2];
    ml_fwd_cols<ML_IN, ML_LDX>(sX, sW1, nt0, lane, z1);
#pragma unroll
    for (int rt = 0; rt < ML_RT; ++rt)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int c = (nt0 + t) * 16 + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          z1[rt][t][r] += sB1[c];
          sH1[(rt * 16 + 4 * q + r) * ML_LDA + c] = silu(z1[rt][t][r]);
        }
      }
    __syncthreads();
    ml_fwd_cols<ML_H, ML_LDA>(sH1, sW2, nt0, lane, z2);
#pragma unroll
    for (int rt = 0; rt < ML_RT; ++rt)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int c = (nt0 + t) * 16 + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          z2[rt][t][r] += sB2[c];
          sH2[(rt * 16 + 4 * q + r) * ML_LDA + c] = silu(z2[rt][t][r]);
        }
      }
    __syncthreads();
    ml_fwd_cols<ML_H, ML_LDA>(sH2, sW3, nt0, lane, z3);
    float h3[ML_RT][2][4], p[ML_RT][4];
#pragma unroll
    for (int rt = 0; rt < ML_RT; ++rt) {
#pragma unroll
      for (int r = 0; r < 4; ++r) p[rt][r] = 0.f;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int c = (nt0 + t) * 16 + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          z3[rt][t][r] += sB3[c];
          h3[rt][t][r] = silu(z3[rt][t][r]);
          p[rt][r] += h3[rt][t][r] * sW4[c];
        }
      }
      // output: row sums over the 16 lanes of a row group, then over the waves
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        p[rt][r] += __shfl_xor(p[rt][r], 1, 64);
        p[rt][r] += __shfl_xor(p[rt][r], 2, 64);
        p[rt][r] += __shfl_xor(p[rt][r], 4, 64);
        p[rt][r] += __shfl_xor(p[rt][r], 8, 64);
      }
      if (i == 0)
#pragma unroll
        for (int r = 0; r < 4; ++r) sRed[wave][rt * 16 + 4 * q + r] = p[rt][r];
    }
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < ML_RT; ++rt) {
      float dout[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rt * 16 + 4 * q + r;
        const float out = sRed[0][row] + sRed[1][row] + sRed[2][row] + sRed[3][row] + b4;
        const float diff = t0 + row < s1 ? out - sY[row] : 0.f;
        dout[r] = diff * inv_batch;
        if (wave == 0 && i == 0) {
          lsum += 0.5f * diff * diff;
          gb4 += dout[r];
        }
      }
      // layer 3 error
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int c = (nt0 + t) * 16 + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          gw4[t] += dout[r] * h3[rt][t][r];
          const float d = dout[r] * sW4[c] * dsilu(z3[rt][t][r]);
          gb3[t] += d;
          sD[(rt * 16 + 4 * q + r) * ML_LDA + c] = d;
        }
      }
    }
    __syncthreads();
    f32x4 e[ML_RT][2];
    ml_wgrad<ML_H, ML_LDA>(sH2, sD, nt0, lane, g3);
    ml_bwd_cols(sD, sW3, nt0, lane, e);
    __syncthreads();  // every wave has read the layer-3 error
#pragma unroll
    for (int rt = 0; rt < ML_RT; ++rt)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int c = (nt0 + t) * 16 + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float d = e[rt][t][r] * dsilu(z2[rt][t][r]);
          gb2[t] += d;
          sD[(rt * 16 + 4 * q + r) * ML_LDA + c] = d;
        }
      }
    __syncthreads();
    ml_wgrad<ML_H, ML_LDA>(sH1, sD, nt0, lane, g2);
    ml_bwd_cols(sD, sW2, nt0, lane, e);
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < ML_RT; ++rt)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int c = (nt0 + t) * 16 + i;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float d = e[rt][t][r] * dsilu(z1[rt][t][r]);
          gb1[t] += d;
          sD[(rt * 16 + 4 * q + r) * ML_LDA + c] = d;
        }
      }
    __syncthreads();
    ml_wgrad<ML_IN, ML_LDX>(sX, sD, nt0, lane, g1);
  }
  // the workgroup's partial gradient
  float* P = partial + (size_t)blockIdx.x * ML_NP;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int u = (nt0 + t) * 16 + i;
#pragma unroll
    for (int mt = 0; mt < ML_H / 16; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = mt * 16 + 4 * q + r;
        P[ML_OFF_W2 + j * ML_H + u] = g2[mt][t][r];
        P[ML_OFF_W3 + j * ML_H + u] = g3[mt][t][r];
        if (mt < ML_IN / 16) P[ML_OFF_W1 + j * ML_H + u] = g1[mt][t][r];
      }
    float v1 = gb1[t], v2 = gb2[t], v3 = gb3[t], v4 = gw4[t];
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      v1 += __shfl_xor(v1, o, 64);
      v2 += __shfl_xor(v2, o, 64);
      v3 += __shfl_xor(v3, o, 64);
      v4 += __shfl_xor(v4, o, 64);
    }
    if (q == 0) {
      P[ML_OFF_B1 + u] = v1;
      P[ML_OFF_B2 + u] = v2;
      P[ML_OFF_B3 + u] = v3;
      P[ML_OFF_W4 + u] = v4;
    }
  }
  if (wave == 0) {
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      gb4 += __shfl_xor(gb4, o, 64);
      lsum += __shfl_xor(lsum, o, 64);
    }
    if (lane == 0) {
      P[ML_OFF_B4] = gb4;
      loss_partial[blockIdx.x] = lsum;
    }
  }
}

// Sum of the workgroup partials, then Adam on the FP32 master weights and the
// BF16 copies (rounded weights as FP32 for training, raw BF16 for inference).
// A 256-thread block takes 32 parameters; its 8 thread rows each sum every
// 8th partial and the 8 sums are added in a fixed order, so the result does
// not depend on scheduling.
constexpr int ML_RED_P = 32, ML_RED_S = 8;

__global__ __launch_bounds__(ML_RED_P * ML_RED_S) void ml_adam_kernel(
    int n_part, const float* partial, float* grad_out, float* master, float* m1, float* m2,
    float* wq, unsigned short* wbf, float lr, float beta1, float beta2, float eps, int t,
    int apply) {
  __shared__ float red[ML_RED_S][ML_RED_P];
  const int lp = threadIdx.x % ML_RED_P, sl = threadIdx.x / ML_RED_P;
  const int p = blockIdx.x * ML_RED_P + lp;
  float g = 0.f;
  if (p < ML_NP)
    for (int w = sl; w < n_part; w += ML_RED_S) g += partial[(size_t)w * ML_NP + p];
  red[sl][lp] = g;
  __syncthreads();
  if (sl != 0 || p >= ML_NP) return;
  g = 0.f;
#pragma unroll
  for (int k = 0; k < ML_RED_S; ++k) g += red[k][lp];
  if (grad_out) grad_out[p] = g;
  if (!apply) return;
  const float a = beta1 * m1[p] + (1 - beta1) * g;
  const float b = beta2 * m2[p] + (1 - beta2) * g * g;
  m1[p] = a;
  m2[p] = b;
  const float ah = a / (1 - powf(beta1, (float)t)), bh = b / (1 - powf(beta2, (float)t));
  const float w = master[p] - lr * ah / (sqrtf(bh) + eps);
  master[p] = w;
  const unsigned short h = f2bf(w);
  wbf[p] = h;
  wq[p] = bf2f(h);
}

__global__ void ml_round_kernel(const float* master, float* wq, unsigned short* wbf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ML_NP) return;
  const unsigned short h = f2bf(master[p]);
  wbf[p] = h;
  wq[p] = bf2f(h);
}

// ------------------------------------------------------ proposal table ----
// q(v) = (1 - mix) softmax(logit / temp)(v) + mix / nvox; cdf[v] = sum_{u<=v} q(u)
// with cdf[nvox-1] = 1.  One workgroup per chain, FP64 sums.
__global__ __launch_bounds__(256) void ml_table_kernel(int nvox, const float* logits, float temp,
                                                       float mix, float* cdf) {
  __shared__ double red[4];
  __shared__ double s_max, s_sum, carry;
  const int ch = blockIdx.x, tid = threadIdx.x;
  const float* lg = logits + (size_t)ch * nvox;
  float* c = cdf + (size_t)ch * nvox;
  double mx = -1e300;
  for (int v = tid; v < nvox; v += 256) mx = fmax(mx, (double)lg[v] / temp);
  for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
  if ((tid & 63) == 0) red[tid >> 6] = mx;
  __syncthreads();
  if (tid == 0) s_max = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  __syncthreads();
  double sm = 0.0;
  for (int v = tid; v < nvox; v += 256) sm += exp((double)lg[v] / temp - s_max);
  sm = wave_sum(sm);
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = sm;
  __syncthreads();
  if (tid == 0) { s_sum = red[0] + red[1] + red[2] + red[3]; carry = 0.0; }
  __syncthreads();
  // chunked inclusive scan in fixed order
  for (int base = 0; base < nvox; base += 256) {
    const int v = base + tid;
    double q = 0.0;
    if (v < nvox) q = (1.0 - mix) * exp((double)lg[v] / temp - s_max) / s_sum + mix / nvox;
    // block inclusive scan of q (wave scan + wave offsets)
    double x = q;
    for (int o = 1; o < 64; o <<= 1) {
      const double y = __shfl_up(x, o, 64);
      if ((tid & 63) >= o) x += y;
    }
    __syncthreads();
    if ((tid & 63) == 63) red[tid >> 6] = x;
    __syncthreads();
    double off = carry;
    for (int w = 0; w < (tid >> 6); ++w) off += red[w];
    if (v < nvox) c[v] = v == nvox - 1 ? 1.0f : (float)(off + x);
    __syncthreads();
    if (tid == 255) carry = off + x;
    __syncthreads();
  }
}

// guide[j] = first v with j / nvox < cdf[v] (binary search per entry).
__global__ __launch_bounds__(256) void ml_guide_kernel(int nvox, const float* cdf, int32_t* guide) {
  const int j = blockIdx.x * 256 + threadIdx.x, ch = blockIdx.y;
  if (j >= nvox) return;
  const float* c = cdf + (size_t)ch * nvox;
  const float u = (float)j / (float)nvox;
  int lo = 0, hi = nvox - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (u < c[mid]) hi = mid; else lo = mid + 1;
  }
  guide[(size_t)ch * nvox + j] = lo;
}

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_mlp_param_count(void) { return ML_NP; }

extern "C" int kupsmi_voxel_descriptors(void* stream, int n_chain, int nv, int max_host,
                                        const int32_t* host_off, const float* host_xyzq,
                                        const double* tril, float rmax, float rho0,
                                        uint16_t* desc) {
  KUPSMI_REQUIRE(n_chain >= 0 && nv >= 1 && max_host >= 0, "bad sizes");
  KUPSMI_REQUIRE((size_t)max_host * 16 <= 64 * 1024, "hosts of at most 4096 atoms");
  if (n_chain == 0) return KUPSMI_OK;
  const int nvox = nv * nv * nv;
  dim3 g((nvox + 255) / 256, n_chain);
  ml_desc_kernel<<<g, 256, (size_t)max_host * 16, (hipStream_t)stream>>>(
      n_chain, nv, host_off, reinterpret_cast<const float4*>(host_xyzq), tril, rmax, rho0, desc);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_mlp_infer(void* stream, int n_rows, const uint16_t* desc,
                                const uint16_t* weights_bf16, const float* weights_f32,
                                float* logits) {
  KUPSMI_REQUIRE(n_rows >= 0, "negative size");
  if (n_rows == 0) return KUPSMI_OK;
  int dev = 0, ncu = 256;
  KUPSMI_HIP(hipGetDevice(&dev));
  KUPSMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int ntiles = (n_rows + 15) / 16;
  int grid = (ntiles + ML_WAVES - 1) / ML_WAVES;
  if (grid > 8 * ncu) grid = 8 * ncu;
  ml_infer_kernel<<<grid, 64 * ML_WAVES, 0, (hipStream_t)stream>>>(
      n_rows, desc, weights_bf16, weights_f32, logits);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_mlp_train_step(kupsmi_ctx* ctx, void* stream, int n_samples,
                                     const int32_t* rows, const float* targets,
                                     const uint16_t* desc, float* master, float* m1, float* m2,
                                     float* weights_f32, uint16_t* weights_bf16, float lr,
                                     int adam_step, int apply, float* grad_out,
                                     float* loss_out) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  KUPSMI_REQUIRE(n_samples >= 1, "empty minibatch");
  KUPSMI_REQUIRE(((uintptr_t)weights_bf16 & 15) == 0, "BF16 weights must be 16-byte aligned");
  const int n_part = (n_samples + ML_SLICE - 1) / ML_SLICE;
  void* buf = nullptr;
  int rc = scratch(ctx, 5, sizeof(float) * ((size_t)n_part * ML_NP + n_part + 64), &buf);
  if (rc) return rc;
  float* partial = static_cast<float*>(buf);
  float* loss_partial = loss_out ? loss_out : partial + (size_t)n_part * ML_NP;
  hipStream_t st = (hipStream_t)stream;
  ml_train_kernel<<<n_part, ML_TW, 0, st>>>(n_samples, rows, targets, desc, weights_bf16,
                                            weights_f32, partial, loss_partial, 1.0f / n_samples);
  KUPSMI_CHECK_LAUNCH();
  ml_adam_kernel<<<(ML_NP + ML_RED_P - 1) / ML_RED_P, ML_RED_P * ML_RED_S, 0, st>>>(
      n_part, partial, grad_out, master, m1, m2, weights_f32, weights_bf16, lr, 0.9f, 0.999f,
      1e-8f, adam_step, apply);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_mlp_round_weights(void* stream, const float* master, float* weights_f32,
                                        uint16_t* weights_bf16) {
  ml_round_kernel<<<(ML_NP + 255) / 256, 256, 0, (hipStream_t)stream>>>(master, weights_f32,
                                                                         weights_bf16);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_proposal_table(void* stream, int n_chain, int nvox, const float* logits,
                                     float temperature, float mix, float* cdf, int32_t* guide) {
  KUPSMI_REQUIRE(n_chain >= 0 && nvox >= 1, "bad sizes");
  KUPSMI_REQUIRE(mix > 0.f && mix <= 1.f, "uniform mixing weight must be in (0, 1]");
  KUPSMI_REQUIRE(temperature > 0.f, "temperature must be positive");
  KUPSMI_REQUIRE(guide != nullptr, "the guide table is required");
  if (n_chain == 0) return KUPSMI_OK;
  ml_table_kernel<<<n_chain, 256, 0, (hipStream_t)stream>>>(nvox, logits, temperature, mix, cdf);
  KUPSMI_CHECK_LAUNCH();
  ml_guide_kernel<<<dim3((nvox + 255) / 256, n_chain), 256, 0, (hipStream_t)stream>>>(nvox, cdf,
                                                                                     guide);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}
