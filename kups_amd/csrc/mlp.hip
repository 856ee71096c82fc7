// Learned GCMC insertion proposal: voxel descriptors, MLP inference (BF16
// MFMA), MLP training (FP32 forward/backward with BF16-rounded weights, Adam)
// and the per-chain proposal table.  No counterpart in the reference: the
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

typedef short bf16x4 __attribute__((ext_vector_type(4)));
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
// One wave per 16-row tile: three 16 x 128 layers as v_mfma_f32_16x16x16_bf16
// (A: 16 x 16 activations, B: 16 x 16 weights, FP32 accumulators).  The eight
// 16-column output tiles of a layer are accumulated together, so eight
// independent MFMA chains share each A fragment; bias and SiLU are applied on
// the accumulators (reciprocal by v_rcp_f32: the result is rounded to BF16),
// activations are re-staged in LDS as BF16.  Weights are held transposed
// ([out][in]) in LDS so that every B fragment is one 8-byte load; rows are
// padded by 8 elements so the 16 rows of a fragment fall in different banks.
constexpr int ML_WAVES = 8;
constexpr int ML_LD_IN = ML_IN + 8;  // padded row lengths (BF16 elements)
constexpr int ML_LD_H = ML_H + 8;

__device__ inline float silu_bf(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

template <int KIN, int LDK>
__device__ inline void ml_layer(const unsigned short* act, const unsigned short* wT,
                                const float* bias, unsigned short* out, int lane) {
  const int row = lane & 15, kq = (lane >> 4) * 4;
  f32x4 acc[ML_H / 16];
#pragma unroll
  for (int nt = 0; nt < ML_H / 16; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KIN / 16; ++ks) {
    const bf16x4 a = *reinterpret_cast<const bf16x4*>(act + row * LDK + ks * 16 + kq);
#pragma unroll
    for (int nt = 0; nt < ML_H / 16; ++nt) {
      const bf16x4 b = *reinterpret_cast<const bf16x4*>(wT + (nt * 16 + row) * LDK + ks * 16 + kq);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
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
// One 128-thread workgroup per slice of a minibatch (thread u = hidden unit
// u), forward and backward of every sample with BF16-rounded weights in FP32
// arithmetic, gradients accumulated in LDS (147 KB) and written as one
// partial per workgroup; loss = 1/2 mean (f(x) - y)^2.
__global__ __launch_bounds__(ML_H) void ml_train_kernel(int n_samples, int slice,
                                                        const int32_t* rows,
                                                        const float* targets,
                                                        const unsigned short* desc,
                                                        const float* wq, float* partial,
                                                        float* loss_partial, float inv_batch) {
  __shared__ float gW1[ML_IN * ML_H];
  __shared__ float gW2[ML_H * ML_H];
  __shared__ float gW3[ML_H * ML_H];
  __shared__ float x[ML_IN], h1[ML_H], h2[ML_H], h3[ML_H], g1[ML_H], g2[ML_H], g3[ML_H];
  __shared__ float red[ML_H / 64];
  const int u = threadIdx.x;
  for (int e = u; e < ML_IN * ML_H; e += ML_H) gW1[e] = 0.f;
  for (int e = u; e < ML_H * ML_H; e += ML_H) { gW2[e] = 0.f; gW3[e] = 0.f; }
  float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gw4 = 0.f, gb4 = 0.f, lsum = 0.f;
  const float* W1 = wq + ML_OFF_W1;
  const float* W2 = wq + ML_OFF_W2;
  const float* W3 = wq + ML_OFF_W3;
  const float w4u = wq[ML_OFF_W4 + u];
  const int s0 = blockIdx.x * slice;
  const int s1 = s0 + slice < n_samples ? s0 + slice : n_samples;
  for (int s = s0; s < s1; ++s) {
    __syncthreads();
    if (u < ML_IN) x[u] = bf2f(desc[(size_t)rows[s] * ML_IN + u]);
    __syncthreads();
    float z1 = wq[ML_OFF_B1 + u];
    for (int k = 0; k < ML_IN; ++k) z1 += x[k] * W1[k * ML_H + u];
    const float a1 = silu(z1);
    h1[u] = a1;
    __syncthreads();
    float z2 = wq[ML_OFF_B2 + u];
    for (int j = 0; j < ML_H; ++j) z2 += h1[j] * W2[j * ML_H + u];
    const float a2 = silu(z2);
    h2[u] = a2;
    __syncthreads();
    float z3 = wq[ML_OFF_B3 + u];
    for (int j = 0; j < ML_H; ++j) z3 += h2[j] * W3[j * ML_H + u];
    const float a3 = silu(z3);
    h3[u] = a3;
    // output = b4 + sum_u h3[u] w4[u]
    float p = a3 * w4u;
    p = wave_sumf(p);
    if ((u & 63) == 0) red[u >> 6] = p;
    __syncthreads();
    const float out = red[0] + red[1] + wq[ML_OFF_B4];
    const float diff = out - targets[s];
    if (u == 0) lsum += 0.5f * diff * diff;
    const float dout = diff * inv_batch;
    gb4 += dout;
    gw4 += dout * a3;
    auto dsilu = [](float z) {
      const float sg = 1.0f / (1.0f + __expf(-z));
      return sg * (1.0f + z * (1.0f - sg));
    };
    g3[u] = dout * w4u * dsilu(z3);
    __syncthreads();
    const float d3 = g3[u];
    gb3 += d3;
    for (int j = 0; j < ML_H; ++j) gW3[j * ML_H + u] += h2[j] * d3;
    float e2 = 0.f;
    for (int v = 0; v < ML_H; ++v) e2 += g3[v] * W3[u * ML_H + v];
    g2[u] = e2 * dsilu(z2);
    __syncthreads();
    const float d2 = g2[u];
    gb2 += d2;
    for (int j = 0; j < ML_H; ++j) gW2[j * ML_H + u] += h1[j] * d2;
    float e1 = 0.f;
    for (int v = 0; v < ML_H; ++v) e1 += g2[v] * W2[u * ML_H + v];
    g1[u] = e1 * dsilu(z1);
    __syncthreads();
    const float d1 = g1[u];
    gb1 += d1;
    for (int k = 0; k < ML_IN; ++k) gW1[k * ML_H + u] += x[k] * d1;
  }
  __syncthreads();
  float* P = partial + (size_t)blockIdx.x * ML_NP;
  for (int e = u; e < ML_IN * ML_H; e += ML_H) P[ML_OFF_W1 + e] = gW1[e];
  for (int e = u; e < ML_H * ML_H; e += ML_H) { P[ML_OFF_W2 + e] = gW2[e]; P[ML_OFF_W3 + e] = gW3[e]; }
  P[ML_OFF_B1 + u] = gb1;
  P[ML_OFF_B2 + u] = gb2;
  P[ML_OFF_B3 + u] = gb3;
  P[ML_OFF_W4 + u] = gw4;
  if (u == 0) {
    P[ML_OFF_B4] = gb4;
    loss_partial[blockIdx.x] = lsum;
  }
}

// Sum of the workgroup partials (fixed order), then Adam on the FP32 master
// weights and the BF16 copies (rounded weights as FP32 for training, raw
// BF16 for inference).
__global__ void ml_adam_kernel(int n_part, const float* partial, float* grad_out, float* master,
                               float* m1, float* m2, float* wq, unsigned short* wbf, float lr,
                               float beta1, float beta2, float eps, int t, int apply) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ML_NP) return;
  float g = 0.f;
  for (int w = 0; w < n_part; ++w) g += partial[(size_t)w * ML_NP + p];
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
  const int slice = 32;
  const int n_part = (n_samples + slice - 1) / slice;
  void* buf = nullptr;
  int rc = scratch(ctx, 5, sizeof(float) * ((size_t)n_part * ML_NP + n_part + 64), &buf);
  if (rc) return rc;
  float* partial = static_cast<float*>(buf);
  float* loss_partial = loss_out ? loss_out : partial + (size_t)n_part * ML_NP;
  hipStream_t st = (hipStream_t)stream;
  ml_train_kernel<<<n_part, ML_H, 0, st>>>(n_samples, slice, rows, targets, desc, weights_f32,
                                           partial, loss_partial, 1.0f / n_samples);
  KUPSMI_CHECK_LAUNCH();
  ml_adam_kernel<<<(ML_NP + 255) / 256, 256, 0, st>>>(n_part, partial, grad_out, master, m1, m2,
                                                      weights_f32, weights_bf16, lr, 0.9f,
                                                      0.999f, 1e-8f, adam_step, apply);
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
