// Reciprocal-space Ewald sum (potential/classical/ewald.py:349-439, 585-601).
//
// k = 2 pi Linv n over the half-space integer grid n in [0..mx] x [-my..my] x
// [-mz..mz] (the grid of kvecs_from_kmax, ewald.py:1322-1339, which keeps the
// vectors with |k| <= kmax; kmax < 0 keeps the whole box).  Phases factorise
// per axis, exp(i k.r) = ex(n0) ey(n1) ez(n2) with e_a(n) = exp(2 pi i n f_a)
// on the wrapped fractional coordinates, so the per-atom work is a product of
// three small tables instead of a sincos per (atom, k).
//
//   ek_sk_kernel      S(k) = sum_i q_i exp(i k.r_i) as an FP32 GEMM on the
//                     matrix cores (v_mfma_f32_16x16x4_f32), one workgroup
//                     per (system, 512-atom chunk), atoms streamed through LDS
//                     in tiles of 8..64 (what 64 KB holds) with their per-axis
//                     phase tables, operands k-contiguous for float4 fragments;
//                     ek_sk_reduce_kernel sums the chunks in fixed order.
//   ek_energy_kernel  P(k) (FP64), E = K sum_k P |S|^2, the reciprocal-space
//                     virial, and S'(k) = P(k) S(k) for the force pass.
//   ek_force_kernel   F_i = -2 K q_i sum_k k Im(S'(k) exp(-i k.r_i)): one thread
//                     per atom, S' of its system resident in LDS.
#include <mutex>
#include <string>
#include "common.hpp"

namespace kupsmi {

constexpr int EK_THREADS = 256;
constexpr int EK_NM = 12;          // max |n| per axis
constexpr int EK_W = 2 * EK_NM + 1;
constexpr int EK_NKMAX = (EK_NM + 1) * EK_W * EK_W;

struct EkArgs {
  int n_sys, mx, my, mz, nk, accumulate, tiles;
  const int32_t* sys_off;
  const float* pos;
  const float* q;
  const double* tril;
  const double* alpha;
  const double* kmax;
  int chunks;       // atom chunks per system of the S(k) pass
  double* sk;       // n_sys x nk x 2 (FP64)
  double* skpart;   // n_sys x chunks x nk x 2: chunk partials of S
  float2* skp;      // n_sys x nk, P(k) S(k)
  float* forces;
  double* energy;
  double* virial;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// exp(2 pi i f) for the wrapped fractional coordinate f of a Cartesian r.
__device__ inline void unit_phases(const double* t, const float* r, float2& ux, float2& uy,
                                   float2& uz) {
  double li[6];
  tril_inverse(t, li);
  float f0, f1, f2;
  tmul(li, r[0], r[1], r[2], f0, f1, f2);
  f0 = wrap_half(f0); f1 = wrap_half(f1); f2 = wrap_half(f2);
  sincospif(2.0f * f0, &ux.y, &ux.x);
  sincospif(2.0f * f1, &uy.y, &uy.x);
  sincospif(2.0f * f2, &uz.y, &uz.x);
}

__device__ inline double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// c u^n for n = 0..m into out[0], out[st], .., out[m st] (FP64 recurrence,
// rounded once).
__device__ inline void powers64_pos(double2 u, double c, int m, float2* out, int st = 1) {
  double2 p = make_double2(c, 0.0);
  for (int n = 0; n <= m; ++n) {
    out[n * st] = make_float2((float)p.x, (float)p.y);
    p = zmul(p, u);
  }
}

// u^n for n = -m..m into out[0..2m] with stride st (n = 0 at out[m st];
// FP64, rounded once).
__device__ inline void powers64(double2 u, int m, float2* out, int st = 1) {
  out[m * st] = make_float2(1.f, 0.f);
  double2 p = make_double2(1.0, 0.0);
  for (int n = 1; n <= m; ++n) {
    p = zmul(p, u);
    out[(m + n) * st] = make_float2((float)p.x, (float)p.y);
    out[(m - n) * st] = make_float2((float)p.x, (float)-p.y);
  }
}

// Powers u^n for n = -m..m into out[0..2m] (n = 0 at out[m]).
__device__ inline void powers(float2 u, int m, float2* out) {
  out[m] = make_float2(1.f, 0.f);
  float2 p = make_float2(1.f, 0.f);
  for (int n = 1; n <= m; ++n) {
    p = cmul(p, u);
    out[m + n] = p;
    out[m - n] = make_float2(p.x, -p.y);
  }
}

// S(k) as a real FP32 GEMM on the matrix cores.  With m = (n0, n1)
// (M = (mx+1)(2my+1) rows) and A_i(m) = q_i X_i(n0) Y_i(n1):
//   S(m, n2) = sum_i A_i(m) Z_i(n2),
// stacked over (atom, re/im) on K = 2 x atoms and (re, im) of S on N = 2 wz:
//   S_re = sum Re A Re Z - Im A Im Z,  S_im = sum Re A Im Z + Im A Re Z.
// One workgroup per (system, chunk of EK_CHUNK atoms).  Per tile of TA atoms
// the workgroup builds the per-axis phase tables (FP64, rounded once; stored
// n-major so that two atoms are one 16-byte word), then packs both GEMM
// operands k-contiguous into LDS: A as [m][k] and B as [column][k], row pitch
// ldk = 2 TA + 4 floats (a quad count that is odd, so sixteen rows at the same
// k land on sixteen disjoint quads of banks).  K is summed in blocks of 16:
// lane (r = l % 16, q = l / 16) loads one float4 of A row r and one of B column
// r at k = k0 + 4q .. k0 + 4q + 3, and component j feeds the j-th
// v_mfma_f32_16x16x4_f32 of the block (each MFMA sums k0 + 4q + j over q; A and
// B use the same k order, so the four MFMAs together sum the whole block).
// D of an MFMA: lane l holds row 4 (l / 16) + r, column l % 16.  The 4 waves
// each own every 4th 16-row tile of M and all N tiles; their FP32 sums are
// added into FP64 registers after at most 64 atoms (K = 128), and the chunk's
// partial S goes to HBM for a fixed-order reduce.
constexpr int EK_SK_THREADS = 256;
constexpr int EK_WAVES = EK_SK_THREADS / 64;
constexpr int EK_CHUNK = 512;

struct EkSkShape {
  int ta, ldk;  // atoms per tile (multiple of 8), panel row pitch (floats)
  size_t lds;   // dynamic LDS bytes
};

template <int MTW, int NT>
__global__ __launch_bounds__(EK_SK_THREADS) void ek_sk_kernel(EkArgs a, EkSkShape sh) {
  extern __shared__ float ek_lds[];
  __shared__ double s_li[6];
  const int TA = sh.ta, ldk = sh.ldk, TH = TA / 2;
  const int wy = 2 * a.my + 1, wz = 2 * a.mz + 1, wx = a.mx + 1;
  const int M = wx * wy, NC = 2 * wz;
  const int mt_n = (M + 15) / 16;
  float* apk = ek_lds;                                   // [16 mt_n][ldk]
  float* bpk = apk + (size_t)16 * mt_n * ldk;            // [16 NT][ldk]
  float2* tx = reinterpret_cast<float2*>(bpk + (size_t)16 * NT * ldk);  // [wx][TA]
  float2* ty = tx + (size_t)wx * TA;                     // [wy][TA]
  float2* tz = ty + (size_t)wy * TA;                     // [wz][TA]
  const int s = blockIdx.y, chunk = blockIdx.x;
  const int i0 = a.sys_off[s], n_all = a.sys_off[s + 1] - i0;
  const int c0 = chunk * EK_CHUNK;
  if (c0 >= n_all) return;
  const int n = n_all - c0 < EK_CHUNK ? n_all - c0 : EK_CHUNK;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = lane & 15, kq = lane >> 4;
  if (tid == 0) tril_inverse(a.tril + 6 * s, s_li);
  // zero the panels once: padding rows / columns are never written again
  for (int e = tid; e < 16 * (mt_n + NT) * ldk; e += EK_SK_THREADS) apk[e] = 0.f;
  double sum[MTW][NT][4];
  f32x4 acc[MTW][NT];
#pragma unroll
  for (int t = 0; t < MTW; ++t)
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) sum[t][u][r] = 0.0;
    }
  // tiles between FP64 flushes: at most 64 atoms in an FP32 partial
  const int flush_every = TA >= 64 ? 1 : 64 / TA;
  int since_flush = 0;
  for (int base = 0; base < n; base += TA) {
    __syncthreads();
    const int cnt = n - base < TA ? n - base : TA;
    // 1. per-axis phase tables, n-major (q folded into x); atoms past the end
    //    are zero
    for (int e = tid; e < 3 * TA; e += EK_SK_THREADS) {
      const int axis = e / TA, at = e % TA;
      if (at < cnt) {
        const float* r = a.pos + 3 * (size_t)(i0 + c0 + base + at);
        double f0, f1, f2;
        tmul(s_li, (double)r[0], (double)r[1], (double)r[2], f0, f1, f2);
        double f = axis == 0 ? f0 : (axis == 1 ? f1 : f2);
        f -= rint(f);
        double2 u;
        sincospi(2.0 * f, &u.y, &u.x);
        if (axis == 0)
          powers64_pos(u, (double)a.q[i0 + c0 + base + at], a.mx, tx + at, TA);
        else if (axis == 1)
          powers64(u, a.my, ty + at, TA);
        else
          powers64(u, a.mz, tz + at, TA);
      } else {
        const int w = axis == 0 ? wx : (axis == 1 ? wy : wz);
        float2* o = axis == 0 ? tx + at : (axis == 1 ? ty + at : tz + at);
        for (int m = 0; m < w; ++m) o[(size_t)m * TA] = make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
    // 2. pack A[m][k] (k = 2 at + re/im), two atoms per item
    for (int e = tid; e < M * TH; e += EK_SK_THREADS) {
      const int m = e / TH, p = e - m * TH;
      const int n0 = m / wy, n1 = m - n0 * wy;
      const float4 x = *reinterpret_cast<const float4*>(tx + (size_t)n0 * TA + 2 * p);
      const float4 y = *reinterpret_cast<const float4*>(ty + (size_t)n1 * TA + 2 * p);
      const float2 A0 = cmul(make_float2(x.x, x.y), make_float2(y.x, y.y));
      const float2 A1 = cmul(make_float2(x.z, x.w), make_float2(y.z, y.w));
      *reinterpret_cast<float4*>(apk + (size_t)m * ldk + 4 * p) = make_float4(A0.x, A0.y, A1.x, A1.y);
    }
    //    and B[j][k]: j < wz -> Re S (Re Z, -Im Z), else Im S (Im Z, Re Z)
    for (int e = tid; e < wz * TH; e += EK_SK_THREADS) {
      const int j = e / TH, p = e - j * TH;
      const float4 z = *reinterpret_cast<const float4*>(tz + (size_t)j * TA + 2 * p);
      *reinterpret_cast<float4*>(bpk + (size_t)j * ldk + 4 * p) = make_float4(z.x, -z.y, z.z, -z.w);
      *reinterpret_cast<float4*>(bpk + (size_t)(wz + j) * ldk + 4 * p) = make_float4(z.y, z.x, z.w, z.z);
    }
    __syncthreads();
    // 3. matrix-core products over this tile's K = 2 TA, 16 k per block
    const float* ap = apk + (size_t)(wave * 16 + row) * ldk + 4 * kq;
    const float* bp = bpk + (size_t)row * ldk + 4 * kq;
    for (int k0 = 0; k0 < 2 * TA; k0 += 16) {
      f32x4 bv[NT];
#pragma unroll
      for (int u = 0; u < NT; ++u)
        bv[u] = *reinterpret_cast<const f32x4*>(bp + (size_t)u * 16 * ldk + k0);
#pragma unroll
      for (int t = 0; t < MTW; ++t) {
        if (wave + t * EK_WAVES >= mt_n) break;
        const f32x4 av =
            *reinterpret_cast<const f32x4*>(ap + (size_t)t * EK_WAVES * 16 * ldk + k0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int u = 0; u < NT; ++u)
            acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[j], bv[u][j], acc[t][u], 0, 0, 0);
      }
    }
    if (++since_flush == flush_every || base + TA >= n) {
      since_flush = 0;
#pragma unroll
      for (int t = 0; t < MTW; ++t)
#pragma unroll
        for (int u = 0; u < NT; ++u) {
#pragma unroll
          for (int r = 0; r < 4; ++r) sum[t][u][r] += (double)acc[t][u][r];
          acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    }
  }
  // D[row 4 (l/16) + r][col l % 16] of tile (t, u): m = 16 mt + row, column j
  double* part = a.skpart + ((size_t)s * a.chunks + chunk) * a.nk * 2;
#pragma unroll
  for (int t = 0; t < MTW; ++t) {
    const int mt = wave + t * EK_WAVES;
    if (mt >= mt_n) continue;
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      const int j = u * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = mt * 16 + (lane >> 4) * 4 + r;
        if (m >= M || j >= NC) continue;
        const int k = m * wz + (j < wz ? j : j - wz);
        part[2 * k + (j < wz ? 0 : 1)] = sum[t][u][r];
      }
    }
  }
}

// S of each system: fixed-order sum of its chunk partials.
__global__ void ek_sk_reduce_kernel(EkArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (k >= a.nk) return;
  const int n = a.sys_off[s + 1] - a.sys_off[s];
  const int nc = (n + EK_CHUNK - 1) / EK_CHUNK;
  double re = 0.0, im = 0.0;
  for (int c = 0; c < nc; ++c) {
    const double* p = a.skpart + (((size_t)s * a.chunks + c) * a.nk + k) * 2;
    re += p[0];
    im += p[1];
  }
  double* o = a.sk + 2 * ((size_t)s * a.nk + k);
  o[0] = re;
  o[1] = im;
}

template <int MTW>
void ek_sk_launch_nt(const EkArgs& a, const EkSkShape& sh, int nt, dim3 g, hipStream_t st) {
  switch (nt) {
    case 1: ek_sk_kernel<MTW, 1><<<g, EK_SK_THREADS, sh.lds, st>>>(a, sh); break;
    case 2: ek_sk_kernel<MTW, 2><<<g, EK_SK_THREADS, sh.lds, st>>>(a, sh); break;
    case 3: ek_sk_kernel<MTW, 3><<<g, EK_SK_THREADS, sh.lds, st>>>(a, sh); break;
    default: ek_sk_kernel<MTW, 4><<<g, EK_SK_THREADS, sh.lds, st>>>(a, sh); break;
  }
}

// Tile size: the largest TA in {64, 56, .., 8} whose panels and tables fit
// 64 KB (two workgroups per CU); ldk = 2 TA + 4 has an odd quad count for TA
// a multiple of 4.
EkSkShape ek_sk_shape(const EkArgs& a) {
  const int wx = a.mx + 1, wy = 2 * a.my + 1, wz = 2 * a.mz + 1;
  const int rows = 16 * ((wx * wy + 15) / 16) + 16 * ((2 * wz + 15) / 16);
  EkSkShape sh{};
  for (int ta = 64; ta >= 8; ta -= 8) {
    const int ldk = 2 * ta + 4;
    const size_t lds = sizeof(float) * (size_t)rows * ldk + sizeof(float2) * (size_t)(wx + wy + wz) * ta;
    sh.ta = ta;
    sh.ldk = ldk;
    sh.lds = lds;
    if (lds <= 64 * 1024) break;
  }
  return sh;
}

void ek_sk_launch(const EkArgs& a, dim3 g, hipStream_t st) {
  const int wy = 2 * a.my + 1, wz = 2 * a.mz + 1;
  const int mt = ((a.mx + 1) * wy + 15) / 16, nt = (2 * wz + 15) / 16;
  const int mtw = (mt + EK_WAVES - 1) / EK_WAVES;
  const EkSkShape sh = ek_sk_shape(a);
  switch (mtw) {
    case 1: ek_sk_launch_nt<1>(a, sh, nt, g, st); break;
    case 2: ek_sk_launch_nt<2>(a, sh, nt, g, st); break;
    case 3: ek_sk_launch_nt<3>(a, sh, nt, g, st); break;
    case 4: ek_sk_launch_nt<4>(a, sh, nt, g, st); break;
    default: ek_sk_launch_nt<6>(a, sh, nt, g, st); break;  // up to 21 tiles (|n| <= 12)
  }
}

__global__ __launch_bounds__(EK_THREADS) void ek_energy_kernel(EkArgs a) {
  __shared__ double red[EK_THREADS / 64][7];
  const int s = blockIdx.x, tid = threadIdx.x;
  const double* t = a.tril + 6 * s;
  double li[6];
  tril_inverse(t, li);
  const double vol = fabs(t[0] * t[2] * t[5]);
  const double al = a.alpha[s];
  const double km = a.kmax ? a.kmax[s] : -1.0;
  const double K = 14.399645351950548;
  const int wy = 2 * a.my + 1, wz = 2 * a.mz + 1;
  double e = 0, w00 = 0, w01 = 0, w02 = 0, w11 = 0, w12 = 0, w22 = 0;
  for (int k = tid; k < a.nk; k += EK_THREADS) {
    const int n0 = k / (wy * wz), n1 = (k / wz) % wy - a.my, n2 = k % wz - a.mz;
    const double kv0 = 2 * M_PI * (li[0] * n0);
    const double kv1 = 2 * M_PI * (li[1] * n0 + li[2] * n1);
    const double kv2 = 2 * M_PI * (li[3] * n0 + li[4] * n1 + li[5] * n2);
    const double k2 = kv0 * kv0 + kv1 * kv1 + kv2 * kv2;
    double P = 0.0;
    const bool in = km < 0 || sqrt(k2) <= km;
    if (k2 > 0 && in) P = 2 * M_PI / vol * exp(-k2 / (4 * al * al)) / k2 * (n0 == 0 ? 1.0 : 2.0);
    const double* sk = a.sk + 2 * ((size_t)s * a.nk + k);
    const double S2 = sk[0] * sk[0] + sk[1] * sk[1];
    const double ek = K * P * S2;
    e += ek;
    if (P > 0) {
      const double f = 2.0 * (1.0 / k2 + 1.0 / (4 * al * al));
      w00 += ek * (1 - f * kv0 * kv0); w11 += ek * (1 - f * kv1 * kv1); w22 += ek * (1 - f * kv2 * kv2);
      w01 -= ek * f * kv0 * kv1; w02 -= ek * f * kv0 * kv2; w12 -= ek * f * kv1 * kv2;
    }
    a.skp[(size_t)s * a.nk + k] = make_float2((float)(P * sk[0]), (float)(P * sk[1]));
  }
  const double v[7] = {e, w00, w01, w02, w11, w12, w22};
  const int lane = tid & 63, wv = tid >> 6;
  for (int c = 0; c < 7; ++c) {
    double x = wave_sum(v[c]);
    if (lane == 0) red[wv][c] = x;
  }
  __syncthreads();
  if (tid == 0) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int w = 0; w < EK_THREADS / 64; ++w)
      for (int c = 0; c < 7; ++c) acc[c] += red[w][c];
    if (a.energy) a.energy[s] = a.accumulate ? a.energy[s] + acc[0] : acc[0];
    if (a.virial) {
      const double W[9] = {acc[1], acc[2], acc[3], acc[2], acc[4], acc[5], acc[3], acc[5], acc[6]};
      for (int c = 0; c < 9; ++c)
        a.virial[9 * (size_t)s + c] = a.accumulate ? a.virial[9 * (size_t)s + c] + W[c] : W[c];
    }
  }
}

// MZ > 0: the k grid's z half width as a constant (inner loop fully
// unrolled, (n2) factors folded); MZ = 0: read from the arguments.
template <int MZ>
__global__ __launch_bounds__(EK_THREADS) void ek_force_kernel(EkArgs a) {
  extern __shared__ float2 skp[];
  const int s = blockIdx.y;
  const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
  const int tid = threadIdx.x;
  const int start = blockIdx.x * EK_THREADS;
  if (start >= n) return;
  for (int k = tid; k < a.nk; k += EK_THREADS) skp[k] = a.skp[(size_t)s * a.nk + k];
  __syncthreads();
  const int i = start + tid;
  if (i >= n) return;
  const double* t = a.tril + 6 * s;
  float2 ux, uy, uz;
  unit_phases(t, a.pos + 3 * (size_t)(i0 + i), ux, uy, uz);
  const int mz = MZ > 0 ? MZ : a.mz;
  const int wy = 2 * a.my + 1, wz = 2 * mz + 1;
  // u^-m starting points
  float2 uy0 = make_float2(1.f, 0.f), uz0 = make_float2(1.f, 0.f);
  const float2 cy = make_float2(uy.x, -uy.y), cz = make_float2(uz.x, -uz.y);
  for (int m = 0; m < a.my; ++m) uy0 = cmul(uy0, cy);
  for (int m = 0; m < mz; ++m) uz0 = cmul(uz0, cz);
  float t0 = 0.f, t1 = 0.f, t2 = 0.f;
  float2 ex = make_float2(1.f, 0.f);
  int k = 0;
  for (int n0 = 0; n0 <= a.mx; ++n0) {
    float2 exy = cmul(ex, uy0);
    for (int j1 = 0; j1 < wy; ++j1) {
      const float n1 = (float)(j1 - a.my);
      float2 e = cmul(exy, uz0);
      float a0 = 0.f, a2 = 0.f;
#pragma unroll
      for (int j2 = 0; j2 < (MZ > 0 ? 2 * MZ + 1 : wz); ++j2) {
        const float2 sp = skp[k + j2];
        // Im(S' conj(e))
        const float v = sp.y * e.x - sp.x * e.y;
        a0 += v;
        a2 += (float)(j2 - mz) * v;
        e = cmul(e, uz);
      }
      k += wz;
      t0 += (float)n0 * a0;
      t1 += n1 * a0;
      t2 += a2;
      exy = cmul(exy, uy);
    }
    ex = cmul(ex, ux);
  }
  // F_a = -2 K q 2 pi sum_b Li[a][b] T_b  (k_a = 2 pi sum_b Li[a][b] n_b)
  double li[6];
  tril_inverse(t, li);
  const float c = -2.0f * 14.399645351950548f * a.q[i0 + i] * 6.283185307179586f;
  const float F0 = c * ((float)li[0] * t0);
  const float F1 = c * ((float)li[1] * t0 + (float)li[2] * t1);
  const float F2 = c * ((float)li[3] * t0 + (float)li[4] * t1 + (float)li[5] * t2);
  float* f = a.forces + 3 * (size_t)(i0 + i);
  if (a.accumulate) {
    f[0] += F0; f[1] += F1; f[2] += F2;
  } else {
    f[0] = F0; f[1] = F1; f[2] = F2;
  }
}

}  // namespace kupsmi

using namespace kupsmi;

// Scratch of kupsmi_ewald_recip_noalloc: a file-scope device array (every
// GPU the library runs on holds its own copy until the process ends), found
// per device with hipGetSymbolAddress on the device's first call and used
// through that address.  Every call on one GPU shares the array: calls that
// overlap on a GPU race.
constexpr size_t EK_STATIC_SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[EK_STATIC_SCRATCH_BYTES];

namespace kupsmi {

constexpr int EK_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[EK_MAX_DEVICES] = {};

// The device address of this device's copy of `scratch` (looked up once per
// device, with the device current).
static int static_scratch(void** out) {
  int dev = 0;
  KUPSMI_HIP(hipGetDevice(&dev));
  KUPSMI_REQUIRE(dev >= 0 && dev < EK_MAX_DEVICES, "device index beyond the library's table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    KUPSMI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
    KUPSMI_REQUIRE(p != nullptr, "no device address for the static scratch");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return KUPSMI_OK;
}

// Both entry points: sizes, then the scratch of the context (ctx) or the
// static array (ctx == nullptr; sizes beyond it are rejected), then the
// launches.
static int ewald_recip(kupsmi_ctx* ctx, void* stream, int n_sys, const int32_t* sys_off,
                       int n_atoms, int max_atoms_per_sys, const float* pos,
                       const float* charges, const double* tril, const double* alpha,
                       const double* kmax, int mx, int my, int mz, int accumulate,
                       float* forces, double* energy, double* virial, double* structure_factor) {
  KUPSMI_REQUIRE(n_sys >= 0 && n_atoms >= 0 && max_atoms_per_sys >= 0, "negative size");
  KUPSMI_REQUIRE(mx >= 0 && my >= 0 && mz >= 0 && mx <= EK_NM && my <= EK_NM && mz <= EK_NM,
                 "k grid half widths must be in [0, 12]");
  KUPSMI_REQUIRE(tril && alpha, "cell and alpha required");
  KUPSMI_REQUIRE(n_atoms == 0 || (pos && charges), "positions and charges required");
  if (n_sys == 0) return KUPSMI_OK;
  hipStream_t st = (hipStream_t)stream;
  EkArgs a{};
  a.n_sys = n_sys; a.mx = mx; a.my = my; a.mz = mz;
  a.nk = (mx + 1) * (2 * my + 1) * (2 * mz + 1);
  a.accumulate = accumulate;
  a.sys_off = sys_off; a.pos = pos; a.q = charges; a.tril = tril; a.alpha = alpha; a.kmax = kmax;
  a.forces = forces; a.energy = energy; a.virial = virial;
  const size_t nsk = (size_t)n_sys * a.nk;
  a.chunks = max_atoms_per_sys > 0 ? (max_atoms_per_sys + EK_CHUNK - 1) / EK_CHUNK : 1;
  void* buf = nullptr;
  const size_t sk_bytes = structure_factor ? 0 : ((sizeof(double) * 2 * nsk + 255) & ~(size_t)255);
  const size_t skp_bytes = (sizeof(float2) * nsk + 255) & ~(size_t)255;
  const size_t part_bytes = sizeof(double) * 2 * nsk * (size_t)a.chunks;
  const size_t need = sk_bytes + skp_bytes + part_bytes;
  if (ctx) {
    int rc = scratch(ctx, 4, need, &buf);
    if (rc) return rc;
  } else {
    if (need > EK_STATIC_SCRATCH_BYTES) {
      set_error("kupsmi_ewald_recip_noalloc: the call needs " + std::to_string(need) +
                " bytes of scratch (n_sys x k vectors x (24 + 16 x ceil(max_atoms_per_sys / " +
                std::to_string(EK_CHUNK) + ")) bytes), more than the " +
                std::to_string(EK_STATIC_SCRATCH_BYTES) + " of its static array: " +
                "split the batch or use kupsmi_ewald_recip");
      return KUPSMI_ERR_ARG;
    }
    int rc = static_scratch(&buf);
    if (rc) return rc;
  }
  char* b = static_cast<char*>(buf);
  a.sk = structure_factor ? structure_factor : reinterpret_cast<double*>(b);
  a.skp = reinterpret_cast<float2*>(b + sk_bytes);
  a.skpart = reinterpret_cast<double*>(b + sk_bytes + skp_bytes);
  ek_sk_launch(a, dim3(a.chunks, n_sys), st);
  KUPSMI_CHECK_LAUNCH();
  ek_sk_reduce_kernel<<<dim3((a.nk + 255) / 256, n_sys), 256, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  ek_energy_kernel<<<n_sys, EK_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  if (forces && max_atoms_per_sys > 0) {
    dim3 g2((max_atoms_per_sys + EK_THREADS - 1) / EK_THREADS, n_sys);
    const size_t lds = sizeof(float2) * a.nk;
    if (a.mz == 8) ek_force_kernel<8><<<g2, EK_THREADS, lds, st>>>(a);
    else ek_force_kernel<0><<<g2, EK_THREADS, lds, st>>>(a);
    KUPSMI_CHECK_LAUNCH();
  }
  return KUPSMI_OK;
}

}  // namespace kupsmi

extern "C" int kupsmi_ewald_recip(kupsmi_ctx* ctx, void* stream, int n_sys, const int32_t* sys_off,
                                  int n_atoms, int max_atoms_per_sys, const float* pos,
                                  const float* charges, const double* tril, const double* alpha,
                                  const double* kmax, int mx, int my, int mz, int accumulate,
                                  float* forces, double* energy, double* virial,
                                  double* structure_factor) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  return kupsmi::ewald_recip(ctx, stream, n_sys, sys_off, n_atoms, max_atoms_per_sys, pos,
                             charges, tril, alpha, kmax, mx, my, mz, accumulate, forces, energy,
                             virial, structure_factor);
}

extern "C" int kupsmi_ewald_recip_noalloc(void* stream, int n_sys, const int32_t* sys_off,
                                          int n_atoms, int max_atoms_per_sys, const float* pos,
                                          const float* charges, const double* tril,
                                          const double* alpha, const double* kmax, int mx, int my,
                                          int mz, int accumulate, float* forces, double* energy,
                                          double* virial, double* structure_factor) {
  return kupsmi::ewald_recip(nullptr, stream, n_sys, sys_off, n_atoms, max_atoms_per_sys, pos,
                             charges, tril, alpha, kmax, mx, my, mz, accumulate, forces, energy,
                             virial, structure_factor);
}
