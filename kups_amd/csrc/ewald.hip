// Reciprocal-space Ewald sum (potential/classical/ewald.py:349-439, 585-601).
//
// k = 2 pi Linv n over the half-space integer grid n in [0..mx] x [-my..my] x
// [-mz..mz] (the grid of kvecs_from_kmax, ewald.py:1322-1339, which keeps the
// vectors with |k| <= kmax; kmax < 0 keeps the whole box).  Phases factorise
// per axis, exp(i k.r) = ex(n0) ey(n1) ez(n2) with e_a(n) = exp(2 pi i n f_a)
// on the wrapped fractional coordinates, so the per-atom work is a product of
// three small tables instead of a sincos per (atom, k).
//
//   ek_sk_kernel      S(k) = sum_i q_i exp(i k.r_i) as a GEMM on the FP16
//                     matrix cores with hi / lo split operands (FP32
//                     accumulation, ~22-bit products), one workgroup
//                     per (system, 512-atom chunk), atoms streamed through LDS
//                     in tiles of up to 128 as per-axis phase planes that
//                     both GEMM operands are read from; S(-n2) comes from
//                     the conjugate symmetry of the z factor;
//                     ek_sk_reduce_kernel sums the chunks in fixed order.
//   ek_energy_kernel  P(k) (FP64), E = K sum_k P |S|^2, the reciprocal-space
//                     virial, and S'(k) = P(k) S(k) for the force pass.
//   ek_force_mfma_kernel / ek_force_kernel
//                     F_i = -2 K q_i sum_k k Im(S'(k) exp(-i k.r_i)): the sum
//                     over n_z as an FP16 hi / lo split GEMM on the matrix
//                     cores (z half width <= 8), else one thread per atom on
//                     the vector ALUs.
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
  float* skc;       // n_sys x M x (4 + 4 mz): force coefficients of P(k) S(k)
  float* forces;
  double* energy;
  double* virial;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// FP16 hi / lo split of eight FP32 values (round to nearest each): v - hi is
// exact in FP32, so hi + lo carries ~22 significant bits.
__device__ inline void split8(const float (&v)[8], f16x8& hi, f16x8& lo) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    hi[u] = (_Float16)v[u];
    lo[u] = (_Float16)(v[u] - (float)hi[u]);
  }
}

// v_permlane32_swap on a 16-byte fragment: returns (x in lanes 0..31, y of
// lane l - 32 in lanes 32..63) and (x of lane l + 32 in lanes 0..31, y in
// lanes 32..63).
__device__ inline void swap32(f16x8 x, f16x8 y, f16x8& r0, f16x8& r1) {
  const u32x4 xu = __builtin_bit_cast(u32x4, x), yu = __builtin_bit_cast(u32x4, y);
  u32x4 a, b;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const auto r = __builtin_amdgcn_permlane32_swap(xu[k], yu[k], false, false);
    a[k] = r[0];
    b[k] = r[1];
  }
  r0 = __builtin_bit_cast(f16x8, a);
  r1 = __builtin_bit_cast(f16x8, b);
}

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

// S(k) as a real GEMM on the matrix cores, using the z-axis symmetry
// Z(-n2) = conj Z(n2).  Rows (m, a): m = (n0, n1) over [0..mx] x [-my..my],
// a = re / im of A_i(m) = q_i X_i(n0) Y_i(n1); columns (n2, b): n2 in
// [0..mz], b = re / im of Z_i(n2); K = atoms:
//   C[(m,a)][(n2,b)] = sum_i A_ia(m) Z_ib(n2)
//   S(m, +n2) = (C_rr - C_ii) + i (C_ri + C_ir)
//   S(m, -n2) = (C_rr + C_ii) + i (C_ir - C_ri)
// One workgroup per (system, chunk of EK_CHUNK atoms, group of row tiles).
// Per tile of TA atoms the workgroup writes the per-axis phase factors (FP64
// recurrences, rounded once) to LDS, four atoms per 16-byte row (layout in
// the kernel); the GEMM operands are then read straight from these rows:
// the A value of row (m, a) is xr yr - xi yi (a = re) or xr yi + xi yr
// (a = im), one multiply and one FMA from rows chosen per lane (yr / -yi or
// yi / yr).  K runs in blocks of 32 atoms; both operands are split into FP16
// hi + lo (round to nearest each, ~22 significant bits) and the three
// products Ah Zh + Al Zh + Ah Zl accumulate in FP32 (the dropped Al Zl is
// below 2^-22 of a term); the MFMAs of one product over all (row, column)
// tiles issue back to back, so no accumulator waits on the previous MFMA
// (details at the loop).  D of an MFMA: lane l holds rows 4 (l / 16) + r,
// column l % 16, so rows (m, re) and (m, im) sit in one lane and columns
// (n2, re) and (n2, im) in lanes l and l ^ 1: one exchange gives the even
// lane S(m, +n2) and the odd lane S(m, -n2).  A chunk's sums stay in the
// FP32 accumulators (at most 512 unit-modulus terms times q each); its
// partial S goes to HBM in FP64 for a fixed-order reduce.
constexpr int EK_SK_THREADS = 256;
constexpr int EK_WAVES = EK_SK_THREADS / 64;
constexpr int EK_CHUNK = 512;
constexpr int EK_MTW_MAX = 5;  // row tiles per wave

struct EkSkShape {
  int ta, pitch;  // atoms per tile (multiple of 32), rows per group of four atoms
  int mtw, groups, nt;
  size_t lds;     // dynamic LDS bytes
};

template <int MTW, int NT>
// Two waves per SIMD (a register budget of 256): the unconstrained build
// takes ~300 registers and runs one wave per SIMD with every LDS and matrix-
// core latency exposed (0.80 -> 0.54 ms per launch on the 512-melt batch,
// 44 bytes per lane spilled).
__global__ __launch_bounds__(EK_SK_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2)))
void ek_sk_kernel(EkArgs a, EkSkShape sh) {
  extern __shared__ float ek_lds[];
  __shared__ double s_li[6];
  const int TA = sh.ta, RP = sh.pitch;
  const int wx = a.mx + 1, wy = 2 * a.my + 1, wz = 2 * a.mz + 1, wzh = a.mz + 1;
  const int M = wx * wy;
  // LDS: per group of four atoms RP rows of four floats (the four atoms):
  // rows 2 n0 / 2 n0 + 1 = Re / Im of q X(n0); rows Y0 + 2 j / + 1 = Re / Im
  // of Y(n1) (j = n1 + my); rows Y1 + 2 j / + 1 = -Im / Re of Y(n1); rows
  // ZR + 2 n2 / + 1 = Re / Im of Z(n2).  The 16 lanes of a quarter-wave read
  // 16 consecutive rows of one group (rows (m, re) and (m, im) of 8
  // consecutive m, or 8 n2 x re / im), i.e. 256 contiguous bytes; RP = 1
  // (mod 16) puts the 16 groups that a wave's phase writes touch in 16
  // different bank quads.
  const int Y0 = 2 * wx, Y1 = Y0 + 2 * wy, ZR = Y1 + 2 * wy;
  // system in x: consecutive workgroups go round-robin to the dies, so a
  // system's chunks all run on die s % 8 (its positions in that die's L2)
  // and the dies get systems of every size
  const int s = blockIdx.x, chunk = blockIdx.y, grp = blockIdx.z;
  const int i0 = a.sys_off[s], n_all = a.sys_off[s + 1] - i0;
  const int c0 = chunk * EK_CHUNK;
  if (c0 >= n_all) return;
  const int n = n_all - c0 < EK_CHUNK ? n_all - c0 : EK_CHUNK;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = lane & 15, kq = lane >> 4;
  if (tid == 0) tril_inverse(a.tril + 6 * s, s_li);
  // operand rows of this lane: row tile T = (grp EK_WAVES + wave) MTW + t
  // holds GEMM rows 16 T .. 16 T + 15; rows past 2 M read row M - 1
  // (discarded)
  int pa[MTW], pb[MTW], pc[MTW];
#pragma unroll
  for (int t = 0; t < MTW; ++t) {
    const int R = 16 * ((grp * EK_WAVES + wave) * MTW + t) + row;
    const int m = min(R >> 1, M - 1), im = R & 1;
    const int n0 = m / wy, j1 = m - n0 * wy;
    pa[t] = 2 * n0;
    pb[t] = Y0 + 2 * j1 + im;
    pc[t] = Y1 + 2 * j1 + im;
  }
  int pz[NT];
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int col = 16 * u + row;
    pz[u] = ZR + 2 * min(col >> 1, a.mz) + (col & 1);
  }
  f32x4 acc[MTW][NT];
#pragma unroll
  for (int t = 0; t < MTW; ++t)
#pragma unroll
    for (int u = 0; u < NT; ++u) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int base = 0; base < n; base += TA) {
    __syncthreads();
    const int cnt = n - base < TA ? n - base : TA;
    // 1. per-axis phase rows (q folded into X); atoms past the end are zero.
    // Two threads per atom: one writes X and Z (2 wx + 2 wzh rows), the other
    // Y (4 wy rows, both signs of n1 from one recurrence).
    for (int e = tid; e < 2 * TA; e += EK_SK_THREADS) {
      const int at = e >> 1, part = e & 1;
      float* o = ek_lds + (size_t)(at >> 2) * RP * 4 + (at & 3);  // row r at o[4 r]
      if (at < cnt) {
        const float* r = a.pos + 3 * (size_t)(i0 + c0 + base + at);
        double f0, f1, f2;
        tmul(s_li, (double)r[0], (double)r[1], (double)r[2], f0, f1, f2);
        if (part == 0) {
          f0 -= rint(f0);
          f2 -= rint(f2);
          double2 ux, uz;
          sincospi(2.0 * f0, &ux.y, &ux.x);
          sincospi(2.0 * f2, &uz.y, &uz.x);
          double2 p = make_double2((double)a.q[i0 + c0 + base + at], 0.0);
          for (int k = 0; k <= a.mx; ++k) {
            o[8 * k] = (float)p.x;
            o[8 * k + 4] = (float)p.y;
            p = zmul(p, ux);
          }
          float* oz = o + 4 * ZR;
          p = make_double2(1.0, 0.0);
          for (int k = 0; k <= a.mz; ++k) {
            oz[8 * k] = (float)p.x;
            oz[8 * k + 4] = (float)p.y;
            p = zmul(p, uz);
          }
        } else {
          f1 -= rint(f1);
          double2 u;
          sincospi(2.0 * f1, &u.y, &u.x);
          float* y0 = o + 4 * (Y0 + 2 * a.my);  // row of n1 = 0
          float* y1 = o + 4 * (Y1 + 2 * a.my);
          y0[0] = 1.f; y0[4] = 0.f;
          y1[0] = 0.f; y1[4] = 1.f;
          double2 p = make_double2(1.0, 0.0);
          for (int k = 1; k <= a.my; ++k) {
            p = zmul(p, u);
            const float re = (float)p.x, im = (float)p.y;
            y0[8 * k] = re;  y0[8 * k + 4] = im;
            y1[8 * k] = -im; y1[8 * k + 4] = re;
            y0[-8 * k] = re; y0[-8 * k + 4] = -im;
            y1[-8 * k] = im; y1[-8 * k + 4] = re;
          }
        }
      } else if (part == 0) {
        for (int k = 0; k < 2 * wx; ++k) o[4 * k] = 0.f;
        for (int k = 0; k < 2 * wzh; ++k) o[4 * (ZR + k)] = 0.f;
      } else {
        for (int k = 0; k < 4 * wy; ++k) o[4 * (Y0 + k)] = 0.f;
      }
    }
    __syncthreads();
    // 2. matrix-core products over the tile's atoms, 32 per block, on FP16
    // hi / lo splits of both operands (C += Ah Zh + Al Zh + Ah Zl):
    // v_mfma_f32_16x16x32_f16 takes Ah Zh + Al Zh of 16 atoms, K slot
    // (kq, u) = atom 8 (kq & 1) + u of the half, hi for kq < 2 and lo above;
    // v_mfma_f32_16x16x16_f16 takes Ah Zl of atoms 8 kq + u (u < 4, then
    // u >= 4).  Lane (row, kq) computes A of atoms k0 + 8 kq .. + 7 and one
    // permlane32 swap hands each half of the wave the other half's values.
    for (int k0 = 0; k0 < cnt; k0 += 32) {
      const int g0 = (k0 >> 2) + 2 * kq;  // group of atoms k0 + 8 kq .. + 3
      f16x8 zb1[NT], zb2[NT];
      f16x4 zla[NT], zlb[NT];
#pragma unroll
      for (int u = 0; u < NT; ++u) {
        const f32x4 z0 = *reinterpret_cast<const f32x4*>(ek_lds + (g0 * RP + pz[u]) * 4);
        const f32x4 z1 = *reinterpret_cast<const f32x4*>(ek_lds + ((g0 + 1) * RP + pz[u]) * 4);
        const float zv[8] = {z0[0], z0[1], z0[2], z0[3], z1[0], z1[1], z1[2], z1[3]};
        f16x8 zh, zl;
        split8(zv, zh, zl);
        swap32(zh, zh, zb1[u], zb2[u]);
        zla[u] = f16x4{zl[0], zl[1], zl[2], zl[3]};
        zlb[u] = f16x4{zl[4], zl[5], zl[6], zl[7]};
      }
      f16x8 a1[MTW], a2[MTW];
      f16x4 aha[MTW], ahb[MTW];
#pragma unroll
      for (int t = 0; t < MTW; ++t) {
        float v[8];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const float* gb = ek_lds + (size_t)(g0 + h) * RP * 4;
          const f32x4 x0 = *reinterpret_cast<const f32x4*>(gb + pa[t] * 4);
          const f32x4 x1 = *reinterpret_cast<const f32x4*>(gb + pa[t] * 4 + 4);
          const f32x4 y0 = *reinterpret_cast<const f32x4*>(gb + pb[t] * 4);
          const f32x4 y1 = *reinterpret_cast<const f32x4*>(gb + pc[t] * 4);
#pragma unroll
          for (int j = 0; j < 4; ++j) v[4 * h + j] = fmaf(x1[j], y1[j], x0[j] * y0[j]);
        }
        f16x8 hi, lo;
        split8(v, hi, lo);
        aha[t] = f16x4{hi[0], hi[1], hi[2], hi[3]};
        ahb[t] = f16x4{hi[4], hi[5], hi[6], hi[7]};
        swap32(hi, lo, a1[t], a2[t]);
      }
#pragma unroll
      for (int t = 0; t < MTW; ++t)
#pragma unroll
        for (int u = 0; u < NT; ++u)
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1[t], zb1[u], acc[t][u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < MTW; ++t)
#pragma unroll
        for (int u = 0; u < NT; ++u)
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2[t], zb2[u], acc[t][u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < MTW; ++t)
#pragma unroll
        for (int u = 0; u < NT; ++u)
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x16f16(aha[t], zla[u], acc[t][u], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < MTW; ++t)
#pragma unroll
        for (int u = 0; u < NT; ++u)
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x16f16(ahb[t], zlb[u], acc[t][u], 0, 0, 0);
    }
  }
  // lane l of tile (t, u): acc[r] = C[(m0 + r / 2, r % 2)][(n2, b)] with
  // m0 = 8 T + 2 (l / 16), n2 = 8 u + (l % 16) / 2, b = l % 2
  double* part = a.skpart + ((size_t)s * a.chunks + chunk) * a.nk * 2;
  const int odd = lane & 1;
#pragma unroll
  for (int t = 0; t < MTW; ++t) {
    const int T = (grp * EK_WAVES + wave) * MTW + t;
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      const int n2 = 8 * u + (row >> 1);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float vr = acc[t][u][2 * h], vi = acc[t][u][2 * h + 1];
        const float wr = __shfl_xor(vr, 1), wi = __shfl_xor(vi, 1);
        const int m = 8 * T + 2 * kq + h;
        if (m >= M || n2 > a.mz || (odd && n2 == 0)) continue;
        // even lane: C_rr = vr, C_ir = vi, C_ri = wr, C_ii = wi -> S(m, +n2)
        // odd lane:  C_ri = vr, C_ii = vi, C_rr = wr, C_ir = wi -> S(m, -n2)
        const double re = odd ? (double)wr + vi : (double)vr - wi;
        const double im = odd ? (double)wi - vr : (double)wr + vi;
        const int k = m * wz + a.mz + (odd ? -n2 : n2);
        part[2 * k] = re;
        part[2 * k + 1] = im;
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

// Tile size: the largest TA in {128, 96, 64, 32} whose phase rows fit
// 64 KB; row tiles per wave up to EK_MTW_MAX, in as many groups as needed.
EkSkShape ek_sk_shape(const EkArgs& a) {
  const int wx = a.mx + 1, wy = 2 * a.my + 1, wzh = a.mz + 1;
  const int rows = 2 * wx + 4 * wy + 2 * wzh;
  const int rp = rows + ((1 - rows) % 16 + 16) % 16;  // rows per atom group, = 1 (mod 16)
  EkSkShape sh{};
  sh.nt = (2 * wzh + 15) / 16;
  // multiples of the 32-atom block of the products
  const int tas[4] = {128, 96, 64, 32};
  for (int c = 0; c < 4; ++c) {
    sh.ta = tas[c];
    sh.pitch = rp;
    sh.lds = sizeof(float) * 4 * (size_t)rp * (sh.ta / 4);
    if (sh.lds <= 64 * 1024) break;
  }
  const int rt = (2 * wx * wy + 15) / 16;
  sh.mtw = min(EK_MTW_MAX, (rt + EK_WAVES - 1) / EK_WAVES);
  sh.groups = (rt + EK_WAVES * sh.mtw - 1) / (EK_WAVES * sh.mtw);
  return sh;
}

template <int MTW>
void ek_sk_launch_nt(const EkArgs& a, const EkSkShape& sh, dim3 g, hipStream_t st) {
  if (sh.nt == 1) ek_sk_kernel<MTW, 1><<<g, EK_SK_THREADS, sh.lds, st>>>(a, sh);
  else ek_sk_kernel<MTW, 2><<<g, EK_SK_THREADS, sh.lds, st>>>(a, sh);
}

void ek_sk_launch(const EkArgs& a, int chunks, int n_sys, hipStream_t st) {
  const EkSkShape sh = ek_sk_shape(a);
  const dim3 g(n_sys, chunks, sh.groups);
  switch (sh.mtw) {
    case 1: ek_sk_launch_nt<1>(a, sh, g, st); break;
    case 2: ek_sk_launch_nt<2>(a, sh, g, st); break;
    case 3: ek_sk_launch_nt<3>(a, sh, g, st); break;
    case 4: ek_sk_launch_nt<4>(a, sh, g, st); break;
    default: ek_sk_launch_nt<5>(a, sh, g, st); break;
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
  // S'(k) = P(k) S(k) and the energy / virial of k; returns S'
  auto term = [&](int m, int n2) -> double2 {
    const int n0 = m / wy, n1 = m % wy - a.my;
    const int k = m * wz + a.mz + n2;
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
    return make_double2(P * sk[0], P * sk[1]);
  };
  // per m: S'(m, 0) (x, y, 0, 0), then for n2 = 1..mz the combinations of
  // A = S'(m, +n2), B = S'(m, -n2) that the force pass contracts with the
  // phase products (ek_force_kernel)
  const int stride = 4 + 4 * a.mz;
  float* out = a.skc + (size_t)s * (a.mx + 1) * wy * stride;
  for (int it = tid; it < (a.mx + 1) * wy * (a.mz + 1); it += EK_THREADS) {
    const int m = it / (a.mz + 1), n2 = it % (a.mz + 1);
    float* o = out + (size_t)m * stride;
    if (n2 == 0) {
      const double2 A = term(m, 0);
      o[0] = (float)A.x; o[1] = (float)A.y; o[2] = 0.f; o[3] = 0.f;
    } else {
      const double2 A = term(m, n2), B = term(m, -n2);
      float* c = o + 4 * n2;
      c[0] = (float)(A.y + B.y); c[1] = (float)(B.y - A.y);
      c[2] = (float)(B.x - A.x); c[3] = (float)(-(A.x + B.x));
    }
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

// F_i = -2 K q_i sum_k k Im(S'(k) conj(e_i(k))), e_i = X(n0) Y(n1) Z(n2):
// one thread per atom.  The +n2 and -n2 terms of a row (n0, n1) share the
// products P1 = xr zr, P2 = xi zi, P3 = xr zi, P4 = xi zr of the row phase
// x = X Y and Z(n2) (e(+n2) = (P1 - P2, P3 + P4), e(-n2) = (P1 + P2, P4 - P3)),
// so with A = S'(+n2), B = S'(-n2) and c = (Ay + By, By - Ay, Bx - Ax,
// -(Ax + Bx)):
//   v+ + v- = P1 c0 + P2 c1 + P3 c2 + P4 c3
//   v+ - v- = -P1 c1 - P2 c0 + P3 c3 + P4 c2
// with v = Im(S' conj(e)); the four coefficients per (row, n2) come from
// ek_energy_kernel.  MZ > 0: the z half width as a constant (Z(n2) kept in
// registers, loop unrolled); MZ = 0: any width, Z(n2) by recurrence per
// row.  LDS: the coefficients staged per workgroup; otherwise read from
// global memory (grids whose coefficients exceed 64 KB).
template <int MZ, bool LDS>
__global__ __launch_bounds__(EK_THREADS) void ek_force_kernel(EkArgs a) {
  extern __shared__ float4 skc_lds[];
  const int s = blockIdx.y;
  const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
  const int tid = threadIdx.x;
  const int start = blockIdx.x * EK_THREADS;
  if (start >= n) return;
  const int mz = MZ > 0 ? MZ : a.mz;
  const int wy = 2 * a.my + 1;
  const int stride4 = 1 + mz;  // float4 per row
  const size_t rows = (size_t)(a.mx + 1) * wy;
  const float4* __restrict__ gsrc =
      reinterpret_cast<const float4*>(a.skc) + (size_t)s * rows * stride4;
  const float4* __restrict__ skc = gsrc;
  if (LDS) {
    for (size_t k = tid; k < rows * stride4; k += EK_THREADS) skc_lds[k] = gsrc[k];
    __syncthreads();
    skc = skc_lds;
  }
  const int i = start + tid;
  if (i >= n) return;
  const double* t = a.tril + 6 * s;
  float2 ux, uy, uz;
  unit_phases(t, a.pos + 3 * (size_t)(i0 + i), ux, uy, uz);
  float2 zk[MZ > 0 ? MZ : 1];
  if (MZ > 0) {
    zk[0] = uz;
#pragma unroll
    for (int k = 1; k < (MZ > 0 ? MZ : 1); ++k) zk[k] = cmul(zk[k - 1], uz);
  }
  // u^-my starting point
  float2 uy0 = make_float2(1.f, 0.f);
  const float2 cy = make_float2(uy.x, -uy.y);
  for (int m = 0; m < a.my; ++m) uy0 = cmul(uy0, cy);
  float t0 = 0.f, t1 = 0.f, t2 = 0.f;
  float2 ex = make_float2(1.f, 0.f);
  const float4* __restrict__ row = skc;
  for (int n0 = 0; n0 <= a.mx; ++n0) {
    float2 x = cmul(ex, uy0);
    for (int j1 = 0; j1 < wy; ++j1) {
      const float n1 = (float)(j1 - a.my);
      const float4 s0 = row[0];
      float a0 = s0.y * x.x - s0.x * x.y;
      float a2 = 0.f;
      auto pair = [&](float2 z, int k) {
        const float4 c = row[1 + k];
        const float P1 = x.x * z.x, P2 = x.y * z.y, P3 = x.x * z.y, P4 = x.y * z.x;
        a0 = fmaf(P1, c.x, a0); a0 = fmaf(P2, c.y, a0);
        a0 = fmaf(P3, c.z, a0); a0 = fmaf(P4, c.w, a0);
        float dm = P3 * c.w;
        dm = fmaf(P4, c.z, dm); dm = fmaf(-P1, c.y, dm); dm = fmaf(-P2, c.x, dm);
        a2 = fmaf((float)(k + 1), dm, a2);
      };
      if (MZ > 0) {
#pragma unroll
        for (int k = 0; k < (MZ > 0 ? MZ : 1); ++k) pair(zk[k], k);
      } else {
        float2 z = uz;
        for (int k = 0; k < mz; ++k) {
          pair(z, k);
          z = cmul(z, uz);
        }
      }
      row += stride4;
      t0 += (float)n0 * a0;
      t1 += n1 * a0;
      t2 += a2;
      x = cmul(x, uy);
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

// ------------------------------------------- force pass on the matrix cores
// The sums over n_z of ek_force_kernel, rewritten per row m = (n0, n1) as
//   a0(m) = xr G0(m) + xi G1(m),   a2(m) = xr G2(m) + xi G3(m)
// (x = X(n0) Y(n1) the row phase of the atom) with
//   G0 = S'y(m,0) + sum_k zr_k c0 + zi_k c2     G1 = -S'x(m,0) + sum_k zi_k c1 + zr_k c3
//   G2 = sum_k k (zi_k c3 - zr_k c1)            G3 = sum_k k (zr_k c2 - zi_k c0)
// (z_k = Z(k) of the atom, c the four coefficients of (m, k), k = 1..mz), are
// a GEMM: G[(m, c)][atom] = sum_f B_f(m, c) phi_f(atom) over the 16 features
// phi = (zr_1, zi_1, ..., zr_8, zi_8) plus the atom-independent n_z = 0 term.
// It runs on v_mfma_f32_16x16x32_f16 + v_mfma_f32_16x16x16_f16 with FP16
// operands split in two (v = hi + lo, each FP16): the three products
// Bhi phi_hi + Bhi phi_lo + Blo phi_hi (K = 32 + 16) keep ~22 bits, the
// dropped Blo phi_lo is below 2^-22 of the row's largest coefficient.  Each
// row (m, c) is scaled by a power of two so its largest |B| lies in
// [0.5, 1) (FP16 range) and unscaled on the accumulator.  Rows of D are the
// 4M rows (m, c) of the system's operand, columns 16 atoms: lane l holds
// G_0..3(m = 4 t + l / 16) of atom l % 16 of row tile t, the epilogue
// contracts them with the row phase into the same t0, t1, t2 sums as
// ek_force_kernel, and the four lanes of an atom sum over l / 16 at the end.
constexpr int EF_TILE_BYTES = 1024;  // 16 rows: Bhi fragments (512 B) + Blo fragments (512 B)
constexpr int EF_MREC = 32;          // per m: 4 unscale factors, S'(m,0) term (2), n0, n1
constexpr int EF_WAVES = 4;
constexpr int EF_NA = 2;             // 16-atom tiles per wave
constexpr int EF_CHUNKS = 4;         // 128-atom chunks per workgroup (grid stride)
constexpr int EF_MAX_MZ = 8;         // 2 mz features fit the K = 16 of one feature set
constexpr size_t EF_MAX_LDS = 150 * 1024;

struct EfShape {
  int M, mpad, tiles;
  size_t block_bytes;  // per system: tiles x EF_TILE_BYTES + mpad x EF_MREC
};

inline EfShape ef_shape(int mx, int my) {
  EfShape s;
  s.M = (mx + 1) * (2 * my + 1);
  s.tiles = (4 * s.M + 15) / 16;
  s.mpad = 4 * s.tiles;
  s.block_bytes = (size_t)s.tiles * EF_TILE_BYTES + (size_t)s.mpad * EF_MREC;
  return s;
}

inline size_t ef_lds_bytes(const EfShape& s) { return s.block_bytes; }

// One thread per (system, m): the 16 x 4 coefficients B_f(m, c) from the
// energy kernel's c, scaled per row (m, c), split into FP16 hi / lo and
// stored in the fragment order of the MFMA A operand: row r = 4 m + c of tile
// t = r / 16 at i = r % 16; hi of features 8 h .. 8 h + 7 (h = 0, 1) at
// t * 1024 + 256 h + 16 i, lo of features 4 p .. 4 p + 3 (p = 0..3) at
// t * 1024 + 512 + 128 p + 8 i; the per-m record after the tiles.
__global__ void ek_coef_kernel(EkArgs a, EfShape sh, unsigned char* blocks) {
  const int s = blockIdx.y;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= sh.mpad) return;
  unsigned char* blk = blocks + (size_t)s * sh.block_bytes;
  float B[4][16];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int f = 0; f < 16; ++f) B[c][f] = 0.f;
  float rec[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (m < sh.M) {
    const int wy = 2 * a.my + 1;
    const float* o = a.skc + ((size_t)s * sh.M + m) * (4 + 4 * a.mz);
    rec[4] = o[1];
    rec[5] = -o[0];
    rec[6] = (float)(m / wy);
    rec[7] = (float)(m % wy - a.my);
#pragma unroll
    for (int k = 1; k <= EF_MAX_MZ; ++k) {
      if (k <= a.mz) {
        const float* c = o + 4 * k;
        const float fk = (float)k;
        const int fr = 2 * (k - 1), fi = fr + 1;
        B[0][fr] = c[0]; B[1][fr] = c[3]; B[2][fr] = -fk * c[1]; B[3][fr] = fk * c[2];
        B[0][fi] = c[2]; B[1][fi] = c[1]; B[2][fi] = fk * c[3]; B[3][fi] = -fk * c[0];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float mxv = 0.f;
#pragma unroll
    for (int f = 0; f < 16; ++f) mxv = fmaxf(mxv, fabsf(B[c][f]));
    int e = 0;
    if (mxv > 0.f) frexpf(mxv, &e);
    rec[c] = m < sh.M ? ldexpf(1.0f, e) : 0.f;
    _Float16 hi[16], lo[16];
#pragma unroll
    for (int f = 0; f < 16; ++f) {
      const float v = ldexpf(B[c][f], -e);
      hi[f] = (_Float16)v;
      lo[f] = (_Float16)(v - (float)hi[f]);
    }
    const int r = 4 * m + c, t = r >> 4, i = r & 15;
    unsigned char* tb = blk + (size_t)t * EF_TILE_BYTES;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      f16x8 v;
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = hi[8 * h + u];
      *reinterpret_cast<f16x8*>(tb + 256 * h + 16 * i) = v;
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      f16x4 v;
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = lo[4 * p + u];
      *reinterpret_cast<f16x4*>(tb + 512 + 128 * p + 8 * i) = v;
    }
  }
  float4* r4 = reinterpret_cast<float4*>(blk + (size_t)sh.tiles * EF_TILE_BYTES +
                                         (size_t)m * EF_MREC);
  r4[0] = make_float4(rec[0], rec[1], rec[2], rec[3]);
  r4[1] = make_float4(rec[4], rec[5], rec[6], rec[7]);
}

__device__ inline double2 zconj(double2 a) { return make_double2(a.x, -a.y); }

// b^e for e >= 0 by squaring (at most 2 log2(e) complex products).
__device__ inline double2 zpow(double2 b, int e) {
  double2 r = make_double2(1.0, 0.0);
  while (e > 0) {
    if (e & 1) r = zmul(r, b);
    b = zmul(b, b);
    e >>= 1;
  }
  return r;
}

// Workgroups of EF_WAVES waves, each wave EF_NA tiles of 16 atoms; workgroup
// (b, s) stages system s's operand block in LDS once and takes the 128-atom
// chunks b, b + gridDim.x, ... of the system.  Atom side of the MFMAs (B
// operand, lane l: atom l % 16, k = 8 (l / 16) ..): the x32 set is phi_hi of
// features 1..8 / 9..16 for l / 16 = 0 / 1 and phi_lo of the same for 2 / 3
// (both against Bhi, whose fragments lanes 32..63 re-read); the x16 set is
// phi_hi of features 4 p + 1 .. 4 p + 4 (p = l / 16) against Blo.  Lane
// (atom, q) meets rows m = q, q + 4, q + 8, ... in order, so its row phase
// x(m) = X(n0) Y(n1) follows an FP64 recurrence: x Y^4 within a row of n0,
// x X Y^(4 - (2 my + 1)) when n1 passes my (one wrap per step for my >= 2),
// seeded from the FP32 unit phases promoted to FP64 and rounded to FP32 once
// per use.  The MFMAs of tile t + 1 are issued before the epilogue of tile t.
__global__ __launch_bounds__(EF_WAVES * 64) void ek_force_mfma_kernel(EkArgs a, EfShape sh,
                                                                      const unsigned char* blocks) {
  extern __shared__ __align__(16) unsigned char ef_lds[];
  constexpr int PER = EF_WAVES * EF_NA * 16;
  const int s = blockIdx.y;
  const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
  if ((int)blockIdx.x * PER >= n) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int j = lane & 15, q = lane >> 4;
  {
    const float4* src = reinterpret_cast<const float4*>(blocks + (size_t)s * sh.block_bytes);
    float4* dst = reinterpret_cast<float4*>(ef_lds);
    const int nv = (int)(sh.block_bytes / 16);
    for (int k = tid; k < nv; k += EF_WAVES * 64) dst[k] = src[k];
  }
  __syncthreads();
  const double* tr = a.tril + 6 * s;
  double li[6];
  tril_inverse(tr, li);
  const int my = a.my, wy = 2 * a.my + 1;
  const float* rec = reinterpret_cast<const float*>(ef_lds + (size_t)sh.tiles * EF_TILE_BYTES);
  const float c0 = -2.0f * 14.399645351950548f * 6.283185307179586f;
  for (int start = blockIdx.x * PER; start < n; start += gridDim.x * PER) {
    f16x8 fa[EF_NA];
    f16x4 fb[EF_NA];
    double2 xd[EF_NA], y4[EF_NA], wr[EF_NA];
#pragma unroll
    for (int at = 0; at < EF_NA; ++at) {
      const int i = start + (w * EF_NA + at) * 16 + j;
      float2 ux = make_float2(1.f, 0.f), uy = ux, uz = ux;
      if (i < n) {
        const float* r = a.pos + 3 * (size_t)(i0 + i);
        float f0, f1, f2;
        tmul(li, r[0], r[1], r[2], f0, f1, f2);
        f0 = wrap_half(f0); f1 = wrap_half(f1); f2 = wrap_half(f2);
        sincospif(2.0f * f0, &ux.y, &ux.x);
        sincospif(2.0f * f1, &uy.y, &uy.x);
        sincospif(2.0f * f2, &uz.y, &uz.x);
      }
      const float2 z2 = cmul(uz, uz), z3 = cmul(z2, uz), z4 = cmul(z2, z2);
      const float2 z5 = cmul(z4, uz), z7 = cmul(z5, z2);
      // x32 set: Z(kb + 1 .. kb + 4), kb = 4 (q & 1); hi for q < 2, lo above
      float2 zc = (q & 1) ? z5 : uz;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const _Float16 hr = (_Float16)zc.x, hi = (_Float16)zc.y;
        if (q < 2) {
          fa[at][2 * u] = hr;
          fa[at][2 * u + 1] = hi;
        } else {
          fa[at][2 * u] = (_Float16)(zc.x - (float)hr);
          fa[at][2 * u + 1] = (_Float16)(zc.y - (float)hi);
        }
        zc = cmul(zc, uz);
      }
      // x16 set: Z(2q + 1), Z(2q + 2), hi
      float2 zb = q == 0 ? uz : (q == 1 ? z3 : (q == 2 ? z5 : z7));
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        fb[at][2 * u] = (_Float16)zb.x;
        fb[at][2 * u + 1] = (_Float16)zb.y;
        zb = cmul(zb, uz);
      }
      // row phase of m = q: Y^(q - my); steps Y^4 and X Y^(4 - wy)
      const double2 yd = make_double2((double)uy.x, (double)uy.y);
      const double2 xdu = make_double2((double)ux.x, (double)ux.y);
      const int e0 = q - my;
      xd[at] = zpow(e0 < 0 ? zconj(yd) : yd, e0 < 0 ? -e0 : e0);
      const double2 y2 = zmul(yd, yd);
      y4[at] = zmul(y2, y2);
      wr[at] = zmul(xdu, zpow(zconj(yd), wy - 4));
    }
    float t0[EF_NA], t1[EF_NA], t2[EF_NA];
#pragma unroll
    for (int at = 0; at < EF_NA; ++at) t0[at] = t1[at] = t2[at] = 0.f;
    f32x4 acc[EF_NA];
    {
      const f16x8 sa = *reinterpret_cast<const f16x8*>(ef_lds + 256 * (q & 1) + 16 * j);
      const f16x4 sb = *reinterpret_cast<const f16x4*>(ef_lds + 512 + 128 * q + 8 * j);
#pragma unroll
      for (int at = 0; at < EF_NA; ++at) {
        acc[at] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sa, fa[at], f32x4{0.f, 0.f, 0.f, 0.f},
                                                         0, 0, 0);
        acc[at] = __builtin_amdgcn_mfma_f32_16x16x16f16(sb, fb[at], acc[at], 0, 0, 0);
      }
    }
    int n1c = q - my;
    for (int tt = 0; tt < sh.tiles; ++tt) {
      f32x4 nxt[EF_NA];
      if (tt + 1 < sh.tiles) {
        const unsigned char* tb = ef_lds + (size_t)(tt + 1) * EF_TILE_BYTES;
        const f16x8 sa = *reinterpret_cast<const f16x8*>(tb + 256 * (q & 1) + 16 * j);
        const f16x4 sb = *reinterpret_cast<const f16x4*>(tb + 512 + 128 * q + 8 * j);
#pragma unroll
        for (int at = 0; at < EF_NA; ++at) {
          nxt[at] = __builtin_amdgcn_mfma_f32_16x16x32_f16(sa, fa[at], f32x4{0.f, 0.f, 0.f, 0.f},
                                                           0, 0, 0);
          nxt[at] = __builtin_amdgcn_mfma_f32_16x16x16f16(sb, fb[at], nxt[at], 0, 0, 0);
        }
      }
      const int m = 4 * tt + q;
      const float4 r0 = *reinterpret_cast<const float4*>(rec + 8 * m);
      const float4 r1 = *reinterpret_cast<const float4*>(rec + 8 * m + 4);
      const bool wrap = n1c + 4 > my;
#pragma unroll
      for (int at = 0; at < EF_NA; ++at) {
        const float xr = (float)xd[at].x, xi = (float)xd[at].y;
        const float g0 = fmaf(acc[at][0], r0.x, r1.x), g1 = fmaf(acc[at][1], r0.y, r1.y);
        const float g2 = acc[at][2] * r0.z, g3 = acc[at][3] * r0.w;
        const float a0 = fmaf(xr, g0, xi * g1), a2 = fmaf(xr, g2, xi * g3);
        t0[at] = fmaf(r1.z, a0, t0[at]);
        t1[at] = fmaf(r1.w, a0, t1[at]);
        t2[at] += a2;
        xd[at] = zmul(xd[at], wrap ? wr[at] : y4[at]);
        acc[at] = nxt[at];
      }
      n1c = wrap ? n1c + 4 - wy : n1c + 4;
    }
#pragma unroll
    for (int at = 0; at < EF_NA; ++at) {
      float u0 = t0[at], u1 = t1[at], u2 = t2[at];
      u0 += __shfl_xor(u0, 16, 64); u1 += __shfl_xor(u1, 16, 64); u2 += __shfl_xor(u2, 16, 64);
      u0 += __shfl_xor(u0, 32, 64); u1 += __shfl_xor(u1, 32, 64); u2 += __shfl_xor(u2, 32, 64);
      const int i = start + (w * EF_NA + at) * 16 + j;
      if (q != 0 || i >= n) continue;
      const float c = c0 * a.q[i0 + i];
      const float F0 = c * ((float)li[0] * u0);
      const float F1 = c * ((float)li[1] * u0 + (float)li[2] * u1);
      const float F2 = c * ((float)li[3] * u0 + (float)li[4] * u1 + (float)li[5] * u2);
      float* f = a.forces + 3 * (size_t)(i0 + i);
      if (a.accumulate) {
        f[0] += F0; f[1] += F1; f[2] += F2;
      } else {
        f[0] = F0; f[1] = F1; f[2] = F2;
      }
    }
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
  const size_t coef_floats = (size_t)(mx + 1) * (2 * my + 1) * (4 + 4 * mz);
  const size_t skp_bytes = (sizeof(float) * coef_floats * n_sys + 255) & ~(size_t)255;
  const size_t part_bytes = (sizeof(double) * 2 * nsk * (size_t)a.chunks + 255) & ~(size_t)255;
  // the force pass runs on the matrix cores when the z half width is at most
  // EF_MAX_MZ and the system's operand block fits LDS, else on the VALU
  const EfShape ef = ef_shape(mx, my);
  const bool ef_mfma = forces && max_atoms_per_sys > 0 && mz <= EF_MAX_MZ && my >= 2 &&
                       ef_lds_bytes(ef) <= EF_MAX_LDS;
  const size_t ef_bytes = ef_mfma ? ef.block_bytes * (size_t)n_sys : 0;
  const size_t need = sk_bytes + skp_bytes + part_bytes + ef_bytes;
  if (ctx) {
    int rc = scratch(ctx, 4, need, &buf);
    if (rc) return rc;
  } else {
    if (need > EK_STATIC_SCRATCH_BYTES) {
      set_error("kupsmi_ewald_recip_noalloc: the call needs " + std::to_string(need) +
                " bytes of scratch (n_sys x k vectors x (24 + 16 x ceil(max_atoms_per_sys / " +
                std::to_string(EK_CHUNK) + ")) bytes, plus n_sys x " +
                std::to_string(ef_mfma ? ef.block_bytes : 0) +
                " bytes of force-pass operands), more than the " +
                std::to_string(EK_STATIC_SCRATCH_BYTES) + " of its static array: " +
                "split the batch or use kupsmi_ewald_recip");
      return KUPSMI_ERR_ARG;
    }
    int rc = static_scratch(&buf);
    if (rc) return rc;
  }
  char* b = static_cast<char*>(buf);
  a.sk = structure_factor ? structure_factor : reinterpret_cast<double*>(b);
  a.skc = reinterpret_cast<float*>(b + sk_bytes);
  a.skpart = reinterpret_cast<double*>(b + sk_bytes + skp_bytes);
  unsigned char* ef_blocks = reinterpret_cast<unsigned char*>(b + sk_bytes + skp_bytes + part_bytes);
  ek_sk_launch(a, a.chunks, n_sys, st);
  KUPSMI_CHECK_LAUNCH();
  ek_sk_reduce_kernel<<<dim3((a.nk + 255) / 256, n_sys), 256, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  // energies, virial and the force coefficients P(k) S(k) of the force pass
  ek_energy_kernel<<<n_sys, EK_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  // the phase event: all but the force pass is done (recorded just before
  // it, so work gated on it starts beside the force pass, and the small
  // kernels before it run on an idle GPU)
  auto phase = [&]() -> int {
    if (ctx && ctx->phase) KUPSMI_HIP(hipEventRecord(ctx->phase, st));
    return KUPSMI_OK;
  };
  if (ef_mfma) {
    const size_t lds = ef_lds_bytes(ef);
    static std::once_flag lds_attr;
    std::call_once(lds_attr, [] {
      (void)hipFuncSetAttribute((const void*)ek_force_mfma_kernel,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)EF_MAX_LDS);
    });
    ek_coef_kernel<<<dim3((ef.mpad + 127) / 128, n_sys), 128, 0, st>>>(a, ef, ef_blocks);
    KUPSMI_CHECK_LAUNCH();
    if (int rc = phase()) return rc;
    const int per_wg = EF_WAVES * EF_NA * 16 * EF_CHUNKS;
    ek_force_mfma_kernel<<<dim3((max_atoms_per_sys + per_wg - 1) / per_wg, n_sys),
                           EF_WAVES * 64, lds, st>>>(a, ef, ef_blocks);
    KUPSMI_CHECK_LAUNCH();
  } else if (forces && max_atoms_per_sys > 0) {
    if (int rc = phase()) return rc;
    dim3 g2((max_atoms_per_sys + EK_THREADS - 1) / EK_THREADS, n_sys);
    const size_t lds = sizeof(float) * coef_floats;
    if (lds <= 64 * 1024) {
      if (a.mz == 8) ek_force_kernel<8, true><<<g2, EK_THREADS, lds, st>>>(a);
      else ek_force_kernel<0, true><<<g2, EK_THREADS, lds, st>>>(a);
    } else {
      ek_force_kernel<0, false><<<g2, EK_THREADS, 0, st>>>(a);
    }
    KUPSMI_CHECK_LAUNCH();
  } else if (int rc = phase()) {
    return rc;
  }
  return KUPSMI_OK;
}

}  // namespace kupsmi

extern "C" int kupsmi_ewald_phase_event(kupsmi_ctx* ctx, void** event) {
  KUPSMI_REQUIRE(ctx != nullptr && event != nullptr, "null context or output");
  *event = (void*)ctx->phase;
  return KUPSMI_OK;
}

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
