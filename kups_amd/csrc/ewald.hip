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
//                     in tiles of up to 128 as per-axis phase planes that
//                     both GEMM operands are read from; S(-n2) comes from
//                     the conjugate symmetry of the z factor;
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
  float* skc;       // n_sys x M x (4 + 4 mz): force coefficients of P(k) S(k)
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

// S(k) as a real FP32 GEMM on the matrix cores, using the z-axis symmetry
// Z(-n2) = conj Z(n2).  Rows (m, a): m = (n0, n1) over [0..mx] x [-my..my],
// a = re / im of A_i(m) = q_i X_i(n0) Y_i(n1); columns (n2, b): n2 in
// [0..mz], b = re / im of Z_i(n2); K = atoms:
//   C[(m,a)][(n2,b)] = sum_i A_ia(m) Z_ib(n2)
//   S(m, +n2) = (C_rr - C_ii) + i (C_ri + C_ir)
//   S(m, -n2) = (C_rr + C_ii) + i (C_ir - C_ri)
// One workgroup per (system, chunk of EK_CHUNK atoms, group of row tiles).
// Per tile of TA atoms the workgroup writes the per-axis phase factors (FP64
// recurrences, rounded once) to LDS as float planes [n][atom] with a row pitch
// of TA + 4 floats; the GEMM operands are then read straight from the planes:
// the A value of row (m, a) is xr yr - xi yi (a = re) or xr yi + xi yr
// (a = im), one multiply and one FMA from planes chosen per lane (yr / -yi or
// yi / yr), so neither operand is packed.  K runs in blocks of 16 atoms: lane
// (r = l % 16, q = l / 16) reads float4s of atoms k0 + 4q .. k0 + 4q + 3 and
// component j feeds the j-th v_mfma_f32_16x16x4_f32 (each MFMA sums
// k0 + 4q + j over q; the four together cover the block); the MFMAs of one j
// over all (row, column) tiles issue back to back, so no accumulator waits
// on the previous MFMA.  D of an MFMA: lane l holds rows 4 (l / 16) + r,
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
  int ta, pitch;  // atoms per tile (multiple of 16), plane row pitch (floats)
  int mtw, groups, nt;
  size_t lds;     // dynamic LDS bytes
};

template <int MTW, int NT>
__global__ __launch_bounds__(EK_SK_THREADS) void ek_sk_kernel(EkArgs a, EkSkShape sh) {
  extern __shared__ float ek_lds[];
  __shared__ double s_li[6];
  const int TA = sh.ta, P = sh.pitch;
  const int wx = a.mx + 1, wy = 2 * a.my + 1, wz = 2 * a.mz + 1, wzh = a.mz + 1;
  const int M = wx * wy;
  float* xr = ek_lds;                  // [wx][P]  Re q X(n0)
  float* xi = xr + (size_t)wx * P;     // [wx][P]  Im q X(n0)
  float* yr = xi + (size_t)wx * P;     // [wy][P]  Re Y(n1), n1 = -my..my
  float* yi = yr + (size_t)wy * P;     // [wy][P]  Im Y
  float* nyi = yi + (size_t)wy * P;    // [wy][P]  -Im Y
  float* zr = nyi + (size_t)wy * P;    // [wzh][P] Re Z(n2), n2 = 0..mz
  float* zi = zr + (size_t)wzh * P;    // [wzh][P] Im Z
  const int s = blockIdx.y, chunk = blockIdx.x, grp = blockIdx.z;
  const int i0 = a.sys_off[s], n_all = a.sys_off[s + 1] - i0;
  const int c0 = chunk * EK_CHUNK;
  if (c0 >= n_all) return;
  const int n = n_all - c0 < EK_CHUNK ? n_all - c0 : EK_CHUNK;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int row = lane & 15, kq = lane >> 4;
  if (tid == 0) tril_inverse(a.tril + 6 * s, s_li);
  // operand planes of this lane: row tile T = (grp EK_WAVES + wave) MTW + t
  // holds rows 16 T .. 16 T + 15; rows past 2 M read row M - 1 (discarded)
  // (float offsets into the LDS planes)
  int pa[MTW], pb[MTW], pc[MTW];
#pragma unroll
  for (int t = 0; t < MTW; ++t) {
    const int R = 16 * ((grp * EK_WAVES + wave) * MTW + t) + row;
    const int m = min(R >> 1, M - 1), im = R & 1;
    const int n0 = m / wy, j1 = m - n0 * wy;
    pa[t] = n0 * P;
    pb[t] = (int)((im ? yi : yr) - ek_lds) + j1 * P;
    pc[t] = (int)((im ? yr : nyi) - ek_lds) + j1 * P;
  }
  int pz[NT];
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const int col = 16 * u + row;
    const int n2 = min(col >> 1, a.mz);
    pz[u] = (int)(((col & 1) ? zi : zr) - ek_lds) + n2 * P;
  }
  const int xoff = wx * P;  // xr -> xi
  f32x4 acc[MTW][NT];
#pragma unroll
  for (int t = 0; t < MTW; ++t)
#pragma unroll
    for (int u = 0; u < NT; ++u) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int base = 0; base < n; base += TA) {
    __syncthreads();
    const int cnt = n - base < TA ? n - base : TA;
    // 1. per-axis phase planes (q folded into X); atoms past the end are zero
    for (int e = tid; e < 3 * TA; e += EK_SK_THREADS) {
      const int axis = e / TA, at = e - axis * TA;
      if (at < cnt) {
        const float* r = a.pos + 3 * (size_t)(i0 + c0 + base + at);
        double f0, f1, f2;
        tmul(s_li, (double)r[0], (double)r[1], (double)r[2], f0, f1, f2);
        double f = axis == 0 ? f0 : (axis == 1 ? f1 : f2);
        f -= rint(f);
        double2 u;
        sincospi(2.0 * f, &u.y, &u.x);
        if (axis == 0) {
          double2 p = make_double2((double)a.q[i0 + c0 + base + at], 0.0);
          for (int k = 0; k <= a.mx; ++k) {
            xr[(size_t)k * P + at] = (float)p.x;
            xi[(size_t)k * P + at] = (float)p.y;
            p = zmul(p, u);
          }
        } else if (axis == 1) {
          const int c = a.my;
          yr[(size_t)c * P + at] = 1.f;
          yi[(size_t)c * P + at] = 0.f;
          nyi[(size_t)c * P + at] = 0.f;
          double2 p = make_double2(1.0, 0.0);
          for (int k = 1; k <= a.my; ++k) {
            p = zmul(p, u);
            const float re = (float)p.x, im = (float)p.y;
            yr[(size_t)(c + k) * P + at] = re;
            yi[(size_t)(c + k) * P + at] = im;
            nyi[(size_t)(c + k) * P + at] = -im;
            yr[(size_t)(c - k) * P + at] = re;
            yi[(size_t)(c - k) * P + at] = -im;
            nyi[(size_t)(c - k) * P + at] = im;
          }
        } else {
          double2 p = make_double2(1.0, 0.0);
          for (int k = 0; k <= a.mz; ++k) {
            zr[(size_t)k * P + at] = (float)p.x;
            zi[(size_t)k * P + at] = (float)p.y;
            p = zmul(p, u);
          }
        }
      } else if (axis == 0) {
        for (int k = 0; k <= a.mx; ++k) xr[(size_t)k * P + at] = xi[(size_t)k * P + at] = 0.f;
      } else if (axis == 1) {
        for (int k = 0; k < wy; ++k)
          yr[(size_t)k * P + at] = yi[(size_t)k * P + at] = nyi[(size_t)k * P + at] = 0.f;
      } else {
        for (int k = 0; k <= a.mz; ++k) zr[(size_t)k * P + at] = zi[(size_t)k * P + at] = 0.f;
      }
    }
    __syncthreads();
    // 2. matrix-core products over the tile's atoms, 16 per block
    for (int k0 = 0; k0 < cnt; k0 += 16) {
      const int at = k0 + 4 * kq;
      f32x4 bv[NT];
#pragma unroll
      for (int u = 0; u < NT; ++u) bv[u] = *reinterpret_cast<const f32x4*>(ek_lds + pz[u] + at);
      f32x4 av[MTW];
#pragma unroll
      for (int t = 0; t < MTW; ++t) {
        const f32x4 x0 = *reinterpret_cast<const f32x4*>(ek_lds + pa[t] + at);
        const f32x4 x1 = *reinterpret_cast<const f32x4*>(ek_lds + pa[t] + xoff + at);
        const f32x4 y0 = *reinterpret_cast<const f32x4*>(ek_lds + pb[t] + at);
        const f32x4 y1 = *reinterpret_cast<const f32x4*>(ek_lds + pc[t] + at);
#pragma unroll
        for (int j = 0; j < 4; ++j) av[t][j] = fmaf(x1[j], y1[j], x0[j] * y0[j]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < MTW; ++t)
#pragma unroll
          for (int u = 0; u < NT; ++u)
            acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[t][j], bv[u][j], acc[t][u], 0, 0, 0);
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

// Tile size: the largest TA in {128, 96, 64, 48, 32, 16} whose planes fit
// 64 KB; row tiles per wave up to EK_MTW_MAX, in as many groups as needed.
EkSkShape ek_sk_shape(const EkArgs& a) {
  const int wx = a.mx + 1, wy = 2 * a.my + 1, wzh = a.mz + 1;
  const int planes = 2 * wx + 3 * wy + 2 * wzh;
  EkSkShape sh{};
  const int tas[6] = {128, 96, 64, 48, 32, 16};
  for (int c = 0; c < 6; ++c) {
    sh.ta = tas[c];
    sh.pitch = sh.ta + 4;
    sh.lds = sizeof(float) * (size_t)planes * sh.pitch;
    if (sh.lds <= 64 * 1024) break;
  }
  const int rt = (2 * wx * wy + 15) / 16;
  sh.mtw = min(EK_MTW_MAX, (rt + EK_WAVES - 1) / EK_WAVES);
  sh.groups = (rt + EK_WAVES * sh.mtw - 1) / (EK_WAVES * sh.mtw);
  sh.nt = (2 * wzh + 15) / 16;
  return sh;
}

template <int MTW>
void ek_sk_launch_nt(const EkArgs& a, const EkSkShape& sh, dim3 g, hipStream_t st) {
  if (sh.nt == 1) ek_sk_kernel<MTW, 1><<<g, EK_SK_THREADS, sh.lds, st>>>(a, sh);
  else ek_sk_kernel<MTW, 2><<<g, EK_SK_THREADS, sh.lds, st>>>(a, sh);
}

void ek_sk_launch(const EkArgs& a, int chunks, int n_sys, hipStream_t st) {
  const EkSkShape sh = ek_sk_shape(a);
  const dim3 g(chunks, n_sys, sh.groups);
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
  a.skc = reinterpret_cast<float*>(b + sk_bytes);
  a.skpart = reinterpret_cast<double*>(b + sk_bytes + skp_bytes);
  ek_sk_launch(a, a.chunks, n_sys, st);
  KUPSMI_CHECK_LAUNCH();
  ek_sk_reduce_kernel<<<dim3((a.nk + 255) / 256, n_sys), 256, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  ek_energy_kernel<<<n_sys, EK_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  if (ctx && ctx->phase) KUPSMI_HIP(hipEventRecord(ctx->phase, st));
  if (forces && max_atoms_per_sys > 0) {
    dim3 g2((max_atoms_per_sys + EK_THREADS - 1) / EK_THREADS, n_sys);
    const size_t lds = sizeof(float) * coef_floats;
    if (lds <= 64 * 1024) {
      if (a.mz == 8) ek_force_kernel<8, true><<<g2, EK_THREADS, lds, st>>>(a);
      else ek_force_kernel<0, true><<<g2, EK_THREADS, lds, st>>>(a);
    } else {
      ek_force_kernel<0, false><<<g2, EK_THREADS, 0, st>>>(a);
    }
    KUPSMI_CHECK_LAUNCH();
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
