// Reciprocal-space Ewald sum (potential/classical/ewald.py:349-439, 585-601).
//
// k = 2 pi Linv n over the half-space integer grid n in [0..mx] x [-my..my] x
// [-mz..mz] (the grid of kvecs_from_kmax, ewald.py:1322-1339, which keeps the
// vectors with |k| <= kmax; kmax < 0 keeps the whole box).  Phases factorise
// per axis, exp(i k.r) = ex(n0) ey(n1) ez(n2) with e_a(n) = exp(2 pi i n f_a)
// on the wrapped fractional coordinates, so the per-atom work is a product of
// three small tables instead of a sincos per (atom, k).
//
//   ek_sk_kernel      S(k) = sum_i q_i exp(i k.r_i): one workgroup per
//                     (system, 256 k vectors), atoms streamed through LDS in
//                     tiles of 64 (their phase tables built once per tile),
//                     FP32 tile sums accumulated in FP64.
//   ek_energy_kernel  P(k) (FP64), E = K sum_k P |S|^2, the reciprocal-space
//                     virial, and S'(k) = P(k) S(k) for the force pass.
//   ek_force_kernel   F_i = -2 K q_i sum_k k Im(S'(k) exp(-i k.r_i)): one thread
//                     per atom, S' of its system resident in LDS.
#include "common.hpp"

namespace kupsmi {

constexpr int EK_THREADS = 256;
constexpr int EK_TA = 64;          // atoms per LDS tile of the S(k) pass
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
  double* sk;       // n_sys x nk x 2 (FP64)
  float2* skp;      // n_sys x nk, P(k) S(k)
  float* forces;
  double* energy;
  double* virial;
};

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

__global__ __launch_bounds__(EK_THREADS) void ek_sk_kernel(EkArgs a) {
  __shared__ float2 tx[EK_TA][EK_NM + 1];
  __shared__ float2 ty[EK_TA][EK_W];
  __shared__ float2 tz[EK_TA][EK_W];
  const int s = blockIdx.y;
  const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
  const int tid = threadIdx.x;
  const int k = blockIdx.x * EK_THREADS + tid;
  const int wy = 2 * a.my + 1, wz = 2 * a.mz + 1;
  int kx = 0, ky = 0, kz = 0;
  if (k < a.nk) {
    kx = k / (wy * wz);
    ky = (k / wz) % wy;   // index into ty (n1 + my)
    kz = k % wz;          // index into tz (n2 + mz)
  }
  const double* t = a.tril + 6 * s;
  double accr = 0.0, acci = 0.0;
  for (int base = 0; base < n; base += EK_TA) {
    __syncthreads();
    // three quarter-warps of 64 threads build the x, y, z tables of 64 atoms
    const int role = tid / EK_TA, at = tid % EK_TA;
    if (role < 3) {
      const int i = base + at;
      if (i < n) {
        float2 ux, uy, uz;
        unit_phases(t, a.pos + 3 * (size_t)(i0 + i), ux, uy, uz);
        if (role == 0) {
          const float qi = a.q[i0 + i];
          float2 p = make_float2(qi, 0.f);
          for (int m = 0; m <= a.mx; ++m) {
            tx[at][m] = p;
            p = cmul(p, ux);
          }
        } else if (role == 1) {
          powers(uy, a.my, &ty[at][0]);
        } else {
          powers(uz, a.mz, &tz[at][0]);
        }
      } else {
        if (role == 0)
          for (int m = 0; m <= a.mx; ++m) tx[at][m] = make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
    if (k < a.nk) {
      const int cnt = n - base < EK_TA ? n - base : EK_TA;
      float sr = 0.f, si = 0.f;
      for (int j = 0; j < cnt; ++j) {
        const float2 v = cmul(cmul(tx[j][kx], ty[j][ky]), tz[j][kz]);
        sr += v.x;
        si += v.y;
      }
      accr += sr;
      acci += si;
    }
  }
  if (k < a.nk) {
    double* o = a.sk + 2 * ((size_t)s * a.nk + k);
    o[0] = accr;
    o[1] = acci;
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
  const int wy = 2 * a.my + 1, wz = 2 * a.mz + 1;
  // u^-m starting points
  float2 uy0 = make_float2(1.f, 0.f), uz0 = make_float2(1.f, 0.f);
  const float2 cy = make_float2(uy.x, -uy.y), cz = make_float2(uz.x, -uz.y);
  for (int m = 0; m < a.my; ++m) uy0 = cmul(uy0, cy);
  for (int m = 0; m < a.mz; ++m) uz0 = cmul(uz0, cz);
  float t0 = 0.f, t1 = 0.f, t2 = 0.f;
  float2 ex = make_float2(1.f, 0.f);
  int k = 0;
  for (int n0 = 0; n0 <= a.mx; ++n0) {
    float2 exy = cmul(ex, uy0);
    for (int j1 = 0; j1 < wy; ++j1) {
      const float n1 = (float)(j1 - a.my);
      float2 e = cmul(exy, uz0);
      float a0 = 0.f, a2 = 0.f;
      for (int j2 = 0; j2 < wz; ++j2, ++k) {
        const float2 sp = skp[k];
        // Im(S' conj(e))
        const float v = sp.y * e.x - sp.x * e.y;
        a0 += v;
        a2 += (float)(j2 - a.mz) * v;
        e = cmul(e, uz);
      }
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

extern "C" int kupsmi_ewald_recip(kupsmi_ctx* ctx, void* stream, int n_sys, const int32_t* sys_off,
                                  int n_atoms, int max_atoms_per_sys, const float* pos,
                                  const float* charges, const double* tril, const double* alpha,
                                  const double* kmax, int mx, int my, int mz, int accumulate,
                                  float* forces, double* energy, double* virial,
                                  double* structure_factor) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
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
  void* buf = nullptr;
  const size_t sk_bytes = structure_factor ? 0 : ((sizeof(double) * 2 * nsk + 255) & ~(size_t)255);
  int rc = scratch(ctx, 4, sk_bytes + sizeof(float2) * nsk, &buf);
  if (rc) return rc;
  a.sk = structure_factor ? structure_factor : static_cast<double*>(buf);
  a.skp = reinterpret_cast<float2*>(static_cast<char*>(buf) + sk_bytes);
  dim3 g1((a.nk + EK_THREADS - 1) / EK_THREADS, n_sys);
  ek_sk_kernel<<<g1, EK_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  ek_energy_kernel<<<n_sys, EK_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  if (forces && max_atoms_per_sys > 0) {
    dim3 g2((max_atoms_per_sys + EK_THREADS - 1) / EK_THREADS, n_sys);
    ek_force_kernel<<<g2, EK_THREADS, sizeof(float2) * a.nk, st>>>(a);
    KUPSMI_CHECK_LAUNCH();
  }
  return KUPSMI_OK;
}
