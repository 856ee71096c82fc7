// Shared device/host helpers of libkupsmi (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include "../../include/kupsmi.h"

namespace kupsmi {

// ------------------------------------------------------------- errors ----
void set_error(const std::string& msg);
#define KUPSMI_HIP(call)                                                   \
  do {                                                                     \
    hipError_t _e = (call);                                                \
    if (_e != hipSuccess) {                                                \
      ::kupsmi::set_error(std::string(#call) + ": " + hipGetErrorString(_e)); \
      return KUPSMI_ERR_HIP;                                               \
    }                                                                      \
  } while (0)
#define KUPSMI_CHECK_LAUNCH() KUPSMI_HIP(hipGetLastError())
#define KUPSMI_REQUIRE(cond, msg)                                          \
  do {                                                                     \
    if (!(cond)) {                                                         \
      ::kupsmi::set_error(msg);                                            \
      return KUPSMI_ERR_ARG;                                               \
    }                                                                      \
  } while (0)

// ------------------------------------------------------------ context ----
// Grow-only device scratch.  Allocation happens only when a call needs more
// than any earlier call; steady-state calls allocate nothing.  Growing a slot
// frees its old buffer, so work recorded with the old address (a captured
// HIP graph) must not outlive it: kupsmi_ctx_freeze turns growth into an
// error.
struct Scratch {
  void* ptr = nullptr;
  size_t bytes = 0;
};

// Parity output of the force kernels (see record_pair below).
struct PairRecord {
  int64_t cap = 0;
  int32_t* ij = nullptr;
  int32_t* shift = nullptr;
  uint8_t* terms = nullptr;
  unsigned long long* count = nullptr;
};
}  // namespace kupsmi

struct kupsmi_ctx {
  int device = 0;
  kupsmi::Scratch slots[16];
  // armed by kupsmi_ctx_record_pairs, taken by the next pair-evaluating call
  kupsmi::PairRecord rec;
  int rec_armed = 0;
  // set by kupsmi_ctx_freeze: a slot that would have to grow is an error
  // (recorded work holds the slots' addresses)
  int frozen = 0;
  // recorded by each kupsmi_ewald_recip on this context between its
  // structure-factor / energy kernels and its force pass
  hipEvent_t phase = nullptr;
};

namespace kupsmi {
// Returns a device buffer of at least `bytes` for scratch slot `slot`.
int scratch(kupsmi_ctx* ctx, int slot, size_t bytes, void** out);

// --------------------------------------------------------- unit cells ----
// Per-system geometry derived on the device from tril and cutoff, the
// quantities core/neighborlist.py:281-292 (cells per axis) and 453-459
// (image counts) compute.
struct SysGeom {
  double L[6];    // L00 L10 L11 L20 L21 L22
  double Li[6];   // inverse, same layout
  float Lf[6];
  float Lif[6];
  double rc, rc2;
  float rc2f;
  int nc[3];      // cells per axis
  int h[3];       // stencil half width per axis (1, or image half count)
  int rep;        // 1 if the system needs periodic images (image grid)
  int ncell;
  int ortho;      // 1 for an orthorhombic cell (L10 = L20 = L21 = 0)
};

__host__ __device__ inline void tril_inverse(const double* t, double* li) {
  // unfused, so the inverse has the same bits as a plain C evaluation
#pragma clang fp contract(off)
  const double a = t[0], b = t[1], c = t[2], d = t[3], e = t[4], f = t[5];
  li[0] = 1.0 / a;
  li[1] = -b / (a * c);
  li[2] = 1.0 / c;
  li[3] = (b * e - c * d) / (a * c * f);
  li[4] = -e / (c * f);
  li[5] = 1.0 / f;
}

// x @ M for lower-triangular M in tril layout (core/utils/math.py:177-239).
template <typename T, typename M>
__host__ __device__ inline void tmul(const M* m, T x0, T x1, T x2, T& y0, T& y1, T& y2) {
  y0 = x0 * (T)m[0] + x1 * (T)m[1] + x2 * (T)m[3];
  y1 = x1 * (T)m[2] + x2 * (T)m[4];
  y2 = x2 * (T)m[5];
}

// max_nc_axis bounds the cells per axis (LDS tables).  Cells are at least
// rc / sub wide and the stencil spans +-sub cells, so every pair within rc is
// visited; sub = 1 is the classic 27-cell stencil, sub = 2 the 125-cell one
// that halves the candidates of dense, long-cutoff systems.
__host__ __device__ inline void make_geom(const double* tril, double rc, int max_nc_axis,
                                          SysGeom& g, int sub = 1) {
  for (int k = 0; k < 6; ++k) {
    g.L[k] = tril[k];
    g.Lf[k] = (float)tril[k];
  }
  tril_inverse(tril, g.Li);
  for (int k = 0; k < 6; ++k) g.Lif[k] = (float)g.Li[k];
  const double* t = tril;
  // rows a=(t0,0,0) b=(t1,t2,0) c=(t3,t4,t5); vol = t0 t2 t5
  double vol = fabs(t[0] * t[2] * t[5]);
  // |b x c|, |a x c|, |a x b|
  double bc0 = t[2] * t[5], bc1 = -t[1] * t[5], bc2 = t[1] * t[4] - t[2] * t[3];
  double ac1 = -t[0] * t[5], ac2 = t[0] * t[4];
  double ab2 = t[0] * t[2];
  double perp[3];
  perp[0] = vol / sqrt(bc0 * bc0 + bc1 * bc1 + bc2 * bc2);
  perp[1] = vol / sqrt(ac1 * ac1 + ac2 * ac2);
  perp[2] = vol / fabs(ab2);
  g.rc = rc;
  g.rc2 = rc * rc;
  g.rc2f = (float)g.rc2;
  int img[3];
  for (int a = 0; a < 3; ++a) {
    int n = isfinite(rc) ? (int)ceil(2.0 * rc / perp[a]) : 1;
    if (n < 1) n = 1;
    if (n % 2 == 0) n += 1;
    img[a] = n;
  }
  g.rep = img[0] * img[1] * img[2] > 1;
  g.ortho = t[1] == 0.0 && t[3] == 0.0 && t[4] == 0.0;
  for (int a = 0; a < 3; ++a) {
    if (g.rep) {
      g.nc[a] = 1;
      g.h[a] = (img[a] - 1) / 2;
    } else {
      // cells at least rc (1 + 1e-6) / sub wide: an FP32 cell assignment off
      // by one at a cell face then still finds every pair within rc
      int n = isfinite(rc) ? (int)(sub * perp[a] / (rc * (1.0 + 1e-6))) : 1;
      if (n < 1) n = 1;
      if (n > max_nc_axis) n = max_nc_axis;
      g.nc[a] = n;
      g.h[a] = sub;
    }
  }
  g.ncell = g.nc[0] * g.nc[1] * g.nc[2];
}

// Fractional wrap to [-0.5, 0.5) (core/unitcell.py:139-171), returns the
// wrapped fractional coordinate.
__device__ inline float wrap_half(float f) {
  float u = f + 0.5f;
  u = u - floorf(u);
  if (u >= 1.0f) u = 0.0f;  // guard the rounding of u - floor(u) to 1
  return u - 0.5f;
}

// Wrap a Cartesian position into the primary cell by whole lattice vectors:
// n = floor(f + 1/2) per axis and r' = r - n @ L (in FP64, rounded once), so a
// position already inside the cell is kept bit for bit.  Returns the wrapped
// fractional coordinates in f0..f2 (for cell indices).
__device__ inline void wrap_shift(const double* L, const float* Lif, float& x, float& y, float& z,
                                  float& f0, float& f1, float& f2) {
  tmul(Lif, x, y, z, f0, f1, f2);
  const float n0 = floorf(f0 + 0.5f), n1 = floorf(f1 + 0.5f), n2 = floorf(f2 + 0.5f);
  f0 -= n0; f1 -= n1; f2 -= n2;
  if (n0 != 0.f || n1 != 0.f || n2 != 0.f) {
    double s0, s1, s2;
    tmul(L, (double)n0, (double)n1, (double)n2, s0, s1, s2);
    x = (float)((double)x - s0);
    y = (float)((double)y - s1);
    z = (float)((double)z - s2);
  }
}

// ------------------------------------------- FP64 pair decisions ----
// The reference's pair test (core/neighborlist.py:542-564, 589): fractional
// coordinates f = r @ L^-1, d = (f_i - f_j - s) @ L, kept when |d|^2 < rc^2
// (strict).  Evaluated in FP64 without multiply-add fusion, operation by
// operation in the order of a plain C evaluation, so every kernel that
// decides a pair here decides it exactly as a CPU restatement does.
__device__ inline void frac64(const double* li, double x, double y, double z, double& f0,
                              double& f1, double& f2) {
#pragma clang fp contract(off)
  f0 = x * li[0] + y * li[1] + z * li[3];
  f1 = y * li[2] + z * li[4];
  f2 = z * li[5];
}

__device__ inline double dist2_64(const double* L, double d0, double d1, double d2) {
#pragma clang fp contract(off)
  const double x = d0 * L[0] + d1 * L[1] + d2 * L[3];
  const double y = d1 * L[2] + d2 * L[4];
  const double z = d2 * L[5];
  return x * x + y * y + z * z;
}

// |(f_i - f_j - s) @ L|^2 for positions pi, pj (FP32 as stored, widened).
__device__ inline double pair_dist2_64(const SysGeom& g, const float* pi, const float* pj, int s0,
                                       int s1, int s2) {
#pragma clang fp contract(off)
  double fi0, fi1, fi2, fj0, fj1, fj2;
  frac64(g.Li, pi[0], pi[1], pi[2], fi0, fi1, fi2);
  frac64(g.Li, pj[0], pj[1], pj[2], fj0, fj1, fj2);
  const double d0 = fi0 - fj0, d1 = fi1 - fj1, d2 = fi2 - fj2;
  return dist2_64(g.L, d0 - s0, d1 - s1, d2 - s2);
}

// Minimum image of the pair: s == rint(f_i - f_j) per axis (jnp.round rounds
// half to even, as rint does in the default rounding mode).
__device__ inline bool pair_is_min_image64(const SysGeom& g, const float* pi, const float* pj,
                                           int s0, int s1, int s2) {
#pragma clang fp contract(off)
  double fi0, fi1, fi2, fj0, fj1, fj2;
  frac64(g.Li, pi[0], pi[1], pi[2], fi0, fi1, fi2);
  frac64(g.Li, pj[0], pj[1], pj[2], fj0, fj1, fj2);
  return rint(fi0 - fj0) == s0 && rint(fi1 - fj1) == s1 && rint(fi2 - fj2) == s2;
}

// Whole lattice vectors n with caller = copy + n @ L, for a working copy of a
// caller's position that was moved by lattice vectors (and rounded once).
__device__ inline void lattice_offset(const SysGeom& g, const float* caller, float cx, float cy,
                                      float cz, int& n0, int& n1, int& n2) {
  double e0, e1, e2;
  frac64(g.Li, (double)caller[0] - cx, (double)caller[1] - cy, (double)caller[2] - cz, e0, e1,
         e2);
  n0 = (int)rint(e0); n1 = (int)rint(e1); n2 = (int)rint(e2);
}

// -------------------------------------------------- pair-set records ----
// Parity output of the force kernels: the directed pairs (i, j, s) a kernel
// evaluated, in global particle indices and the caller's frame, with a bit
// mask of the terms that kept the pair (bit 0 LJ, bit 1 real-space
// Coulomb).  Appended in no particular order; entries past cap are counted,
// not written.
__device__ inline void record_pair(const PairRecord& r, int gi, int gj, int s0, int s1, int s2,
                                   int terms) {
  const unsigned long long e = atomicAdd(r.count, 1ull);
  if (e < (unsigned long long)r.cap) {
    r.ij[2 * e] = gi; r.ij[2 * e + 1] = gj;
    r.shift[3 * e] = s0; r.shift[3 * e + 1] = s1; r.shift[3 * e + 2] = s2;
    if (r.terms) r.terms[e] = (uint8_t)terms;
  }
}

static __global__ void record_clear_kernel(unsigned long long* count) { *count = 0ull; }

__device__ inline int cell_index(float fw, int n) {
  int b = (int)((fw + 0.5f) * (float)n);
  return b < 0 ? 0 : (b >= n ? n - 1 : b);
}

// --------------------------------------------------------- Philox RNG ----
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), counter-based.
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * c[0];
    uint64_t p1 = (uint64_t)M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += W0; k1 += W1;
  }
}

// Uniform in (0, 1): (x + 0.5) 2^-32.
__host__ __device__ inline float u01(uint32_t x) {
  return ((float)x + 0.5f) * 2.3283064365386963e-10f;
}

// Three standard normals for (atom, step) by Box-Muller.
__device__ inline void normals3(uint64_t seed, uint32_t atom, int64_t step, float& n0,
                                float& n1, float& n2) {
  uint32_t c[4] = {atom, (uint32_t)step, (uint32_t)((uint64_t)step >> 32), 0x4B555053u};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  float r0 = sqrtf(-2.0f * logf(u01(c[0])));
  float r1 = sqrtf(-2.0f * logf(u01(c[2])));
  float s0, c0, s1, c1;
  sincospif(2.0f * u01(c[1]), &s0, &c0);
  sincospif(2.0f * u01(c[3]), &s1, &c1);
  n0 = r0 * c0;
  n1 = r0 * s0;
  n2 = r1 * c1;
}

// ------------------------------------------------------------- CSVR ----
// Stochastic velocity rescaling (Bussi, Donadio and Parrinello, J. Chem.
// Phys. 126, 014101 (2007); md/integrators.py:470-585): per system
//   alpha^2 = c1 + c2 (R1^2 + R2) + 2 R1 sqrt(c1 c2), c1 = exp(-dt / tau),
//   c2 = (1 - c1) K_target / (K N_dof), K_target = N_dof kT / 2, N_dof = 3N - 3,
// with ONE R1 ~ N(0, 1) for the whole batch (the reference draws a scalar)
// and R2 ~ chi^2(N_dof - 1) = 2 Gamma((N_dof - 1) / 2) per system by
// Marsaglia and Tsang (ACM TOMS 26 (2000) 363-372).  Counters: R1 from
// {0xFFFFFFFF, t, t >> 32, CSVR_TAG_R1}; attempt k of system s from
// {s, t, t >> 32, CSVR_TAG_R2 + k} (x = Box-Muller of words 0, 1; u = word 2);
// shape < 1 boosts from shape + 1 with U^(1/shape), U = word 0 of attempt 255.
constexpr uint32_t CSVR_TAG_R1 = 0x43535631u;
constexpr uint32_t CSVR_TAG_R2 = 0x43535700u;
// the batch-shared normal of the stochastic cell rescaling, counter
// {0xFFFFFFFE, t, t >> 32, SCR_TAG}
constexpr uint32_t SCR_TAG = 0x53435231u;

__device__ inline double csvr_u(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }

__device__ inline double csvr_bm(uint32_t a, uint32_t b) {
  return sqrt(-2.0 * log(csvr_u(a))) * cospi(2.0 * csvr_u(b));
}

__device__ inline void csvr_draw(uint64_t seed, uint32_t a, int64_t t, uint32_t tag, uint32_t c[4]) {
  c[0] = a; c[1] = (uint32_t)t; c[2] = (uint32_t)((uint64_t)t >> 32); c[3] = tag;
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__device__ inline double csvr_gamma(double shape, uint64_t seed, uint32_t sys, int64_t t) {
  uint32_t c[4];
  double boost = 1.0;
  if (shape < 1.0) {
    csvr_draw(seed, sys, t, CSVR_TAG_R2 + 255u, c);
    boost = pow(csvr_u(c[0]), 1.0 / shape);
    shape += 1.0;
  }
  const double d = shape - 1.0 / 3.0, cc = 1.0 / sqrt(9.0 * d);
  for (uint32_t k = 0; k < 255u; ++k) {
    csvr_draw(seed, sys, t, CSVR_TAG_R2 + k, c);
    const double x = csvr_bm(c[0], c[1]);
    double v = 1.0 + cc * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    if (log(csvr_u(c[2])) < 0.5 * x * x + d - d * v + d * log(v)) return boost * d * v;
  }
  return boost * d;
}

// alpha of system `sys` at global step t from its kinetic energy K (eV); 1
// for a system without degrees of freedom or kinetic energy.
__device__ inline double csvr_alpha(double K, int n_atoms, double kT, double dt, double tau,
                                    uint64_t seed, uint32_t sys, int64_t t) {
  const int ndof = 3 * n_atoms - 3;
  if (ndof <= 0 || !(K > 0.0)) return 1.0;
  uint32_t c[4];
  csvr_draw(seed, 0xFFFFFFFFu, t, CSVR_TAG_R1, c);
  const double r1 = csvr_bm(c[0], c[1]);
  const double r2 = ndof - 1 > 0 ? 2.0 * csvr_gamma(0.5 * (ndof - 1), seed, sys, t) : 0.0;
  const double c1 = exp(-dt / tau);
  const double c2 = (1.0 - c1) * (0.5 * ndof * kT) / (K * ndof);
  const double a2 = c1 + c2 * (r1 * r1 + r2) + 2.0 * r1 * sqrt(c1 * c2);
  return sqrt(a2 > 0.0 ? a2 : 0.0);
}

// erfc(x), x >= 0, by the Chebyshev fit of Press, Teukolsky, Vetterling and
// Flannery, Numerical Recipes in C, 2nd ed. (1992), section 6.2 (erfcc):
// fractional error below 1.2e-7 in exact arithmetic.  In FP32 the rounding of
// the exponent (about -13 at x = 3.6, the usual alpha rc) adds ~1e-6
// relative where erfc itself is ~4e-7, i.e. ~1e-12 eV per pair.
__device__ inline float erfc_cheb(float x) {
  const float t = __builtin_amdgcn_rcpf(1.0f + 0.5f * x);
  float p = 0.17087277f;
  p = fmaf(p, t, -0.82215223f);
  p = fmaf(p, t, 1.48851587f);
  p = fmaf(p, t, -1.13520398f);
  p = fmaf(p, t, 0.27886807f);
  p = fmaf(p, t, -0.18628806f);
  p = fmaf(p, t, 0.09678418f);
  p = fmaf(p, t, 0.37409196f);
  p = fmaf(p, t, 1.00002368f);
  p = fmaf(p, t, -1.26551223f);
  return t * __expf(p - x * x);
}

// erfcx(x) = exp(x^2) erfc(x) for x >= 0 as t Q(t), t = 1 / (1 + 0.6 x), Q
// the degree-9 polynomial below, fitted by iteratively reweighted least
// squares for relative error over x in [0, inf) (below 2e-7 in exact
// arithmetic, 3.5e-7 evaluated in FP32; erfcx -> 1 / (sqrt(pi) x) for large
// x is the limit t Q(t) -> Q(0) t, Q(0) = 0.6 / sqrt(pi)).  erfc(x) = erfcx(x) exp(-x^2) then shares
// its exponential with the derivative d erfc / dx = -2 / sqrt(pi) exp(-x^2).
__device__ inline float erfcx_fit(float x) {
  const float t = __builtin_amdgcn_rcpf(fmaf(0.6f, x, 1.0f));
  float q = 7.742052525e-02f;
  q = fmaf(q, t, -3.791217506e-01f);
  q = fmaf(q, t, 6.996498704e-01f);
  q = fmaf(q, t, -5.089484453e-01f);
  q = fmaf(q, t, 9.372118860e-03f);
  q = fmaf(q, t, -8.299292997e-03f);
  q = fmaf(q, t, 1.549688876e-01f);
  q = fmaf(q, t, 2.779625654e-01f);
  q = fmaf(q, t, 3.384810984e-01f);
  q = fmaf(q, t, 3.385145068e-01f);
  return t * q;
}

// ---------------------------------------------------------- reductions ----
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline float wave_sumf(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace kupsmi
