// Pair-path helpers shared by the cell-list kernel (pairs.hip) and the
// Verlet-list kernel (verlet.hip): constants, block scan, cell binning, the
// FP64 re-decision of the cutoff band, the minimum-image test and the LJ +
// real-space Ewald terms of one directed pair.
#pragma once
#include "common.hpp"

namespace kupsmi {

constexpr int PC_BIN_THREADS = 1024;
constexpr int PC_NMAX = 16384;
constexpr int PC_MAX_NC_AXIS = 16;
constexpr int PC_NCMAX = PC_MAX_NC_AXIS * PC_MAX_NC_AXIS * PC_MAX_NC_AXIS;
constexpr int PC_TILE = 256;
constexpr int PC_NSPEC = 16;
// relative half width of the band around each cutoff whose pairs are
// re-decided on the caller's positions: it covers the rounding of a wrapped
// copy (half an FP32 ulp per coordinate, ~6e-7 of rc^2 at rc = 12 A in a
// 30 A cell) with a wide margin
constexpr double PC_BAND64 = 4e-6;
constexpr int PC_NPART = 8;  // e_lj, e_real, w00 w01 w02 w11 w12 w22

// Exclusive scan of cnt[0..n) into start[0..n] by the whole workgroup, in
// chunks of 4 * blockDim entries (four per lane) carried over; wsum holds
// blockDim / 64 ints.
__device__ void block_scan4(const int* cnt, int* start, int n, int* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x / 64;
  int carry = 0;
  for (int base = 0; base < n; base += 4 * (int)blockDim.x) {
    int v4[4], own = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int c = base + tid * 4 + k;
      v4[k] = c < n ? cnt[c] : 0;
      own += v4[k];
    }
    int v = own;
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(v, o, 64);
      if (lane >= o) v += y;
    }
    if (lane == 63) wsum[wv] = v;
    __syncthreads();
    if (wv == 0) {
      int w = lane < nw ? wsum[lane] : 0;
      for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(w, o, 64);
        if (lane >= o) w += y;
      }
      if (lane < nw) wsum[lane] = w;
    }
    __syncthreads();
    int ex = carry + v - own + (wv > 0 ? wsum[wv - 1] : 0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int c = base + tid * 4 + k;
      if (c < n) start[c] = ex;
      ex += v4[k];
      if (c == n - 1) start[n] = ex;
    }
    carry += wsum[nw - 1];
    __syncthreads();  // wsum is rewritten by the next chunk
  }
}

__device__ inline int cell_of(const SysGeom& g, float x, float y, float z, float& wx, float& wy,
                              float& wz) {
  float f0, f1, f2;
  wx = x; wy = y; wz = z;
  wrap_shift(g.L, g.Lif, wx, wy, wz, f0, f1, f2);
  return (cell_index(f2, g.nc[2]) * g.nc[1] + cell_index(f1, g.nc[1])) * g.nc[0] +
         cell_index(f0, g.nc[0]);
}

// LDS of the per-system counting sort.
struct PcBinLds {
  int cnt[PC_NCMAX];
  int start[PC_NCMAX + 1];
  short keys[PC_NMAX];
  int wsum[PC_BIN_THREADS / 64];
};

// Sorted copies of a system's atoms (global arrays indexed i0 + d) and its
// cell starts cs[0..ncell].
struct PcSorted {
  float4* spos;  // wrapped (x, y, z, q)
  int* sidx;     // local index of the atom in the caller's order
  int* scell;
  int* sspec;
  int* slj;      // NULL when unused
  int* scl;
  int* cs;
};

// One workgroup (any size up to PC_BIN_THREADS) bins system i0..i0+n into cells of g:
// wrap, stable counting sort (ties keep atom order), sorted copies.
__device__ inline void pc_bin_system(const SysGeom& g, int i0, int n, const float* pos,
                                     const float* q, const int32_t* species,
                                     const int32_t* lj_excl, const int32_t* cl_excl, PcSorted o,
                                     PcBinLds& L) {
  const int tid = threadIdx.x;
  const int ncell = g.ncell;
  if (n <= 0 || n > PC_NMAX) {
    for (int c = tid; c <= ncell; c += (int)blockDim.x) o.cs[c] = 0;
    return;
  }
  for (int c = tid; c < ncell; c += (int)blockDim.x) L.cnt[c] = 0;
  __syncthreads();
  float wx, wy, wz;
  for (int i = tid; i < n; i += (int)blockDim.x) {
    const float* r = pos + 3 * (size_t)(i0 + i);
    atomicAdd(&L.cnt[cell_of(g, r[0], r[1], r[2], wx, wy, wz)], 1);
  }
  __syncthreads();
  block_scan4(L.cnt, L.start, ncell, L.wsum);
  __syncthreads();
  for (int c = tid; c < ncell; c += (int)blockDim.x) L.cnt[c] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += (int)blockDim.x) {
    const float* r = pos + 3 * (size_t)(i0 + i);
    const int c = cell_of(g, r[0], r[1], r[2], wx, wy, wz);
    L.keys[L.start[c] + atomicAdd(&L.cnt[c], 1)] = (short)i;
  }
  __syncthreads();
  for (int i = tid; i < n; i += (int)blockDim.x) {
    const float* r = pos + 3 * (size_t)(i0 + i);
    const int c = cell_of(g, r[0], r[1], r[2], wx, wy, wz);
    const int b = L.start[c], e = L.start[c + 1];
    int rank = 0;
    for (int t = b; t < e; ++t) rank += (int)(unsigned short)L.keys[t] < i;
    const int d = i0 + b + rank;
    o.spos[d] = make_float4(wx, wy, wz, q ? q[i0 + i] : 0.0f);
    o.sidx[d] = i;
    o.scell[d] = c;
    o.sspec[d] = species ? species[i0 + i] : 0;
    if (lj_excl) o.slj[d] = lj_excl[i0 + i];
    if (cl_excl) o.scl[d] = cl_excl[i0 + i];
  }
  for (int c = tid; c <= ncell; c += (int)blockDim.x) o.cs[c] = L.start[c];
}

// Band decisions on the caller's positions.  The pair kernels work on
// copies of the caller's positions that may have been moved by whole lattice
// vectors (wrapped, or kept next to a list's build positions) and rounded
// once; a pair within the band of a cutoff is re-decided with the
// reference's FP64 formula |(f_i - f_j - s) @ L|^2 < rc2
// (core/neighborlist.py:542-564, 589) on the caller's own positions, where
// r_j - r_i + s @ L on the copies is r_j - r_i + (s - n_j + n_i) @ L with
// caller = copy + n @ L.
//
// The out-of-line helpers take only the partner's sorted index and the
// shift: the copies, the caller's positions and the sorted -> caller map of
// the workgroup's system are in band_ctx (set by the kernel), and the left
// atom is sorted atom band_ctx.t0 + threadIdx.x, so the rare call
// costs the pair loops no registers.
struct BandCtx {
  const float4* spos;  // copies of the system (sorted)
  const float* cpos;   // caller's positions of the system
  const int* csidx;    // sorted -> caller index within the system
  int t0;              // the workgroup's first sorted atom
};
static __shared__ BandCtx band_ctx;

// The caller's positions of the pair and its caller-frame shift.  Written
// as a sequence of short phases (compiler barriers between them), so the
// helpers hold few values live at once and claim few registers.
__device__ inline void band_pair(const SysGeom& g, int tj, int sx, int sy, int sz,
                                 const float*& pi, const float*& pj, int& c0, int& c1, int& c2) {
  const BandCtx& c = band_ctx;
  const int ti = c.t0 + threadIdx.x;
  pi = c.cpos + 3 * (size_t)c.csidx[ti];
  pj = c.cpos + 3 * (size_t)c.csidx[tj];
  int a0, a1, a2, b0, b1, b2;
  {
    const float4 r = c.spos[ti];
    lattice_offset(g, pi, r.x, r.y, r.z, a0, a1, a2);
  }
  asm volatile("" ::: "memory");
  {
    const float4 r = c.spos[tj];
    lattice_offset(g, pj, r.x, r.y, r.z, b0, b1, b2);
  }
  asm volatile("" ::: "memory");
  c0 = sx - b0 + a0; c1 = sy - b1 + a1; c2 = sz - b2 + a2;
}

__device__ __attribute__((noinline)) bool band_keep(const SysGeom& g, int tj, int sx, int sy,
                                                    int sz, double rc2) {
  const float *pi, *pj;
  int c0, c1, c2;
  band_pair(g, tj, sx, sy, sz, pi, pj, c0, c1, c2);
  return pair_dist2_64(g, pi, pj, c0, c1, c2) < rc2;
}

// Minimum-image test of a replicated system's pair on the caller's
// positions: s == round(f_i - f_j) (the exclusion rule, 731-735).
__device__ __attribute__((noinline)) bool band_min_image(const SysGeom& g, int tj, int sx, int sy,
                                                         int sz) {
  const float *pi, *pj;
  int c0, c1, c2;
  band_pair(g, tj, sx, sy, sz, pi, pj, c0, c1, c2);
  return pair_is_min_image64(g, pi, pj, c0, c1, c2);
}

// Appends the pair (sorted left atom of this lane, tj, s) to a record in the
// caller's indices and frame; i0 is the system's first particle.
__device__ __attribute__((noinline)) void band_record(const PairRecord& r, const SysGeom& g, int i0,
                                                      int tj, int sx, int sy, int sz, int terms) {
  const float *pi, *pj;
  int c0, c1, c2;
  band_pair(g, tj, sx, sy, sz, pi, pj, c0, c1, c2);
  const BandCtx& c = band_ctx;
  record_pair(r, i0 + (int)((pi - c.cpos) / 3), i0 + (int)((pj - c.cpos) / 3), c0, c1, c2, terms);
}

// Pair vectors with every fused operation spelled out: the main pass and the
// band pass of a pair kernel form bitwise the same vector and r^2, so they
// agree on which pairs lie inside a cutoff band.
__device__ inline void shift_vec64(const double* L, int sx, int sy, int sz, double& x, double& y,
                                   double& z) {
  x = fma((double)sz, L[3], fma((double)sy, L[1], (double)sx * L[0]));
  y = fma((double)sz, L[4], (double)sy * L[2]);
  z = (double)sz * L[5];
}

__device__ inline double norm2_64(double x, double y, double z) {
  return fma(z, z, fma(y, y, x * x));
}

__device__ inline void shift_vec32(const float* L, int sx, int sy, int sz, float& x, float& y,
                                   float& z) {
  x = fmaf((float)sz, L[3], fmaf((float)sy, L[1], (float)sx * L[0]));
  y = fmaf((float)sz, L[4], (float)sy * L[2]);
  z = (float)sz * L[5];
}

__device__ inline float norm2_32(float x, float y, float z) { return fmaf(z, z, fmaf(y, y, x * x)); }

// Per-system constants of the pair terms (cutoffs, FP64 bands, Ewald alpha).
struct PairParams {
  int has_lj, has_coul, has_lj_excl, has_coul_excl, ns;
  double lj_rc2, lj_lo, lj_hi, cl_rc2, cl_lo, cl_hi, all_hi;
  float alpha, two_a_sqrtpi, alpha2;
  int has_sw;                                 // LJ smooth switch from sw_rtr2 to the cutoff
  float sw_rtr2, sw_rc2, sw_3rtr2, sw_inv_div;
  const float *sig2, *e4, *e24;  // LDS tables
};

// Smooth switch of pair_tail_corrected_lennard_jones_energy
// (potential/classical/lennard_jones.py:200-222): for r_tr < r < rc the pair
// energy u becomes u S with S = (rc^2 - r^2)^2 (2 r^2 + rc^2 - 3 r_tr^2) /
// (rc^2 - r_tr^2)^3 (S = 1, S' = 0 at r_tr; S = S' = 0 at rc).  gl holds
// 2 du/d(r^2) and becomes 2 d(uS)/d(r^2), with dS/d(r^2) =
// 6 (rc^2 - r^2)(r_tr^2 - r^2) / (rc^2 - r_tr^2)^3.
__device__ inline void lj_switch(const PairParams& pp, float r2, float& e, float& gl) {
  const float x = pp.sw_rc2 - r2;
  const float S = x * x * (2.0f * r2 + pp.sw_rc2 - pp.sw_3rtr2) * pp.sw_inv_div;
  const float dS = 6.0f * x * (pp.sw_rtr2 - r2) * pp.sw_inv_div;
  gl = gl * S + 2.0f * e * dS;
  e *= S;
}

__device__ inline void pc_pair_params(PairParams& pp, int has_lj, int has_coul, int has_lj_excl,
                                      int has_coul_excl, const double* lj_rc, const double* cl_rc,
                                      const double* alpha, int s, int ns, const float* sig2,
                                      const float* e4, const float* e24,
                                      const double* lj_rtr = nullptr) {
  pp.has_lj = has_lj; pp.has_coul = has_coul;
  pp.has_lj_excl = has_lj_excl; pp.has_coul_excl = has_coul_excl; pp.ns = ns;
  pp.lj_rc2 = has_lj ? lj_rc[s] * lj_rc[s] : 0.0;
  const double crc = has_coul ? cl_rc[s] : 0.0;
  pp.cl_rc2 = crc * crc;
  // FP64 band around each cutoff re-decided with the reference's fractional
  // formula (the Cartesian FP64 r^2 differs from it by a few ulp)
  pp.lj_hi = pp.lj_rc2 * (1.0 + PC_BAND64); pp.lj_lo = pp.lj_rc2 * (1.0 - PC_BAND64);
  pp.cl_hi = pp.cl_rc2 * (1.0 + PC_BAND64); pp.cl_lo = pp.cl_rc2 * (1.0 - PC_BAND64);
  pp.all_hi = fmax(pp.lj_hi, pp.cl_hi);
  pp.alpha = has_coul ? (float)alpha[s] : 0.0f;
  pp.two_a_sqrtpi = 2.0f * pp.alpha * 0.5641895835477563f;
  pp.alpha2 = pp.alpha * pp.alpha;
  pp.sig2 = sig2; pp.e4 = e4; pp.e24 = e24;
  const double rtr = has_lj && lj_rtr ? lj_rtr[s] : -1.0;
  pp.has_sw = rtr >= 0.0 && rtr * rtr < pp.lj_rc2;
  pp.sw_rtr2 = (float)(rtr * rtr);
  pp.sw_rc2 = (float)pp.lj_rc2;
  pp.sw_3rtr2 = (float)(3.0 * rtr * rtr);
  const double span = pp.lj_rc2 - rtr * rtr;
  pp.sw_inv_div = pp.has_sw ? (float)(1.0 / (span * span * span)) : 0.0f;
}

// LJ and real-space Ewald terms of the directed pair (i -> j + s) of working
// copies ri, rj (i the lane's sorted atom, j sorted atom tj): adds the pair
// energies to elj / ecl, sets kept (bit 0 LJ, bit 1 Coulomb) and returns
// 2 d(E_directed)/d(r^2).  r2d is the FP64 squared distance of the copies,
// r2 its FP32 rounding; exclusion ids are compared on the minimum image only
// (core/neighborlist.py:731-735).  Two passes: PASS 0 keeps the pairs below
// the band of a cutoff and sets inband for a pair inside it; PASS 1 (run
// only by lanes that met such a pair) keeps exactly the band pairs the
// caller-frame FP64 decision keeps.  The pair loops of PASS 0 then make no
// calls, which would pin registers across every trip.
template <bool EXCL, int PASS>
__device__ inline float pc_pair_terms(const PairParams& pp, const SysGeom& g, int tj, float4 ri,
                                      float4 rj, int sx, int sy, int sz, double r2d, float r2,
                                      bool self, int si, int sj, int xlj_i, int xlj_j, int xcl_i,
                                      int xcl_j, float& elj, float& ecl, int& kept, bool& inband) {
  const float K = 14.399645351950548f;  // HARTREE * BOHR, eV A
  bool min_img = true;
  if (EXCL && g.rep) min_img = band_min_image(g, tj, sx, sy, sz);
  float gsum = 0.f;
  kept = 0;
  if (pp.has_lj && r2d < pp.lj_hi && !self) {
    const bool mid = r2d > pp.lj_lo;
    bool ok;
    if (PASS == 0) { ok = !mid; inband |= mid; }
    else ok = mid && band_keep(g, tj, sx, sy, sz, pp.lj_rc2);
    if (EXCL && pp.has_lj_excl && min_img && xlj_j == xlj_i) ok = false;
    if (ok) {
      kept |= 1;
      const int p = si * pp.ns + sj;
      const float ir2 = 1.0f / r2;
      float c6 = pp.sig2[p] * ir2;
      c6 = c6 * c6 * c6;
      float e = pp.e4[p] * c6 * (c6 - 1.0f);
      float gl = pp.e24[p] * c6 * (1.0f - 2.0f * c6) * ir2;
      if (pp.has_sw && r2 > pp.sw_rtr2) lj_switch(pp, r2, e, gl);
      elj += e;
      gsum += gl;
    }
  }
  if (pp.has_coul && r2d < pp.cl_hi && !self) {
    const bool mid = r2d > pp.cl_lo;
    bool ok;
    if (PASS == 0) { ok = !mid; inband |= mid; }
    else ok = mid && band_keep(g, tj, sx, sy, sz, pp.cl_rc2);
    if (EXCL && pp.has_coul_excl && min_img && xcl_j == xcl_i) ok = false;
    if (ok) {
      kept |= 2;
      const float r = sqrtf(r2);
      const float ir = 1.0f / r;
      const float qq = ri.w * rj.w * K;
      const float ec = erfcf(pp.alpha * r);
      ecl += qq * ec * ir;
      // dU/dr = qq (-2a/sqrt(pi) e^{-a^2 r^2} / r - erfc / r^2); 2 u' = dU/dr / r
      gsum += qq * (-pp.two_a_sqrtpi * __expf(-pp.alpha2 * r2) - ec * ir) * ir * ir;
    }
  }
  return gsum;
}

// Per-block FP64 partials of a tile (fixed order: wave sums, then the waves
// in order) written to part[0..PC_NPART).
template <int TILE>
__device__ inline void pc_block_partials(const double (&v)[PC_NPART], double (*red)[PC_NPART],
                                         double* part) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int k = 0; k < PC_NPART; ++k) {
    const double x = wave_sum(v[k]);
    if (lane == 0) red[wv][k] = x;
  }
  __syncthreads();
  if (tid < PC_NPART) {
    double x = 0.0;
    for (int w = 0; w < TILE / 64; ++w) x += red[w][tid];
    part[tid] = x;
  }
}

// Fixed-order sum of the tile partials of system s (then those of the band
// pass, if any) into energy[2 s..] (LJ, real space) and the symmetric
// virial[9 s..].
__device__ inline void pc_reduce_system(int s, int tiles, const double* partial, int accumulate,
                                        double* energy, double* virial,
                                        const double* partial2 = nullptr) {
  double acc[PC_NPART] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = 0; t < tiles; ++t)
    for (int k = 0; k < PC_NPART; ++k) acc[k] += partial[((size_t)s * tiles + t) * PC_NPART + k];
  // the band pass's partials, after the main pass's
  if (partial2)
    for (int t = 0; t < tiles; ++t)
      for (int k = 0; k < PC_NPART; ++k) acc[k] += partial2[((size_t)s * tiles + t) * PC_NPART + k];
  if (energy) {
    double* e = energy + 2 * (size_t)s;
    if (accumulate) { e[0] += acc[0]; e[1] += acc[1]; } else { e[0] = acc[0]; e[1] = acc[1]; }
  }
  if (virial) {
    const double w[9] = {acc[2], acc[3], acc[4], acc[3], acc[5], acc[6], acc[4], acc[6], acc[7]};
    double* W = virial + 9 * (size_t)s;
    for (int k = 0; k < 9; ++k) W[k] = accumulate ? W[k] + w[k] : w[k];
  }
}

}  // namespace kupsmi
