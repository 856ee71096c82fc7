// Batched LJ molecular dynamics with every system resident in LDS.
//
// One 1024-thread workgroup owns one system at a time (<= 4096 atoms, 4 slots
// per thread) and runs all requested MD steps without leaving the CU: state
// (positions, momenta, forces) stays in LDS and registers, the cell list is
// rebuilt in LDS every step, and HBM is touched only to load the system and
// to store it back.  Systems are taken from a largest-first work queue so the
// 64..4096-atom spread of a batch balances over the 256 CUs.
//
// Reference semantics (md/integrators.py:273-317 velocity Verlet, 405-453
// BAOAB; potential/classical/lennard_jones.py:149-178 pair energy; neighbor
// set of core/neighborlist.py:692-786 with particle exclusion).
#include "common.hpp"

// Multiply-adds are fused only within one source expression: the forces of
// a step are then the same bits whether or not that step also accumulates
// the virial (a separate instantiation of the pair loops).
#pragma clang fp contract(on)

namespace kupsmi {

constexpr int MD_THREADS = 1024;
constexpr int MD_NMAX = KUPSMI_LDS_MAX_ATOMS;
constexpr int MD_SLOTS = MD_NMAX / MD_THREADS;
// cells are at least rc/2 wide (stencil of +-2 cells per axis)
constexpr int MD_SUB = 2;
constexpr int MD_MAX_NC_AXIS = 13;  // <= 2197 cells per system
constexpr int MD_NCMAX = MD_MAX_NC_AXIS * MD_MAX_NC_AXIS * MD_MAX_NC_AXIS;
constexpr int MD_CPT = (MD_NCMAX + MD_THREADS - 1) / MD_THREADS;
constexpr int MD_NSPEC = 8;
// relative half width of the FP32 cutoff band re-decided in FP64
constexpr float CUT_BAND = 4e-6f;

struct MdArgs {
  int integrator, n_steps, n_sys, n_species, want_virial;
  // Verlet lists: skin > 0 and nblist != NULL enable them; nbmax entries per
  // atom, MD_NMAX atoms per workgroup
  float skin;
  int nbmax;
  unsigned short* nblist;
  int64_t step0;
  uint64_t seed;
  const int32_t* sys_off;
  const int32_t* order;
  int* queue;
  const SysGeom* geom;
  float* pos;
  float* mom;
  float* forces;
  const float* mass;
  const int32_t* species;
  const double* sigma;
  const double* epsilon;
  const double* tril;
  const double* cutoff;
  const double* dt;
  const double* kT;
  const double* gamma;
  const double* tau;  // CSVR coupling time per system (internal units)
  // stochastic cell rescaling (integrator 3) per system: target pressure,
  // coupling time, compressibility, minimum scale factor; the cell out
  const double* baro;
  double* tril_out;
  double* pot;
  double* kin;
  double* virial;
  // parity output: the pair set of the last step's final force pass
  PairRecord rec;
};

// Largest-first order of the systems (counting sort of the sizes, clamped to
// the LDS capacity) and reset of the work-queue counter.  One workgroup.
__global__ __launch_bounds__(1024) void order_kernel(int n_sys, const int32_t* sys_off,
                                                     int32_t* order, int* queue) {
  __shared__ int hist[MD_NMAX + 2];
  const int tid = threadIdx.x;
  for (int b = tid; b < MD_NMAX + 2; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  for (int s = tid; s < n_sys; s += blockDim.x) {
    int n = sys_off[s + 1] - sys_off[s];
    n = n < 0 ? 0 : (n > MD_NMAX + 1 ? MD_NMAX + 1 : n);
    atomicAdd(&hist[MD_NMAX + 1 - n], 1);  // descending size
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int b = 0; b < MD_NMAX + 2; ++b) {
      int c = hist[b];
      hist[b] = acc;
      acc += c;
    }
    *queue = 0;
  }
  __syncthreads();
  for (int s = tid; s < n_sys; s += blockDim.x) {
    int n = sys_off[s + 1] - sys_off[s];
    n = n < 0 ? 0 : (n > MD_NMAX + 1 ? MD_NMAX + 1 : n);
    int at = atomicAdd(&hist[MD_NMAX + 1 - n], 1);
    order[at] = s;
  }
}

// Geometry of every system (cells per axis, image stencil, inverse cell).
// Cells and images are sized for the list radius cutoff + skin; rc / rc2 /
// rc2f hold the cutoff itself, and `rep` tells whether the cutoff itself
// needs periodic images (a box thinner than 2 rc: no Verlet list).
__global__ void geom_kernel(int n_sys, const double* tril, const double* cutoff, int max_nc,
                            double skin, SysGeom* geom) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_sys) return;
  const double rc = cutoff[s];
  SysGeom g0;
  make_geom(tril + 6 * s, rc, max_nc, g0, MD_SUB);
  SysGeom g;
  make_geom(tril + 6 * s, rc + (skin > 0.0 ? skin : 0.0), max_nc, g, MD_SUB);
  g.rc = rc;
  g.rc2 = rc * rc;
  g.rc2f = (float)g.rc2;
  g.rep = g0.rep;
  geom[s] = g;
}

// Exclusive scan of cnt[0..n) into start[0..n] (n <= MD_NCMAX) by the
// block: MD_CPT consecutive cells per thread, then a scan of the thread sums.
__device__ inline void block_scan_cells(const int* cnt, int* start, int n, int* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int c0 = tid * MD_CPT;
  int loc[MD_CPT];
  int own = 0;
#pragma unroll
  for (int k = 0; k < MD_CPT; ++k) {
    loc[k] = c0 + k < n ? cnt[c0 + k] : 0;
    own += loc[k];
  }
  int v = own;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) wsum[wv] = v;
  __syncthreads();
  if (wv == 0) {
    int w = lane < (MD_THREADS / 64) ? wsum[lane] : 0;
    for (int o = 1; o < 16; o <<= 1) {
      int y = __shfl_up(w, o, 64);
      if (lane >= o) w += y;
    }
    if (lane < MD_THREADS / 64) wsum[lane] = w;
  }
  __syncthreads();
  int ex = v - own + (wv > 0 ? wsum[wv - 1] : 0);
#pragma unroll
  for (int k = 0; k < MD_CPT; ++k) {
    if (c0 + k < n) start[c0 + k] = ex;
    ex += loc[k];
    if (c0 + k == n - 1) start[n] = ex;
  }
  if (n == 0 && tid == 0) start[0] = 0;
}

__device__ inline double block_sum(double v, double* red) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0.0;
  if (tid < 64) {
    t = tid < MD_THREADS / 64 ? red[tid] : 0.0;
    t = wave_sum(t);
  }
  return t;  // valid in thread 0
}

// FP64 re-decision of a pair whose FP32 r^2 lies in the cutoff band, with
// the reference's fractional formula |(f_i - f_j - s) @ L|^2 < rc^2
// (core/neighborlist.py:542-564, 589).  Rare: only pairs within a few ulp of
// the cutoff take it.
__device__ inline bool band_keep(const SysGeom& g, float4 ri, float4 rj, int sx, int sy,
                                 int sz) {
  // keep the FP64 cell matrices out of the loop's live registers: reload them
  asm volatile("" ::: "memory");
  const float pi[3] = {ri.x, ri.y, ri.z}, pj[3] = {rj.x, rj.y, rj.z};
  return pair_dist2_64(g, pi, pj, sx, sy, sz) < g.rc2;
}

// floor(x / n) for a small integer x and 1 <= n <= MD_MAX_NC_AXIS: the
// quotient of (x + 1/2) is at least 1/(2n) away from every integer, far more
// than the rounding of the FP32 product.
__device__ inline int floor_div_small(int x, float inv_n) {
  return (int)floorf(((float)x + 0.5f) * inv_n);
}

// ---- per-atom force walks (one lane per atom) ------------------------------

// LJ terms of one candidate with edge vector e = r_j + s@L - r_i and FP32 r^2.
// The keep decision is folded into 1/r^2 (zero for a rejected candidate).
template <bool VIR>
__device__ inline void md_lj_terms(float ex, float ey, float ez, float r2, bool keep, float4 pp,
                                   float& ax, float& ay, float& az, float& ei, float (&w)[6]) {
  const float ir2 = keep ? __builtin_amdgcn_rcpf(r2) : 0.0f;
  float c6 = pp.x * ir2;
  c6 = c6 * c6 * c6;
  ei += pp.y * c6 * (c6 - 1.0f);
  const float g2 = pp.z * c6 * (1.0f - 2.0f * c6) * ir2;  // u'(r) / r
  ax += g2 * ex; ay += g2 * ey; az += g2 * ez;
  if (VIR) {
    const float hg = -0.5f * g2;
    w[0] += hg * ex * ex; w[1] += hg * ex * ey; w[2] += hg * ex * ez;
    w[3] += hg * ey * ey; w[4] += hg * ey * ez; w[5] += hg * ez * ez;
  }
}

// FP64 re-decision of a band pair on its nearest image: s = rint(f_i - f_j)
// per axis, kept when |(f_i - f_j - s) @ L|^2 < rc^2 (the reference's
// fractional formula, core/neighborlist.py:542-564, 589).  Out of line: the
// rare call claims no registers in the pair loops.
// Positions are passed as x, y, z only, so the pair loop's loads of the
// neighbour need not fetch the w word.
// The FP64 nearest image s = rint(f_i - f_j) of a pair.
__device__ inline void md_nearest64(const SysGeom& g, float xi, float yi, float zi, float xj,
                                    float yj, float zj, int& s0, int& s1, int& s2) {
#pragma clang fp contract(off)
  double fi0, fi1, fi2, fj0, fj1, fj2;
  frac64(g.Li, xi, yi, zi, fi0, fi1, fi2);
  frac64(g.Li, xj, yj, zj, fj0, fj1, fj2);
  s0 = (int)rint(fi0 - fj0); s1 = (int)rint(fi1 - fj1); s2 = (int)rint(fi2 - fj2);
}

__device__ __attribute__((noinline)) bool md_band_nearest(const SysGeom& g, float xi, float yi,
                                                          float zi, float xj, float yj, float zj) {
#pragma clang fp contract(off)
  double fi0, fi1, fi2, fj0, fj1, fj2;
  frac64(g.Li, xi, yi, zi, fi0, fi1, fi2);
  frac64(g.Li, xj, yj, zj, fj0, fj1, fj2);
  const double t0 = fi0 - fj0, t1 = fi1 - fj1, t2 = fi2 - fj2;
  return dist2_64(g.L, t0 - rint(t0), t1 - rint(t1), t2 - rint(t2)) < g.rc2;
}

// Where a record goes: the pair set of a force walk in global indices
// (local ids in the w word of the momenta buffer).
// caller_pos (a force evaluation without steps, whose positions are not
// written back): shifts are converted to the frame of the caller's positions,
// of which the working copies may differ by the lattice vectors of the wrap.
struct MdRec {
  const PairRecord* rec;
  const float4* pb;
  const float4* sp;
  const SysGeom* g;
  const float* caller_pos;
  int i0;
};

__device__ __attribute__((noinline)) void md_record(const MdRec& r, int slot, int t, int s0, int s1,
                                                    int s2) {
  const int gi = r.i0 + __float_as_int(r.pb[slot].w), gj = r.i0 + __float_as_int(r.pb[t].w);
  if (r.caller_pos) {
    int a0, a1, a2, b0, b1, b2;
    const float4 ri = r.sp[slot], rj = r.sp[t];
    lattice_offset(*r.g, r.caller_pos + 3 * (size_t)gi, ri.x, ri.y, ri.z, a0, a1, a2);
    lattice_offset(*r.g, r.caller_pos + 3 * (size_t)gj, rj.x, rj.y, rj.z, b0, b1, b2);
    s0 += a0 - b0; s1 += a1 - b1; s2 += a2 - b2;
  }
  record_pair(*r.rec, gi, gj, s0, s1, s2, 1);
}

__device__ __attribute__((noinline)) bool md_band_shift(const SysGeom& g, float4 ri, float4 rj,
                                                        int sx, int sy, int sz) {
  return band_keep(g, ri, rj, sx, sy, sz);
}

// Forces on the atom in sorted slot `slot` from every candidate of its cell
// stencil.  Atoms are sorted by cell index (z ny + y) nx + x, so for a
// stencil row (dz, dy) the cells cx-h .. cx+h of one image shift sx are one
// contiguous run of sp[]: the row is walked as one run per image.
template <bool MULTI, bool VIR, bool REC = false>
__device__ inline void md_stencil_forces(const SysGeom& g, const float4* sp, const int* cstart,
                                         const float4* s_par, int ns, int slot, float rc2_lo,
                                         float rc2_hi, float& ax, float& ay, float& az, float& ei,
                                         float (&w)[6], const MdRec* rec = nullptr) {
  const float4 ri = sp[slot];
  const int wi = __float_as_int(ri.w);
  const int c = (wi >> 8) & 0xfff;
  const int si = wi & 0xff;
  const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];
  const float ix = 1.0f / (float)ncx, iy = 1.0f / (float)ncy, iz = 1.0f / (float)ncz;
  const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
  const int h0 = g.h[0], h1 = g.h[1], h2 = g.h[2];
  const float4 p0 = s_par[0];
  const int sx_lo = floor_div_small(cx - h0, ix), sx_hi = floor_div_small(cx + h0, ix);
  for (int dz = -h2; dz <= h2; ++dz) {
    int zz = cz + dz;
    const int sz = floor_div_small(zz, iz);
    zz -= sz * ncz;
    for (int dy = -h1; dy <= h1; ++dy) {
      int yy = cy + dy;
      const int sy = floor_div_small(yy, iy);
      yy -= sy * ncy;
      const int row = (zz * ncy + yy) * ncx;
      for (int sx = sx_lo; sx <= sx_hi; ++sx) {
        const int x0 = max(cx - h0 - sx * ncx, 0), x1 = min(cx + h0 - sx * ncx, ncx - 1);
        float Sx, Sy, Sz;
        tmul(g.Lf, (float)sx, (float)sy, (float)sz, Sx, Sy, Sz);
        const float qx = ri.x - Sx, qy = ri.y - Sy, qz = ri.z - Sz;
        const int self = (sx | sy | sz) == 0 ? slot : -1;
        const int tb = cstart[row + x0], te = cstart[row + x1 + 1];
        for (int t = tb; t < te; ++t) {
          const float4 rj = sp[t];
          const float ex = rj.x - qx, ey = rj.y - qy, ez = rj.z - qz;
          const float r2 = ex * ex + ey * ey + ez * ez;
          const bool lo = r2 < rc2_lo;
          bool keep = lo && t != self;
          if (__builtin_expect(!lo && r2 < rc2_hi, 0)) keep = md_band_shift(g, ri, rj, sx, sy, sz);
          if constexpr (REC) {
            if (keep) md_record(*rec, slot, t, sx, sy, sz);
          } else {
            const float4 pp = MULTI ? s_par[si * ns + (__float_as_int(rj.w) & 0xff)] : p0;
            md_lj_terms<VIR>(ex, ey, ez, r2, keep, pp, ax, ay, az, ei, w);
          }
        }
      }
    }
  }
}

// List entries per atom as allocated: the capacity rounded up to whole
// groups of four.
__host__ __device__ inline int md_list_cap(int nbmax) { return (nbmax + 3) & ~3; }

// Word index of list entry e of slot s: entries are grouped by four, the four
// entries of a group contiguous (one 8-byte load), groups entry-major over
// the slots (a wave's loads of one group are contiguous).
__device__ inline size_t md_list_word(int e, int s) {
  return ((size_t)(e >> 2) * MD_NMAX + s) * 4 + (e & 3);
}

// Verlet list of the atom in sorted slot `slot`: every atom j with an image
// within the list radius, once.  When every axis has >= 5 cells, the +-2-cell
// stencil visits each (j, image) once and distinct images of j cannot both be
// within rc + skin <= perp / 2, so the stencil image is tested directly
// (DEDUP = false).  Otherwise the stencil may visit j under several images;
// only the visit whose shift is the per-axis nearest image,
// s = -rint((r_j - r_i) @ L^-1) (one FP32 evaluation per pair, the same for
// every visit), is kept, and it is kept when that image lies within the list
// radius (orthorhombic cells, where it is the nearest image) or within the
// slab |f - s| < (rc + skin) / perp on every axis (triclinic cells, a
// superset of the sphere).  Either way an atom within the cutoff at a later
// step (every atom moved by less than skin/2) is in the list.  The last group
// is padded with the atom's own slot (masked off by the count).  Returns the
// count, which may exceed cap (entries past cap are not written).
template <bool DEDUP>
__device__ inline int md_build_list(const SysGeom& g, const float4* sp, const int* cstart, int slot,
                                    unsigned short* nbl, int cap, float list_r2, bool ortho,
                                    float fx_max, float fy_max, float fz_max) {
  const float4 ri = sp[slot];
  const int c = (__float_as_int(ri.w) >> 8) & 0xfff;
  const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];
  const float ix = 1.0f / (float)ncx, iy = 1.0f / (float)ncy, iz = 1.0f / (float)ncz;
  const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
  const int h0 = g.h[0], h1 = g.h[1], h2 = g.h[2];
  const int sx_lo = floor_div_small(cx - h0, ix), sx_hi = floor_div_small(cx + h0, ix);
  int cnt = 0;
  for (int dz = -h2; dz <= h2; ++dz) {
    int zz = cz + dz;
    const int sz = floor_div_small(zz, iz);
    zz -= sz * ncz;
    for (int dy = -h1; dy <= h1; ++dy) {
      int yy = cy + dy;
      const int sy = floor_div_small(yy, iy);
      yy -= sy * ncy;
      const int row = (zz * ncy + yy) * ncx;
      for (int sx = sx_lo; sx <= sx_hi; ++sx) {
        const int x0 = max(cx - h0 - sx * ncx, 0), x1 = min(cx + h0 - sx * ncx, ncx - 1);
        float Sx, Sy, Sz;
        tmul(g.Lf, (float)sx, (float)sy, (float)sz, Sx, Sy, Sz);
        const float qx = ri.x - Sx, qy = ri.y - Sy, qz = ri.z - Sz;
        const int tb = cstart[row + x0], te = cstart[row + x1 + 1];
        for (int t = tb; t < te; ++t) {
          if (t == slot) continue;  // self images are never within 2 rc
          const float4 rj = sp[t];
          const float ex = rj.x - qx, ey = rj.y - qy, ez = rj.z - qz;
          const float r2 = ex * ex + ey * ey + ez * ez;
          bool take;
          if (DEDUP) {
            float f0, f1, f2;
            tmul(g.Lif, rj.x - ri.x, rj.y - ri.y, rj.z - ri.z, f0, f1, f2);
            const float n0 = rintf(f0), n1 = rintf(f1), n2 = rintf(f2);
            const bool nearest = (int)n0 == -sx && (int)n1 == -sy && (int)n2 == -sz;
            const bool near = ortho ? r2 < list_r2
                                    : (fabsf(f0 - n0) < fx_max && fabsf(f1 - n1) < fy_max &&
                                       fabsf(f2 - n2) < fz_max);
            take = nearest && near;
          } else {
            take = r2 < list_r2;
          }
          if (take) {
            if (cnt < cap) nbl[md_list_word(cnt, slot)] = (unsigned short)t;
            ++cnt;
          }
        }
      }
    }
  }
  for (int e = cnt; e < min((cnt + 3) & ~3, cap); ++e) nbl[md_list_word(e, slot)] = (unsigned short)slot;
  return cnt;
}

// Forces on the atom in sorted slot `slot` from its Verlet list: each entry
// on its per-axis nearest image (the only image within the cutoff when every
// perpendicular width is at least 2 rc).  Four entries per trip, their list
// words and positions loaded before any is used.
template <bool MULTI, bool VIR, bool ORTHO, bool REC = false>
__device__ inline void md_list_forces(const SysGeom& g, const float4* sp, const unsigned short* nbl,
                                      int cnt, int slot, const float4* s_par, int ns,
                                      float rc2_lo, float rc2_hi, float& ax, float& ay, float& az,
                                      float& ei, float (&w)[6], const MdRec* rec = nullptr) {
  const float4 ri = sp[slot];
  const int si = __float_as_int(ri.w) & 0xff;
  const float4 p0 = s_par[0];
  // group q of this slot: one 8-byte word of four entries; the next group's
  // word is loaded one trip ahead
  const uint2* nl = reinterpret_cast<const uint2*>(nbl) + slot;
  const int ng = (cnt + 3) >> 2;
  uint2 cur = ng > 0 ? nl[0] : make_uint2(0u, 0u);
  for (int q = 0; q < ng; ++q) {
    const uint2 nxt = nl[(size_t)min(q + 1, ng - 1) * MD_NMAX];
    const int j[4] = {(int)(cur.x & 0xffffu), (int)(cur.x >> 16), (int)(cur.y & 0xffffu),
                      (int)(cur.y >> 16)};
    float4 rj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) rj[u] = sp[j[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      float ex = rj[u].x - ri.x, ey = rj[u].y - ri.y, ez = rj[u].z - ri.z;
      // nearest image n (fractional units); the edge shift is -n
      float n0, n1, n2;
      if (ORTHO) {
        n0 = rintf(ex * g.Lif[0]); n1 = rintf(ey * g.Lif[2]); n2 = rintf(ez * g.Lif[5]);
        ex -= n0 * g.Lf[0];
        ey -= n1 * g.Lf[2];
        ez -= n2 * g.Lf[5];
      } else {
        float f0, f1, f2;
        tmul(g.Lif, ex, ey, ez, f0, f1, f2);
        n0 = rintf(f0); n1 = rintf(f1); n2 = rintf(f2);
        float sx, sy, sz;
        tmul(g.Lf, n0, n1, n2, sx, sy, sz);
        ex -= sx; ey -= sy; ez -= sz;
      }
      const float r2 = ex * ex + ey * ey + ez * ez;
      const bool valid = 4 * q + u < cnt;
      const bool lo = r2 < rc2_lo;
      bool keep = valid && lo;
      const bool band = valid && !lo && r2 < rc2_hi;
      if (__builtin_expect(band, 0))
        keep = md_band_nearest(g, ri.x, ri.y, ri.z, rj[u].x, rj[u].y, rj[u].z);
      if constexpr (REC) {
        if (keep) {
          int s0 = -(int)n0, s1 = -(int)n1, s2 = -(int)n2;
          if (band) md_nearest64(g, ri.x, ri.y, ri.z, rj[u].x, rj[u].y, rj[u].z, s0, s1, s2);
          md_record(*rec, slot, j[u], s0, s1, s2);
        }
      } else {
        const float4 pp = MULTI ? s_par[si * ns + (__float_as_int(rj[u].w) & 0xff)] : p0;
        md_lj_terms<VIR>(ex, ey, ez, r2, keep, pp, ax, ay, az, ei, w);
      }
    }
    cur = nxt;
  }
}

// Forces, closing B(dt/2) (+ the next step's opening B(dt/2)) and kinetic
// energy of every slot of the thread; the forces of the last step go to HBM.
template <bool MULTI, bool VIR, bool ORTHO>
__device__ inline void md_force_phase(const MdArgs& a, const SysGeom& g, const float4* sp,
                                      float4* pb, const int* cstart, const short* nbcount,
                                      const unsigned short* nbl, bool use_list,
                                      const float4* s_par, int ns, int n, int i0, float rc2_lo,
                                      float rc2_hi, float hdt, float dtf, bool last, bool split,
                                      bool kick_now, float& epot, float& ekin, float (&w)[6]) {
  const int tid = threadIdx.x;
#pragma unroll 1
  for (int k = 0; k < MD_SLOTS; ++k) {
    const int slot = tid + k * MD_THREADS;
    if (slot >= n) continue;
    float ax = 0.f, ay = 0.f, az = 0.f, ei = 0.f;
    if (use_list)
      md_list_forces<MULTI, VIR, ORTHO>(g, sp, nbl, nbcount[slot], slot, s_par, ns, rc2_lo, rc2_hi,
                                        ax, ay, az, ei, w);
    else
      md_stencil_forces<MULTI, VIR>(g, sp, cstart, s_par, ns, slot, rc2_lo, rc2_hi, ax, ay, az, ei,
                                    w);
    epot += 0.5f * ei;
    float4 p = pb[slot];
    const int gi = i0 + __float_as_int(p.w);
    if (kick_now && (a.n_steps > 0 || a.kin)) {
      const float im = 1.0f / a.mass[gi];
      const float vx = p.x + ax * hdt, vy = p.y + ay * hdt, vz = p.z + az * hdt;
      ekin += 0.5f * (vx * vx + vy * vy + vz * vz) * im;
    }
    if (kick_now && a.n_steps > 0) {
      const float kick = last || split ? hdt : dtf;
      p.x += ax * kick; p.y += ay * kick; p.z += az * kick;
      pb[slot] = p;
    }
    if (last || split) {
      a.forces[3 * gi] = ax; a.forces[3 * gi + 1] = ay; a.forces[3 * gi + 2] = az;
    }
  }
}

// The pair set of the current state, walked exactly as the force phase walks
// it (same lists or stencil, same decisions), appended to a.rec.
template <bool ORTHO>
__device__ inline void md_record_phase(const MdArgs& a, const SysGeom& g, const float4* sp,
                                       const float4* pb, const int* cstart, const short* nbcount,
                                       const unsigned short* nbl, bool use_list, const float4* s_par,
                                       int ns, int n, int i0, float rc2_lo, float rc2_hi) {
  const int tid = threadIdx.x;
  const MdRec rec{&a.rec, pb, sp, &g, a.n_steps == 0 ? a.pos : nullptr, i0};
  float ax = 0.f, ay = 0.f, az = 0.f, ei = 0.f;
  float w[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
  for (int k = 0; k < MD_SLOTS; ++k) {
    const int slot = tid + k * MD_THREADS;
    if (slot >= n) continue;
    if (use_list)
      md_list_forces<false, false, ORTHO, true>(g, sp, nbl, nbcount[slot], slot, s_par, ns, rc2_lo,
                                                rc2_hi, ax, ay, az, ei, w, &rec);
    else
      md_stencil_forces<false, false, true>(g, sp, cstart, s_par, ns, slot, rc2_lo, rc2_hi, ax, ay,
                                            az, ei, w, &rec);
  }
}

// Maximum of v over the workgroup, returned in every thread.
__device__ inline float block_max_all(float v, double* red) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  __syncthreads();
  if (lane == 0) red[wv] = (double)v;
  __syncthreads();
  double m = red[0];
#pragma unroll
  for (int k = 1; k < MD_THREADS / 64; ++k) m = fmax(m, red[k]);
  return (float)m;
}

// buf[dst[k]] = buf[tid + k * MD_THREADS] for every dst[k] >= 0, in place:
// every read completes at the barrier before any write.
// The copies use a native 4-float vector: a float4 struct copy becomes a
// memcpy through private memory that is not promoted to registers.
__device__ inline void permute_slots(float4* buf, const int (&dst)[MD_SLOTS], int tid) {
  typedef float v4f __attribute__((ext_vector_type(4)));
  v4f* vb = reinterpret_cast<v4f*>(buf);
  v4f t[MD_SLOTS];
#pragma unroll
  for (int k = 0; k < MD_SLOTS; ++k) t[k] = vb[min(tid + k * MD_THREADS, MD_NMAX - 1)];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < MD_SLOTS; ++k)
    if (dst[k] >= 0) vb[dst[k]] = t[k];
  __syncthreads();
}

// THERMO: 0 for velocity Verlet / BAOAB, 1 for CSVR, 2 for CSVR + cell
// rescaling; each is its own instance, so the plain integrators carry none of
// the thermostat's and barostat's state.
// REC: the instantiation that also writes the pair set of the last step's
// final force pass to a.rec (parity output); the MD itself is the same code.
template <int THERMO, bool REC>
__global__ __launch_bounds__(MD_THREADS) void md_lj_kernel(MdArgs a) {
  __shared__ float4 sp[MD_NMAX];   // sorted x,y,z ; w = bits(species | cell << 8)
  __shared__ float4 pb[MD_NMAX];   // sorted px,py,pz ; w = bits(local id)
  // while sorting: previous slot of each provisional entry; between list
  // builds: the Verlet list length of each slot
  __shared__ short keys[MD_NMAX];
  __shared__ int cstart[MD_NCMAX + 1];
  __shared__ int ccount[MD_NCMAX];
  __shared__ int wsum[MD_THREADS / 64];
  __shared__ double red[MD_THREADS / 64];
  __shared__ float4 s_par[MD_NSPEC * MD_NSPEC];  // sigma^2, 4 eps, 24 eps
  __shared__ SysGeom g;
  __shared__ int s_q;
  __shared__ float s_alpha;
  __shared__ double s_mu, s_tril[6];

  const int tid = threadIdx.x;
  const int ns = a.n_species;
  unsigned short* nbl =
      a.nblist ? a.nblist + (size_t)blockIdx.x * (size_t)md_list_cap(a.nbmax) * MD_NMAX : nullptr;
  for (int k = tid; k < ns * ns; k += MD_THREADS) {
    double sg = a.sigma[k];
    s_par[k] = make_float4((float)(sg * sg), (float)(4.0 * a.epsilon[k]),
                           (float)(24.0 * a.epsilon[k]), 0.0f);
  }

  for (;;) {
    __syncthreads();
    if (tid == 0) s_q = atomicAdd(a.queue, 1);
    __syncthreads();
    const int q = s_q;
    if (q >= a.n_sys) break;
    const int s = a.order[q];
    const int i0 = a.sys_off[s];
    const int n = a.sys_off[s + 1] - i0;
    if (n <= 0) continue;
    if (n > MD_NMAX) {  // the host checks sizes; never drop atoms silently
      if (tid == 0 && a.pot) a.pot[(size_t)s * (a.n_steps > 0 ? a.n_steps : 1)] = __builtin_nan("");
      continue;
    }
    {
      const int* src = reinterpret_cast<const int*>(a.geom + s);
      int* dstw = reinterpret_cast<int*>(&g);
      for (int w = tid; w < (int)(sizeof(SysGeom) / sizeof(int)); w += MD_THREADS) dstw[w] = src[w];
    }

    // ---- load the system: state in LDS (slot = local atom index).  Forces
    // are not kept between steps: the closing B(dt/2) of a step and the
    // opening B(dt/2) of the next are applied together as one kick right
    // after the force evaluation; only the first step reads F(t0).
#pragma unroll
    for (int k = 0; k < MD_SLOTS; ++k) {
      const int li = tid + k * MD_THREADS;
      if (li < n) {
        const int gi = i0 + li;
        sp[li] = make_float4(a.pos[3 * gi], a.pos[3 * gi + 1], a.pos[3 * gi + 2],
                             __int_as_float(a.species[gi]));
        pb[li] = make_float4(a.mom[3 * gi], a.mom[3 * gi + 1], a.mom[3 * gi + 2], __int_as_float(li));
      }
    }
    __syncthreads();
    const float rc2 = g.rc2f;
    const float rc2_hi = rc2 * (1.0f + CUT_BAND), rc2_lo = rc2 * (1.0f - CUT_BAND);
    const float dtf = a.n_steps > 0 ? (float)a.dt[s] : 0.0f;
    const float hdt = 0.5f * dtf;
    float c1 = 1.0f, c2 = 0.0f;
    // CSVR (integrator 2): the thermostat scales the momenta at the start of
    // every step, between the previous step's closing half kick and this
    // step's opening one, so the two half kicks are not fused and F goes to
    // HBM every step
    const bool npt = THERMO == 2 && a.n_steps > 0;
    const bool csvr = THERMO >= 1 && a.n_steps > 0;
    if (npt && tid < 6) s_tril[tid] = a.tril[6 * (size_t)s + tid];
    double kin_prev = 0.0;  // K after the previous closing kick (thread 0)
    if (a.integrator == 1 && a.n_steps > 0) {
      double e1 = exp(-a.gamma[s] * a.dt[s]);
      c1 = (float)e1;
      c2 = (float)sqrt(a.kT[s] * (1.0 - e1 * e1));
    }
    // Verlet lists for boxes whose perpendicular widths are all >= 2 rc
    // (one image per pair within the cutoff) and runs of >= 2 steps
    bool use_list = nbl != nullptr && a.skin > 0.0f && g.rep == 0 && a.n_steps >= 2 && !npt;
    bool listed = false;
    const double lr = g.rc + (double)a.skin;
    const float list_r2 = (float)(lr * lr) * (1.0f + 1e-5f);
    const bool ortho = g.L[1] == 0.0 && g.L[3] == 0.0 && g.L[4] == 0.0;
    const bool dedup = g.nc[0] < 5 || g.nc[1] < 5 || g.nc[2] < 5;
    // triclinic list slab: |f| < (rc + skin) / perp per axis, perp = 1 / |row of L^-T|
    const float fx_max = (float)(lr * sqrt(g.Li[0] * g.Li[0] + g.Li[1] * g.Li[1] + g.Li[3] * g.Li[3])) * 1.00001f;
    const float fy_max = (float)(lr * sqrt(g.Li[2] * g.Li[2] + g.Li[4] * g.Li[4])) * 1.00001f;
    const float fz_max = (float)(lr * fabs(g.Li[5])) * 1.00001f;
    // per slot: sum of the step displacements since the list was built, an
    // upper bound of the displacement
    float dacc[MD_SLOTS];
#pragma unroll
    for (int k = 0; k < MD_SLOTS; ++k) dacc[k] = 0.0f;

    const int total = a.n_steps > 0 ? a.n_steps : 1;
    for (int st = 0; st < total; ++st) {
      const bool last = st == total - 1;
      // every force loop of the previous step has finished reading sp[]
      __syncthreads();
      // reload the LDS-resident geometry in every step instead of keeping
      // the hoisted copies live across the whole step loop (and because the
      // barostat changes it)
      asm volatile("" ::: "memory");
      const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2], ncell = g.ncell;
      float alpha = 1.0f;
      if (csvr) {
        double K = kin_prev;
        if (st == 0) {
          float k0 = 0.0f;
#pragma unroll
          for (int k = 0; k < MD_SLOTS; ++k) {
            const int slot = tid + k * MD_THREADS;
            if (slot >= n) continue;
            const float4 p = pb[slot];
            k0 += 0.5f * (p.x * p.x + p.y * p.y + p.z * p.z) / a.mass[i0 + __float_as_int(p.w)];
          }
          K = block_sum((double)k0, red);
        }
        if (tid == 0)
          s_alpha = (float)csvr_alpha(K, n, a.kT[s], a.dt[s], a.tau[s], a.seed, (uint32_t)s,
                                      a.step0 + st);
        __syncthreads();
        alpha = s_alpha;
      }
      // ---- [B(dt/2) at st = 0] A(...) [O A] in place, wrap, cell of every atom
      int cel[MD_SLOTS];
      float dmax = 0.0f;
#pragma unroll
      for (int k = 0; k < MD_SLOTS; ++k) {
        const int slot = tid + k * MD_THREADS;
        cel[k] = 0;
        if (slot >= n) continue;
        float4 r = sp[slot];
        if (a.n_steps > 0) {
          float4 p = pb[slot];
          const int gi = i0 + __float_as_int(p.w);
          const float im = 1.0f / a.mass[gi];
          if (csvr) {
            p.x = alpha * p.x + a.forces[3 * gi] * hdt;
            p.y = alpha * p.y + a.forces[3 * gi + 1] * hdt;
            p.z = alpha * p.z + a.forces[3 * gi + 2] * hdt;
          } else if (st == 0) {
            p.x += a.forces[3 * gi] * hdt;
            p.y += a.forces[3 * gi + 1] * hdt;
            p.z += a.forces[3 * gi + 2] * hdt;
          }
          const float ta = a.integrator == 1 ? hdt : dtf;
          r.x += p.x * im * ta; r.y += p.y * im * ta; r.z += p.z * im * ta;
          float moved = sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * im * ta;
          if (a.integrator == 1) {
            float f0, f1, f2;
            wrap_shift(g.L, g.Lif, r.x, r.y, r.z, f0, f1, f2);
            float e0, e1, e2;
            normals3(a.seed, (uint32_t)gi, a.step0 + st, e0, e1, e2);
            const float sm = c2 * sqrtf(1.0f / im);
            p.x = c1 * p.x + sm * e0;
            p.y = c1 * p.y + sm * e1;
            p.z = c1 * p.z + sm * e2;
            r.x += p.x * im * hdt; r.y += p.y * im * hdt; r.z += p.z * im * hdt;
            moved += sqrtf(p.x * p.x + p.y * p.y + p.z * p.z) * im * hdt;
          }
          pb[slot] = p;
          dacc[k] += moved;
          dmax = fmaxf(dmax, dacc[k]);
        }
        float f0, f1, f2;
        wrap_shift(g.L, g.Lif, r.x, r.y, r.z, f0, f1, f2);
        cel[k] = (cell_index(f2, ncz) * ncy + cell_index(f1, ncy)) * ncx + cell_index(f0, ncx);
        r.w = __int_as_float((__float_as_int(r.w) & 0xff) | (cel[k] << 8));
        sp[slot] = r;
      }
      // ---- stable counting sort into cells (rank by previous slot), slots
      // permuted in place
      auto sort_cells = [&](const int (&cel)[MD_SLOTS], int ncell) {
        __syncthreads();
        for (int c = tid; c < ncell; c += MD_THREADS) ccount[c] = 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MD_SLOTS; ++k)
          if (tid + k * MD_THREADS < n) atomicAdd(&ccount[cel[k]], 1);
        __syncthreads();
        block_scan_cells(ccount, cstart, ncell, wsum);
        __syncthreads();
        for (int c = tid; c < ncell; c += MD_THREADS) ccount[c] = 0;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MD_SLOTS; ++k) {
          const int slot = tid + k * MD_THREADS;
          if (slot < n) keys[cstart[cel[k]] + atomicAdd(&ccount[cel[k]], 1)] = (short)slot;
        }
        __syncthreads();
        // destination of every slot, then the positions and the momenta
        // permuted one array at a time: only MD_SLOTS float4 are held in
        // registers across a barrier (the 128-VGPR budget of a 1024-thread
        // workgroup has no room for both)
        int dst[MD_SLOTS];
#pragma unroll
        for (int k = 0; k < MD_SLOTS; ++k) {
          const int slot = tid + k * MD_THREADS;
          dst[k] = -1;
          if (slot < n) {
            const int b = cstart[cel[k]], e = cstart[cel[k] + 1];
            int rank = 0;
            for (int t = b; t < e; ++t) rank += keys[t] < slot;
            dst[k] = b + rank;
          }
        }
        permute_slots(sp, dst, tid);
        permute_slots(pb, dst, tid);
      };
      // the lists stay valid while no atom can have moved by skin/2
      bool sort_now = true;
      if (use_list && listed) sort_now = block_max_all(dmax, red) > 0.5f * a.skin;
      if (sort_now) {
        sort_cells(cel, ncell);
        if (use_list) {
          // ---- Verlet lists of the new order; an overflowing list turns
          // the lists off for the rest of the run (the stencil is exact)
          int over = 0;
#pragma unroll 1
          for (int k = 0; k < MD_SLOTS; ++k) {
            const int slot = tid + k * MD_THREADS;
            dacc[k] = 0.0f;
            if (slot >= n) continue;
            const int cap = md_list_cap(a.nbmax);
            const int cnt = dedup ? md_build_list<true>(g, sp, cstart, slot, nbl, cap, list_r2, ortho,
                                                        fx_max, fy_max, fz_max)
                                  : md_build_list<false>(g, sp, cstart, slot, nbl, cap, list_r2,
                                                         ortho, fx_max, fy_max, fz_max);
            keys[slot] = (short)min(cnt, cap);
            over |= cnt > cap;
          }
          use_list = !__syncthreads_or(over);
          listed = use_list;
        }
      }
      // ---- forces, kick and kinetic energy; with the barostat a second pass
      // recomputes forces and energies at the rescaled state (one call site,
      // so each pair-loop variant is inlined once)
      float epot = 0.0f, ekin = 0.0f;
      for (int pass = 0; pass < (npt ? 2 : 1); ++pass) {
        const bool kick_now = pass == 0;
        // NPT needs the virial of every first pass (the barostat's pressure)
        const bool do_w = (last && a.want_virial && pass + 1 == (npt ? 2 : 1)) || (npt && pass == 0);
        float w[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        epot = 0.0f;
#define MD_PHASE(M, V, O)                                                                      \
  md_force_phase<M, V, O>(a, g, sp, pb, cstart, keys, nbl, use_list, s_par, ns, n, i0, rc2_lo, \
                          rc2_hi, hdt, dtf, last, csvr, kick_now, epot, ekin, w)
        if (ns == 1) {
          if (do_w) { if (ortho) MD_PHASE(false, true, true); else MD_PHASE(false, true, false); }
          else { if (ortho) MD_PHASE(false, false, true); else MD_PHASE(false, false, false); }
        } else {
          if (do_w) { if (ortho) MD_PHASE(true, true, true); else MD_PHASE(true, true, false); }
          else { if (ortho) MD_PHASE(true, false, true); else MD_PHASE(true, false, false); }
        }
#undef MD_PHASE
        if (kick_now && (a.kin || csvr)) {
          double e = block_sum((double)ekin, red);
          if (tid == 0 && a.kin) a.kin[(size_t)s * total + st] = e;
          kin_prev = e;
        }
        // the system virial (valid in thread 0): the output of the last
        // step's final pass, and the pressure of the barostat
        double ptr_w = 0.0;
        if (do_w) {
          double wsys[6];
#pragma unroll
          for (int c = 0; c < 6; ++c) wsys[c] = block_sum((double)w[c], red);
          ptr_w = wsys[0] + wsys[3] + wsys[5];
          if (tid == 0 && last && a.want_virial && a.virial && pass + 1 == (npt ? 2 : 1)) {
            double* W = a.virial + 9 * (size_t)s;
            W[0] = wsys[0]; W[1] = wsys[1]; W[2] = wsys[2];
            W[3] = wsys[1]; W[4] = wsys[3]; W[5] = wsys[4];
            W[6] = wsys[2]; W[7] = wsys[4]; W[8] = wsys[5];
          }
        }
        if (npt && pass == 0) {
          // ---- stochastic cell rescaling (Bernetti and Bussi 2020;
          // md/integrators.py:657-811): P = 2K / (3V) + tr(W) / (3V),
          // d eps = (dt / tau_P) beta (P - P0) + sqrt(2 kT beta dt / (tau_P V)) R
          // with ONE R ~ N(0, 1) for the batch, mu = exp(d eps / 3) clipped to
          // [mu_min, 1 / mu_min]; cell and positions scaled by mu (momenta
          // kept), then the second pass at the scaled state
          if (tid == 0) {
            const double* bp = a.baro + 4 * (size_t)s;
            const double V = fabs(s_tril[0] * s_tril[2] * s_tril[5]);
            const double P = 2.0 * kin_prev / (3.0 * V) + ptr_w / (3.0 * V);
            uint32_t c4[4];
            csvr_draw(a.seed, 0xFFFFFFFEu, a.step0 + st, SCR_TAG, c4);
            const double R = csvr_bm(c4[0], c4[1]);
            const double dt = a.dt[s], kT = a.kT[s];
            const double de = dt / bp[1] * bp[2] * (P - bp[0]) +
                              sqrt(2.0 * kT * bp[2] * dt / (bp[1] * V)) * R;
            const double mu = fmin(fmax(exp(de / 3.0), bp[3]), 1.0 / bp[3]);
            s_mu = mu;
            for (int c = 0; c < 6; ++c) s_tril[c] *= mu;
            make_geom(s_tril, g.rc, MD_MAX_NC_AXIS, g, MD_SUB);
          }
          __syncthreads();
          asm volatile("" ::: "memory");
          const float mu = (float)s_mu;
          const int nx = g.nc[0], ny = g.nc[1], nz = g.nc[2];
#pragma unroll
          for (int k = 0; k < MD_SLOTS; ++k) {
            const int slot = tid + k * MD_THREADS;
            cel[k] = 0;
            if (slot >= n) continue;
            float4 r = sp[slot];
            r.x *= mu; r.y *= mu; r.z *= mu;
            float f0, f1, f2;
            wrap_shift(g.L, g.Lif, r.x, r.y, r.z, f0, f1, f2);
            cel[k] = (cell_index(f2, nz) * ny + cell_index(f1, ny)) * nx + cell_index(f0, nx);
            r.w = __int_as_float((__float_as_int(r.w) & 0xff) | (cel[k] << 8));
            sp[slot] = r;
          }
          sort_cells(cel, g.ncell);
        }
      }
      if (a.pot) {
        double e = block_sum((double)epot, red);
        if (tid == 0) a.pot[(size_t)s * total + st] = e;
      }
      if (REC && last) {
        __syncthreads();
        if (ortho)
          md_record_phase<true>(a, g, sp, pb, cstart, keys, nbl, use_list, s_par, ns, n, i0,
                                rc2_lo, rc2_hi);
        else
          md_record_phase<false>(a, g, sp, pb, cstart, keys, nbl, use_list, s_par, ns, n, i0,
                                 rc2_lo, rc2_hi);
      }
    }
    if (npt && tid == 0)
      for (int c = 0; c < 6; ++c) a.tril_out[6 * (size_t)s + c] = s_tril[c];
    // ---- store the system back in its original atom order (forces were
    // written by the last force pass; every slot is stored by the thread
    // that kicked it)
    if (a.n_steps > 0) {
#pragma unroll
      for (int k = 0; k < MD_SLOTS; ++k) {
        const int slot = tid + k * MD_THREADS;
        if (slot >= n) continue;
        const float4 r = sp[slot];
        const float4 p = pb[slot];
        const int gi = i0 + __float_as_int(p.w);
        a.pos[3 * gi] = r.x; a.pos[3 * gi + 1] = r.y; a.pos[3 * gi + 2] = r.z;
        a.mom[3 * gi] = p.x; a.mom[3 * gi + 1] = p.y; a.mom[3 * gi + 2] = p.z;
      }
    }
  }
}

__global__ void fill_ones_zeros(float* o, float* z, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = 1.0f;
  if (i < 3 * n) z[i] = 0.0f;
}

__global__ void philox_kernel(int n, const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c[4] = {ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]};
  philox4x32_10(c, key[2 * i], key[2 * i + 1]);
  for (int k = 0; k < 4; ++k) out[4 * i + k] = c[k];
}

__global__ void normals_kernel(uint64_t seed, int64_t step, int n, float* eta) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a, b, c;
  normals3(seed, (uint32_t)i, step, a, b, c);
  eta[3 * i] = a; eta[3 * i + 1] = b; eta[3 * i + 2] = c;
}

static int md_launch(kupsmi_ctx* ctx, hipStream_t stream, MdArgs& a) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  KUPSMI_REQUIRE(a.n_sys >= 0, "n_sys < 0");
  KUPSMI_REQUIRE(a.n_species >= 1 && a.n_species <= MD_NSPEC, "n_species must be in [1, 8]");
  KUPSMI_REQUIRE(a.integrator >= 0 && a.integrator <= 3,
                 "integrator must be 0 (VV), 1 (BAOAB), 2 (CSVR) or 3 (CSVR + cell rescaling)");
  // the barostat rebuilds the cell list every step: no Verlet lists
  if (a.integrator == 3) a.skin = 0.0f;
  KUPSMI_REQUIRE(a.n_steps >= 0, "n_steps < 0");
  // a pending pair record is taken by this call
  if (ctx->rec_armed) {
    a.rec = ctx->rec;
    ctx->rec_armed = 0;
    record_clear_kernel<<<1, 1, 0, stream>>>(a.rec.count);
    KUPSMI_CHECK_LAUNCH();
  }
  if (a.n_sys == 0) return KUPSMI_OK;
  void* buf = nullptr;
  const size_t geom_bytes = sizeof(SysGeom) * (size_t)a.n_sys;
  int rc = scratch(ctx, 0, geom_bytes + sizeof(int32_t) * (size_t)(a.n_sys + 64), &buf);
  if (rc) return rc;
  SysGeom* geom = static_cast<SysGeom*>(buf);
  int32_t* order = reinterpret_cast<int32_t*>(static_cast<char*>(buf) + geom_bytes);
  int* queue = reinterpret_cast<int*>(order + a.n_sys);
  order_kernel<<<1, 1024, 0, stream>>>(a.n_sys, a.sys_off, order, queue);
  KUPSMI_CHECK_LAUNCH();
  geom_kernel<<<(a.n_sys + 255) / 256, 256, 0, stream>>>(a.n_sys, a.tril, a.cutoff, MD_MAX_NC_AXIS,
                                                         (double)a.skin, geom);
  KUPSMI_CHECK_LAUNCH();
  a.order = order;
  a.queue = queue;
  a.geom = geom;
  int dev = 0, ncu = 256;
  KUPSMI_HIP(hipGetDevice(&dev));
  KUPSMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int grid = a.n_sys < ncu ? a.n_sys : ncu;
  a.nblist = nullptr;
  if (a.skin > 0.0f && a.nbmax > 0 && a.n_steps >= 2) {
    // one list region (nbmax entries x MD_NMAX slots) per resident workgroup
    void* lb = nullptr;
    rc = scratch(ctx, 6, sizeof(unsigned short) * (size_t)grid * md_list_cap(a.nbmax) * MD_NMAX,
                 &lb);
    if (rc) return rc;
    a.nblist = static_cast<unsigned short*>(lb);
  }
  const bool rec = a.rec.count != nullptr;
  if (a.integrator == 3) {
    if (rec) md_lj_kernel<2, true><<<grid, MD_THREADS, 0, stream>>>(a);
    else md_lj_kernel<2, false><<<grid, MD_THREADS, 0, stream>>>(a);
  } else if (a.integrator == 2) {
    if (rec) md_lj_kernel<1, true><<<grid, MD_THREADS, 0, stream>>>(a);
    else md_lj_kernel<1, false><<<grid, MD_THREADS, 0, stream>>>(a);
  } else {
    if (rec) md_lj_kernel<0, true><<<grid, MD_THREADS, 0, stream>>>(a);
    else md_lj_kernel<0, false><<<grid, MD_THREADS, 0, stream>>>(a);
  }
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_md_lj(kupsmi_ctx* ctx, void* stream, int integrator, int n_steps,
                            int64_t step0, uint64_t seed, int n_sys, const int32_t* sys_off,
                            int n_atoms, float* pos, float* mom, float* forces,
                            const float* mass, const int32_t* species, int n_species,
                            const double* sigma, const double* epsilon, const double* tril,
                            const double* cutoff, double skin, int max_neighbors,
                            const double* dt, const double* kT, const double* gamma,
                            const double* tau, const double* baro, double* tril_out,
                            double* pot_energy, double* kin_energy, double* virial) {
  KUPSMI_REQUIRE(n_atoms >= 0, "n_atoms < 0");
  KUPSMI_REQUIRE(n_steps == 0 || (dt != nullptr), "dt required");
  KUPSMI_REQUIRE(skin >= 0.0 && max_neighbors >= 0, "skin and max_neighbors must be >= 0");
  KUPSMI_REQUIRE(skin == 0.0 || (max_neighbors > 0 && max_neighbors <= 32767),
                 "a skin needs max_neighbors in [1, 32767]");
  KUPSMI_REQUIRE(integrator != 1 || n_steps == 0 || (kT && gamma), "kT and gamma required for BAOAB");
  KUPSMI_REQUIRE(integrator < 2 || n_steps == 0 || (kT && tau), "kT and tau required for CSVR");
  KUPSMI_REQUIRE(integrator != 3 || n_steps == 0 || (baro && tril_out),
                 "baro and tril_out required for cell rescaling");
  MdArgs a{};
  a.integrator = integrator; a.n_steps = n_steps; a.n_sys = n_sys; a.n_species = n_species;
  a.want_virial = virial != nullptr;
  a.step0 = step0; a.seed = seed;
  a.sys_off = sys_off; a.pos = pos; a.mom = mom; a.forces = forces; a.mass = mass;
  a.species = species; a.sigma = sigma; a.epsilon = epsilon; a.tril = tril; a.cutoff = cutoff;
  a.dt = dt; a.kT = kT; a.gamma = gamma; a.tau = tau; a.baro = baro; a.tril_out = tril_out;
  a.pot = pot_energy; a.kin = kin_energy; a.virial = virial;
  a.skin = (float)skin; a.nbmax = max_neighbors;
  return md_launch(ctx, (hipStream_t)stream, a);
}

extern "C" int kupsmi_lj_forces(kupsmi_ctx* ctx, void* stream, int n_sys,
                                const int32_t* sys_off, int n_atoms, const float* pos,
                                const int32_t* species, int n_species, const double* sigma,
                                const double* epsilon, const double* tril,
                                const double* cutoff, float* forces, double* energy,
                                double* virial) {
  KUPSMI_REQUIRE(n_atoms >= 0, "n_atoms < 0");
  // A force evaluation is the zero-step MD launch: positions are read but not
  // written back and the kinetic term is skipped (kin = NULL); momenta and
  // masses come from private zeros/ones arrays.
  void* buf = nullptr;
  int rc = scratch(ctx, 1, sizeof(float) * (size_t)(4 * (size_t)n_atoms + 4), &buf);
  if (rc) return rc;
  float* ones = static_cast<float*>(buf);
  float* zeros = ones + n_atoms + 1;
  hipStream_t s = (hipStream_t)stream;
  // fill kernel (no hipMemsetAsync: graph-capture safe)
  if (n_atoms > 0) {
    fill_ones_zeros<<<(3 * n_atoms + 255) / 256, 256, 0, s>>>(ones, zeros, n_atoms);
    KUPSMI_CHECK_LAUNCH();
  }
  MdArgs a{};
  a.integrator = 0; a.n_steps = 0; a.n_sys = n_sys; a.n_species = n_species;
  a.want_virial = virial != nullptr;
  a.sys_off = sys_off; a.pos = const_cast<float*>(pos); a.mom = zeros; a.forces = forces;
  a.mass = ones; a.species = species; a.sigma = sigma; a.epsilon = epsilon; a.tril = tril;
  a.cutoff = cutoff; a.pot = energy; a.kin = nullptr; a.virial = virial;
  return md_launch(ctx, s, a);
}

extern "C" int kupsmi_normals(void* stream, uint64_t seed, int64_t step, int n_atoms, float* eta) {
  KUPSMI_REQUIRE(n_atoms >= 0, "n_atoms < 0");
  if (n_atoms == 0) return KUPSMI_OK;
  normals_kernel<<<(n_atoms + 255) / 256, 256, 0, (hipStream_t)stream>>>(seed, step, n_atoms, eta);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_philox(void* stream, int n, const uint32_t* counters, const uint32_t* keys,
                             uint32_t* out) {
  KUPSMI_REQUIRE(n >= 0, "n < 0");
  if (n == 0) return KUPSMI_OK;
  KUPSMI_REQUIRE(counters && keys && out, "null buffer");
  philox_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(n, counters, keys, out);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}
