// Neighbor-list export: the directed pair set of every system, decided in
// FP64 with the reference's fractional-coordinate formula
// (core/neighborlist.py:542-564, 589, 595-624, 731-735).
//
// The export exists for parity and for callers that consume kUPS Edges; the
// MD and GCMC kernels do not materialise the list.  Three exports:
//   dense   every ordered pair of a system (DenseNearestNeighborList,
//           902-982), one thread per left atom, the system's atoms streamed
//           through L1/L2;
//   cells   a cell list (CellListNeighborList, 985-1077): atoms counting-
//           sorted into cells at least rc wide, each atom tests the atoms of
//           the distinct cells among its 27 neighbours (the reference's
//           jnp.unique over the stencil hashes, 295-364);
//   query   an rh subset against the full table (the rh / rh_index_remap
//           form of NearestNeighborList.__call__, 233-255, with the remap
//           rule 536-537 and both directions 680-685).
// Each runs in two passes: count per atom, exclusive scan, then write at the
// scanned offsets, so the output order is deterministic.
#include <hipcub/hipcub.hpp>
#include "pair_common.hpp"

namespace kupsmi {

constexpr int NL_THREADS = 256;

struct NlArgs {
  int n_sys;
  const int32_t* sys_off;
  const float* pos;
  const int32_t* excl;
  const double* tril;
  const double* cutoff;
  int64_t capacity;
  const int64_t* offsets;  // write pass: first edge of each atom
  int64_t* counts;         // count pass: edges per atom
  int32_t* ij;
  int32_t* shift;
};

// FP64 fractional coordinates of a stored FP32 position.
__device__ inline void frac64p(const double* li, const float* r, double& f0, double& f1,
                               double& f2) {
  frac64(li, r[0], r[1], r[2], f0, f1, f2);
}

#pragma clang fp contract(off)
// One workgroup per system (grid-stride over systems), one thread per left
// atom.  WRITE selects the pass.
template <bool WRITE>
__global__ __launch_bounds__(NL_THREADS) void nl_dense_kernel(NlArgs a) {
  for (int s = blockIdx.x; s < a.n_sys; s += gridDim.x) {
    const int i0 = a.sys_off[s], i1 = a.sys_off[s + 1];
    if (i1 <= i0) continue;
    SysGeom g;
    make_geom(a.tril + 6 * s, a.cutoff[s], 1, g);
    const int h0 = g.rep ? g.h[0] : 0, h1 = g.rep ? g.h[1] : 0, h2 = g.rep ? g.h[2] : 0;
    for (int i = i0 + threadIdx.x; i < i1; i += NL_THREADS) {
      double fi0, fi1, fi2;
      frac64p(g.Li, a.pos + 3 * i, fi0, fi1, fi2);
      const int ei = a.excl[i];
      int64_t cnt = 0;
      int64_t out = WRITE ? a.offsets[i] : 0;
      for (int j = i0; j < i1; ++j) {
        double fj0, fj1, fj2;
        frac64p(g.Li, a.pos + 3 * j, fj0, fj1, fj2);
        const double d0 = fi0 - fj0, d1 = fi1 - fj1, d2 = fi2 - fj2;
        const double m0 = rint(d0), m1 = rint(d1), m2 = rint(d2);
        const bool same = ei == a.excl[j];
        if (!g.rep) {
          const double r2 = dist2_64(g.L, d0 - m0, d1 - m1, d2 - m2);
          if (r2 < g.rc2 && !same) {
            if (WRITE && out < a.capacity) {
              a.ij[2 * out] = i; a.ij[2 * out + 1] = j;
              a.shift[3 * out] = (int)m0; a.shift[3 * out + 1] = (int)m1; a.shift[3 * out + 2] = (int)m2;
            }
            ++out; ++cnt;
          }
          continue;
        }
        for (int o0 = -h0; o0 <= h0; ++o0)
          for (int o1 = -h1; o1 <= h1; ++o1)
            for (int o2 = -h2; o2 <= h2; ++o2) {
              const double r2 = dist2_64(g.L, d0 - o0, d1 - o1, d2 - o2);
              const bool is_min = o0 == m0 && o1 == m1 && o2 == m2;
              if (r2 < g.rc2 && !(same && is_min)) {
                if (WRITE && out < a.capacity) {
                  a.ij[2 * out] = i; a.ij[2 * out + 1] = j;
                  a.shift[3 * out] = o0; a.shift[3 * out + 1] = o1; a.shift[3 * out + 2] = o2;
                }
                ++out; ++cnt;
              }
            }
      }
      if (!WRITE) a.counts[i] = cnt;
    }
  }
}
#pragma clang fp contract(on)

__global__ void nl_total_kernel(int n, const int64_t* offsets, const int64_t* counts,
                                int64_t* total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *total = n > 0 ? offsets[n - 1] + counts[n - 1] : 0;
}


// ---------------------------------------------------------------- cells ----
// Geometry of the cell export: cells at least rc wide per axis (the
// reference's _num_cells, 281-292, plus the face margin of make_geom), at
// most PC_MAX_NC_AXIS per axis; systems narrower than 2 rc use the image grid.
__global__ void nlc_geom_kernel(int n_sys, const double* tril, const double* cutoff,
                                SysGeom* geom) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n_sys) make_geom(tril + 6 * s, cutoff[s], PC_MAX_NC_AXIS, geom[s], 1);
}

__global__ __launch_bounds__(PC_BIN_THREADS) void nlc_bin_kernel(int n_sys, const int32_t* sys_off,
                                                                const float* pos,
                                                                const SysGeom* geom,
                                                                PcSorted base) {
  __shared__ PcBinLds L;
  __shared__ SysGeom g;
  for (int s = blockIdx.x; s < n_sys; s += gridDim.x) {
    const int i0 = sys_off[s], n = sys_off[s + 1] - i0;
    __syncthreads();
    if (threadIdx.x == 0) g = geom[s];
    __syncthreads();
    PcSorted o = base;
    o.cs = base.cs + (size_t)s * (PC_NCMAX + 1);
    pc_bin_system(g, i0, n, pos, nullptr, nullptr, nullptr, nullptr, o, L);
  }
}

// The distinct cell coordinates among c - 1, c, c + 1 (mod n) on one axis.
__device__ inline int stencil_axis(int c, int n, int (&out)[3]) {
  if (n >= 3) {
    out[0] = c == 0 ? n - 1 : c - 1; out[1] = c; out[2] = c == n - 1 ? 0 : c + 1;
    return 3;
  }
  for (int k = 0; k < n; ++k) out[k] = k;
  return n;
}

struct NlcArgs {
  int n_sys;
  const int32_t* sys_off;
  const float* pos;
  const int32_t* excl;
  const double* cutoff;
  const SysGeom* geom;
  const int* sidx;
  const int* scell;
  const int* cstart;
  int64_t capacity;
  const int64_t* offsets;
  int64_t* counts;
  int32_t* ij;
  int32_t* shift;
};

#pragma clang fp contract(off)
// One thread per sorted atom of a system (grid: tiles x systems): the left
// atom i = its caller index; candidates j from the distinct stencil cells
// with the minimum-image shift round(f_i - f_j), or the image grid of a
// narrow system (every atom of the system, one cell).
template <bool WRITE>
__global__ __launch_bounds__(NL_THREADS) void nlc_pair_kernel(NlcArgs a) {
  const int s = blockIdx.y;
  const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
  const int d = blockIdx.x * NL_THREADS + threadIdx.x;
  if (d >= n || n > PC_NMAX) return;
  const SysGeom& g = a.geom[s];
  const double rc2 = a.cutoff[s] * a.cutoff[s];
  const int* cs = a.cstart + (size_t)s * (PC_NCMAX + 1);
  const int i = i0 + a.sidx[i0 + d];
  double fi0, fi1, fi2;
  frac64p(g.Li, a.pos + 3 * i, fi0, fi1, fi2);
  const int ei = a.excl[i];
  int64_t out = WRITE ? a.offsets[i] : 0;
  int64_t cnt = 0;
  auto emit = [&](int j, int s0, int s1, int s2) {
    if (WRITE && out < a.capacity) {
      a.ij[2 * out] = i; a.ij[2 * out + 1] = j;
      a.shift[3 * out] = s0; a.shift[3 * out + 1] = s1; a.shift[3 * out + 2] = s2;
    }
    ++out; ++cnt;
  };
  if (g.rep) {
    const int h0 = g.h[0], h1 = g.h[1], h2 = g.h[2];
    for (int j = i0; j < i0 + n; ++j) {
      double fj0, fj1, fj2;
      frac64p(g.Li, a.pos + 3 * j, fj0, fj1, fj2);
      const double d0 = fi0 - fj0, d1 = fi1 - fj1, d2 = fi2 - fj2;
      const double m0 = rint(d0), m1 = rint(d1), m2 = rint(d2);
      const bool same = ei == a.excl[j];
      for (int o0 = -h0; o0 <= h0; ++o0)
        for (int o1 = -h1; o1 <= h1; ++o1)
          for (int o2 = -h2; o2 <= h2; ++o2) {
            const bool is_min = o0 == m0 && o1 == m1 && o2 == m2;
            if (dist2_64(g.L, d0 - o0, d1 - o1, d2 - o2) < rc2 && !(same && is_min))
              emit(j, o0, o1, o2);
          }
    }
  } else {
    const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];
    const int c = a.scell[i0 + d];
    int xs[3], ys[3], zs[3];
    const int nx = stencil_axis(c % ncx, ncx, xs);
    const int ny = stencil_axis((c / ncx) % ncy, ncy, ys);
    const int nz = stencil_axis(c / (ncx * ncy), ncz, zs);
    for (int iz = 0; iz < nz; ++iz)
      for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
          const int cell = (zs[iz] * ncy + ys[iy]) * ncx + xs[ix];
          for (int t = cs[cell]; t < cs[cell + 1]; ++t) {
            const int j = i0 + a.sidx[i0 + t];
            double fj0, fj1, fj2;
            frac64p(g.Li, a.pos + 3 * j, fj0, fj1, fj2);
            const double d0 = fi0 - fj0, d1 = fi1 - fj1, d2 = fi2 - fj2;
            const double m0 = rint(d0), m1 = rint(d1), m2 = rint(d2);
            if (dist2_64(g.L, d0 - m0, d1 - m1, d2 - m2) < rc2 && ei != a.excl[j])
              emit(j, (int)m0, (int)m1, (int)m2);
          }
        }
  }
  if (!WRITE) a.counts[i] = cnt;
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------- query ----
// Edges between the lh table (every particle, systems by sys_off) and k rh
// particles (position, system, exclusion id each): for rh particle r in
// system s and every lh particle i of s, the shift is round(f_i - f_r) (or
// each offset of the image grid of a narrow system), the pair is kept when
// |(f_i - f_r - s) @ L|^2 < rc^2 and not (equal exclusion ids on the
// minimum image) (core/neighborlist.py:505-564, 692-740).  With a remap
// (rh particle r is lh particle remap[r]): an lh particle that some remap
// entry names pairs only with entries whose remap is <= its own index
// (536-537), and every edge is written in both directions, all edges
// (i, remap[r], s) first, then the reversed (remap[r], i, -s) in the same
// order (680-685).  Without one, edges are (i, r, s) with r indexing rh.
// A remap entry outside [0, n_atoms) stands for a particle outside the lh
// table (the reference's out-of-bounds index n_atoms).
struct NqArgs {
  int n_sys, n_atoms, k;
  const int32_t* sys_off;
  const float* pos;
  const int32_t* excl;
  const double* tril;
  const double* cutoff;
  const float* rh_pos;
  const int32_t* rh_sys;
  const int32_t* rh_excl;
  const int32_t* remap;     // NULL: rh disjoint from lh
  const int32_t* in_remap;  // per lh particle: named by some remap entry
  int64_t capacity;
  const int64_t* offsets;
  int64_t* counts;
  int32_t* ij;
  int32_t* shift;
};

__global__ void nq_flag_kernel(int n_atoms, int32_t* in_remap) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_atoms) in_remap[t] = 0;
}

__global__ void nq_flag_set_kernel(int n_atoms, int k, const int32_t* remap, int32_t* in_remap) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < k && remap[r] >= 0 && remap[r] < n_atoms) in_remap[remap[r]] = 1;
}

#pragma clang fp contract(off)
template <bool WRITE>
__global__ __launch_bounds__(NL_THREADS) void nq_kernel(NqArgs a) {
  const int r = blockIdx.x * NL_THREADS + threadIdx.x;
  if (r >= a.k) return;
  int64_t cnt = 0;
  const int s = a.rh_sys[r];
  if (s >= 0 && s < a.n_sys) {
    const int i0 = a.sys_off[s], i1 = a.sys_off[s + 1];
    SysGeom g;
    make_geom(a.tril + 6 * s, a.cutoff[s], 1, g);
    const int h0 = g.rep ? g.h[0] : 0, h1 = g.rep ? g.h[1] : 0, h2 = g.rep ? g.h[2] : 0;
    double fr0, fr1, fr2;
    frac64p(g.Li, a.rh_pos + 3 * r, fr0, fr1, fr2);
    const int er = a.rh_excl[r];
    int out_r = r;
    if (a.remap) {
      out_r = a.remap[r];
      if (out_r < 0 || out_r >= a.n_atoms) out_r = a.n_atoms;
    }
    int64_t out = WRITE ? a.offsets[r] : 0;
    const int64_t total = WRITE ? a.offsets[a.k - 1] + a.counts[a.k - 1] : 0;
    auto emit = [&](int i, int s0, int s1, int s2) {
      if (WRITE) {
        if (out < a.capacity) {
          a.ij[2 * out] = i; a.ij[2 * out + 1] = out_r;
          a.shift[3 * out] = s0; a.shift[3 * out + 1] = s1; a.shift[3 * out + 2] = s2;
        }
        const int64_t rev = total + out;
        if (a.remap && rev < a.capacity) {
          a.ij[2 * rev] = out_r; a.ij[2 * rev + 1] = i;
          a.shift[3 * rev] = -s0; a.shift[3 * rev + 1] = -s1; a.shift[3 * rev + 2] = -s2;
        }
      }
      ++out; ++cnt;
    };
    for (int i = i0; i < i1; ++i) {
      if (a.remap && a.in_remap[i] && !(i >= out_r)) continue;
      double fi0, fi1, fi2;
      frac64p(g.Li, a.pos + 3 * i, fi0, fi1, fi2);
      const double d0 = fi0 - fr0, d1 = fi1 - fr1, d2 = fi2 - fr2;
      const double m0 = rint(d0), m1 = rint(d1), m2 = rint(d2);
      const bool same = a.excl[i] == er;
      if (!g.rep) {
        if (dist2_64(g.L, d0 - m0, d1 - m1, d2 - m2) < g.rc2 && !same)
          emit(i, (int)m0, (int)m1, (int)m2);
        continue;
      }
      for (int o0 = -h0; o0 <= h0; ++o0)
        for (int o1 = -h1; o1 <= h1; ++o1)
          for (int o2 = -h2; o2 <= h2; ++o2) {
            const bool is_min = o0 == m0 && o1 == m1 && o2 == m2;
            if (dist2_64(g.L, d0 - o0, d1 - o1, d2 - o2) < g.rc2 && !(same && is_min))
              emit(i, o0, o1, o2);
          }
    }
  }
  if (!WRITE) a.counts[r] = cnt;
}
#pragma clang fp contract(on)

__global__ void nq_total_kernel(int k, int both, const int64_t* offsets, const int64_t* counts,
                                int64_t* total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    const int64_t f = k > 0 ? offsets[k - 1] + counts[k - 1] : 0;
    *total = both ? 2 * f : f;
  }
}

// ------------------------------------------------------------- changes ----
// Edges that moving k particles removes and adds (neighborlist_changes,
// core/neighborlist.py:1424-1515).  removed = the directed edges of the
// current configuration with an end among the moved particles; added = the
// same for the configuration with the moved particles at their new positions
// (indices are the particles' own, as the reference remaps them).  Both use
// the full list's pair decision above (FP64 fractional formula, image grid,
// minimum-image exclusion).  One thread per (moved particle, configuration)
// scans its system: an edge to an unmoved partner is written in both
// directions, an edge between two moved particles (or a moved particle and
// its own image) only from its first end, so every directed edge appears
// once.  Order: moved particle, partner, image, then the edge before its
// reverse.
struct NcArgs {
  int n_sys, k;
  const int32_t* sys_off;
  const float* pos;
  const int32_t* excl;
  const double* tril;
  const double* cutoff;
  const int32_t* moved;     // k particle indices
  const float* moved_pos;   // k x 3 new positions
  const int32_t* slot;      // per particle: its index in moved, or -1
  int64_t cap[2];           // removed, added
  int64_t* counts;          // 2 x k
  const int64_t* offsets;   // 2 x k
  int32_t* ij[2];
  int32_t* shift[2];
};

__global__ void nc_slot_clear_kernel(int n, int32_t* slot, int32_t* bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) slot[i] = -1;
  if (i == 0) *bad = 0;
}

// slot[moved[m]] = m; an index outside [0, n) or named twice sets *bad
__global__ void nc_slot_set_kernel(int k, int n, const int32_t* moved, int32_t* slot,
                                   int32_t* bad) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= k) return;
  const int c = moved[m];
  if (c < 0 || c >= n) { *bad = 1; return; }
  if (atomicExch(slot + c, m) != -1) *bad = 1;
}

#pragma clang fp contract(off)
template <bool WRITE>
__global__ __launch_bounds__(NL_THREADS) void nc_kernel(NcArgs a, int n_atoms, const int32_t* bad) {
  const int m = blockIdx.x * NL_THREADS + threadIdx.x;
  const int cfg = blockIdx.y;  // 0: current positions (removed), 1: moved (added)
  if (m >= a.k) return;
  const int c = a.moved[m];
  int64_t cnt = 0;
  if (*bad == 0 && c >= 0 && c < n_atoms) {
    int lo = 0, hi = a.n_sys - 1;  // the system of c: largest s with sys_off[s] <= c
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (a.sys_off[mid] <= c) lo = mid; else hi = mid - 1;
    }
    const int i0 = a.sys_off[lo], i1 = a.sys_off[lo + 1];
    SysGeom g;
    make_geom(a.tril + 6 * lo, a.cutoff[lo], 1, g);
    const int h0 = g.rep ? g.h[0] : 0, h1 = g.rep ? g.h[1] : 0, h2 = g.rep ? g.h[2] : 0;
    double fc0, fc1, fc2;
    frac64p(g.Li, cfg ? a.moved_pos + 3 * m : a.pos + 3 * c, fc0, fc1, fc2);
    const int ec = a.excl[c];
    const int64_t cap = a.cap[cfg];
    int32_t* ij = a.ij[cfg];
    int32_t* sh = a.shift[cfg];
    int64_t out = WRITE ? a.offsets[cfg * a.k + m] : 0;
    auto emit = [&](int p, int q, int s0, int s1, int s2) {
      if (WRITE && out < cap) {
        ij[2 * out] = p; ij[2 * out + 1] = q;
        sh[3 * out] = s0; sh[3 * out + 1] = s1; sh[3 * out + 2] = s2;
      }
      ++out; ++cnt;
    };
    for (int j = i0; j < i1; ++j) {
      const int sj = a.slot[j];
      const bool moved_j = sj >= 0;
      double fj0, fj1, fj2;
      frac64p(g.Li, cfg && moved_j ? a.moved_pos + 3 * sj : a.pos + 3 * j, fj0, fj1, fj2);
      const double d0 = fc0 - fj0, d1 = fc1 - fj1, d2 = fc2 - fj2;
      const double m0 = rint(d0), m1 = rint(d1), m2 = rint(d2);
      const bool same = ec == a.excl[j];
      if (!g.rep) {
        if (dist2_64(g.L, d0 - m0, d1 - m1, d2 - m2) < g.rc2 && !same) {
          emit(c, j, (int)m0, (int)m1, (int)m2);
          if (!moved_j) emit(j, c, -(int)m0, -(int)m1, -(int)m2);
        }
        continue;
      }
      for (int o0 = -h0; o0 <= h0; ++o0)
        for (int o1 = -h1; o1 <= h1; ++o1)
          for (int o2 = -h2; o2 <= h2; ++o2) {
            const bool is_min = o0 == m0 && o1 == m1 && o2 == m2;
            if (dist2_64(g.L, d0 - o0, d1 - o1, d2 - o2) < g.rc2 && !(same && is_min)) {
              emit(c, j, o0, o1, o2);
              if (!moved_j) emit(j, c, -o0, -o1, -o2);
            }
          }
    }
  }
  if (!WRITE) a.counts[cfg * a.k + m] = cnt;
}
#pragma clang fp contract(on)

// totals; -1 in both when an index was out of range or repeated
__global__ void nc_total_kernel(int k, const int64_t* offsets, const int64_t* counts,
                                const int32_t* bad, int64_t* removed, int64_t* added) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const bool b = *bad != 0;
  *removed = b ? -1 : (k > 0 ? offsets[k - 1] + counts[k - 1] : 0);
  *added = b ? -1 : (k > 0 ? offsets[2 * k - 1] + counts[2 * k - 1] : 0);
}

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_neighborlist(kupsmi_ctx* ctx, void* stream, int n_sys,
                                   const int32_t* sys_off, int n_atoms, const float* pos,
                                   const int32_t* exclusion, const double* tril,
                                   const double* cutoff, int64_t capacity,
                                   int32_t* edges_ij, int32_t* edges_shift, int64_t* count) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  KUPSMI_REQUIRE(n_sys >= 0 && n_atoms >= 0 && capacity >= 0, "negative size");
  KUPSMI_REQUIRE(count != nullptr, "count pointer required");
  hipStream_t st = (hipStream_t)stream;
  void* buf = nullptr;
  const size_t nn = (size_t)(n_atoms > 0 ? n_atoms : 1);
  size_t temp_bytes = 0;
  KUPSMI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, (int64_t*)nullptr,
                                              (int64_t*)nullptr, (int)nn, st));
  const size_t need = 2 * nn * sizeof(int64_t) + temp_bytes + 256;
  int rc = scratch(ctx, 2, need, &buf);
  if (rc) return rc;
  int64_t* counts = static_cast<int64_t*>(buf);
  int64_t* offsets = counts + nn;
  void* temp = reinterpret_cast<char*>(offsets + nn);
  NlArgs a{};
  a.n_sys = n_sys; a.sys_off = sys_off; a.pos = pos; a.excl = exclusion; a.tril = tril;
  a.cutoff = cutoff; a.capacity = capacity; a.counts = counts; a.offsets = offsets;
  a.ij = edges_ij; a.shift = edges_shift;
  if (n_atoms == 0 || n_sys == 0) {
    nl_total_kernel<<<1, 1, 0, st>>>(0, offsets, counts, count);
    KUPSMI_CHECK_LAUNCH();
    return KUPSMI_OK;
  }
  const int grid = n_sys < 4096 ? n_sys : 4096;
  nl_dense_kernel<false><<<grid, NL_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  KUPSMI_HIP(hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, offsets, n_atoms, st));
  nl_dense_kernel<true><<<grid, NL_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  nl_total_kernel<<<1, 1, 0, st>>>(n_atoms, offsets, counts, count);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_neighborlist_cells(kupsmi_ctx* ctx, void* stream, int n_sys,
                                         const int32_t* sys_off, int n_atoms, int max_atoms_per_sys,
                                         const float* pos, const int32_t* exclusion,
                                         const double* tril, const double* cutoff,
                                         int64_t capacity, int32_t* edges_ij,
                                         int32_t* edges_shift, int64_t* count) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  KUPSMI_REQUIRE(n_sys >= 0 && n_atoms >= 0 && capacity >= 0, "negative size");
  KUPSMI_REQUIRE(count != nullptr, "count pointer required");
  KUPSMI_REQUIRE(max_atoms_per_sys <= PC_NMAX, "systems of the cell export hold at most 16384 atoms");
  hipStream_t st = (hipStream_t)stream;
  const size_t nn = (size_t)(n_atoms > 0 ? n_atoms : 1), ns = (size_t)(n_sys > 0 ? n_sys : 1);
  size_t temp_bytes = 0;
  KUPSMI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, (int64_t*)nullptr,
                                              (int64_t*)nullptr, (int)nn, st));
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
  const size_t o_cnt = take(nn * sizeof(int64_t)), o_off = take(nn * sizeof(int64_t));
  const size_t o_geom = take(ns * sizeof(SysGeom)), o_spos = take(nn * sizeof(float4));
  const size_t o_sidx = take(nn * sizeof(int)), o_scell = take(nn * sizeof(int));
  const size_t o_sspec = take(nn * sizeof(int)), o_cs = take(ns * (PC_NCMAX + 1) * sizeof(int));
  const size_t o_temp = take(temp_bytes);
  void* buf = nullptr;
  int rc = scratch(ctx, 2, off, &buf);
  if (rc) return rc;
  char* b = static_cast<char*>(buf);
  int64_t* counts = reinterpret_cast<int64_t*>(b + o_cnt);
  int64_t* offsets = reinterpret_cast<int64_t*>(b + o_off);
  if (n_atoms == 0 || n_sys == 0) {
    nl_total_kernel<<<1, 1, 0, st>>>(0, offsets, counts, count);
    KUPSMI_CHECK_LAUNCH();
    return KUPSMI_OK;
  }
  SysGeom* geom = reinterpret_cast<SysGeom*>(b + o_geom);
  PcSorted srt{reinterpret_cast<float4*>(b + o_spos), reinterpret_cast<int*>(b + o_sidx),
               reinterpret_cast<int*>(b + o_scell), reinterpret_cast<int*>(b + o_sspec),
               nullptr, nullptr, reinterpret_cast<int*>(b + o_cs)};
  nlc_geom_kernel<<<(n_sys + 255) / 256, 256, 0, st>>>(n_sys, tril, cutoff, geom);
  KUPSMI_CHECK_LAUNCH();
  int dev = 0, ncu = 256;
  KUPSMI_HIP(hipGetDevice(&dev));
  KUPSMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  nlc_bin_kernel<<<n_sys < ncu ? n_sys : ncu, PC_BIN_THREADS, 0, st>>>(n_sys, sys_off, pos, geom,
                                                                       srt);
  KUPSMI_CHECK_LAUNCH();
  NlcArgs a{};
  a.n_sys = n_sys; a.sys_off = sys_off; a.pos = pos; a.excl = exclusion; a.cutoff = cutoff;
  a.geom = geom; a.sidx = srt.sidx; a.scell = srt.scell; a.cstart = srt.cs;
  a.capacity = capacity; a.offsets = offsets; a.counts = counts; a.ij = edges_ij;
  a.shift = edges_shift;
  const int tiles = (max_atoms_per_sys + NL_THREADS - 1) / NL_THREADS;
  const dim3 grid(tiles > 0 ? tiles : 1, n_sys);
  nlc_pair_kernel<false><<<grid, NL_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  KUPSMI_HIP(hipcub::DeviceScan::ExclusiveSum(b + o_temp, temp_bytes, counts, offsets, n_atoms, st));
  nlc_pair_kernel<true><<<grid, NL_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  nl_total_kernel<<<1, 1, 0, st>>>(n_atoms, offsets, counts, count);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_neighborlist_query(
    kupsmi_ctx* ctx, void* stream, int n_sys, const int32_t* sys_off, int n_atoms,
    const float* pos, const int32_t* exclusion, int k, const float* rh_pos,
    const int32_t* rh_system, const int32_t* rh_exclusion, const int32_t* rh_index_remap,
    const double* tril, const double* cutoff, int64_t capacity, int32_t* edges_ij,
    int32_t* edges_shift, int64_t* count) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  KUPSMI_REQUIRE(n_sys >= 0 && n_atoms >= 0 && k >= 0 && capacity >= 0, "negative size");
  KUPSMI_REQUIRE(count != nullptr, "count pointer required");
  KUPSMI_REQUIRE(k == 0 || (rh_pos && rh_system && rh_exclusion), "rh particles required");
  hipStream_t st = (hipStream_t)stream;
  const size_t kk = (size_t)(k > 0 ? k : 1), nn = (size_t)(n_atoms > 0 ? n_atoms : 1);
  size_t temp_bytes = 0;
  KUPSMI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, (int64_t*)nullptr,
                                              (int64_t*)nullptr, (int)kk, st));
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
  const size_t o_cnt = take(kk * sizeof(int64_t)), o_off = take(kk * sizeof(int64_t));
  const size_t o_flag = take(nn * sizeof(int32_t)), o_temp = take(temp_bytes);
  void* buf = nullptr;
  int rc = scratch(ctx, 8, off, &buf);
  if (rc) return rc;
  char* b = static_cast<char*>(buf);
  int64_t* counts = reinterpret_cast<int64_t*>(b + o_cnt);
  int64_t* offsets = reinterpret_cast<int64_t*>(b + o_off);
  int32_t* flag = reinterpret_cast<int32_t*>(b + o_flag);
  if (k == 0 || n_sys == 0) {
    nq_total_kernel<<<1, 1, 0, st>>>(0, 0, offsets, counts, count);
    KUPSMI_CHECK_LAUNCH();
    return KUPSMI_OK;
  }
  if (rh_index_remap) {
    nq_flag_kernel<<<(int)((nn + 255) / 256), 256, 0, st>>>(n_atoms, flag);
    KUPSMI_CHECK_LAUNCH();
    nq_flag_set_kernel<<<(k + 255) / 256, 256, 0, st>>>(n_atoms, k, rh_index_remap, flag);
    KUPSMI_CHECK_LAUNCH();
  }
  NqArgs a{};
  a.n_sys = n_sys; a.n_atoms = n_atoms; a.k = k; a.sys_off = sys_off; a.pos = pos;
  a.excl = exclusion; a.tril = tril; a.cutoff = cutoff; a.rh_pos = rh_pos; a.rh_sys = rh_system;
  a.rh_excl = rh_exclusion; a.remap = rh_index_remap; a.in_remap = flag; a.capacity = capacity;
  a.offsets = offsets; a.counts = counts; a.ij = edges_ij; a.shift = edges_shift;
  const int grid = (k + NL_THREADS - 1) / NL_THREADS;
  nq_kernel<false><<<grid, NL_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  KUPSMI_HIP(hipcub::DeviceScan::ExclusiveSum(b + o_temp, temp_bytes, counts, offsets, k, st));
  nq_kernel<true><<<grid, NL_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  nq_total_kernel<<<1, 1, 0, st>>>(k, rh_index_remap != nullptr, offsets, counts, count);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_neighborlist_changes(
    kupsmi_ctx* ctx, void* stream, int n_sys, const int32_t* sys_off, int n_atoms,
    const float* pos, const int32_t* exclusion, const double* tril, const double* cutoff,
    int k, const int32_t* moved, const float* moved_pos, int64_t capacity_added,
    int32_t* added_ij, int32_t* added_shift, int64_t* count_added, int64_t capacity_removed,
    int32_t* removed_ij, int32_t* removed_shift, int64_t* count_removed) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  KUPSMI_REQUIRE(n_sys >= 1 && n_atoms >= 1 && k >= 0, "bad sizes");
  KUPSMI_REQUIRE(capacity_added >= 0 && capacity_removed >= 0, "negative capacity");
  KUPSMI_REQUIRE(count_added != nullptr && count_removed != nullptr, "count pointers required");
  KUPSMI_REQUIRE(k == 0 || (moved != nullptr && moved_pos != nullptr), "moved particles required");
  hipStream_t st = (hipStream_t)stream;
  const size_t kk = (size_t)(k > 0 ? k : 1);
  size_t temp_bytes = 0;
  KUPSMI_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, (int64_t*)nullptr,
                                              (int64_t*)nullptr, (int)kk, st));
  // scratch: counts | offsets | slot | bad | scan temp, each 256-byte aligned
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_off = al(2 * kk * sizeof(int64_t));
  const size_t o_slot = o_off + o_off;
  const size_t o_bad = o_slot + al((size_t)n_atoms * sizeof(int32_t));
  const size_t o_temp = o_bad + 256;
  void* buf = nullptr;
  int rc = scratch(ctx, 7, o_temp + temp_bytes, &buf);
  if (rc) return rc;
  char* base = static_cast<char*>(buf);
  int64_t* counts = reinterpret_cast<int64_t*>(base);
  int64_t* offsets = reinterpret_cast<int64_t*>(base + o_off);
  int32_t* slot = reinterpret_cast<int32_t*>(base + o_slot);
  int32_t* bad = reinterpret_cast<int32_t*>(base + o_bad);
  void* temp = base + o_temp;
  nc_slot_clear_kernel<<<(n_atoms + 255) / 256, 256, 0, st>>>(n_atoms, slot, bad);
  KUPSMI_CHECK_LAUNCH();
  if (k == 0) {
    nc_total_kernel<<<1, 1, 0, st>>>(0, offsets, counts, bad, count_removed, count_added);
    KUPSMI_CHECK_LAUNCH();
    return KUPSMI_OK;
  }
  nc_slot_set_kernel<<<(k + 255) / 256, 256, 0, st>>>(k, n_atoms, moved, slot, bad);
  KUPSMI_CHECK_LAUNCH();
  NcArgs a{};
  a.n_sys = n_sys; a.k = k; a.sys_off = sys_off; a.pos = pos; a.excl = exclusion;
  a.tril = tril; a.cutoff = cutoff; a.moved = moved; a.moved_pos = moved_pos; a.slot = slot;
  a.cap[0] = capacity_removed; a.cap[1] = capacity_added;
  a.counts = counts; a.offsets = offsets;
  a.ij[0] = removed_ij; a.ij[1] = added_ij;
  a.shift[0] = removed_shift; a.shift[1] = added_shift;
  const dim3 grid((k + NL_THREADS - 1) / NL_THREADS, 2);
  nc_kernel<false><<<grid, NL_THREADS, 0, st>>>(a, n_atoms, bad);
  KUPSMI_CHECK_LAUNCH();
  KUPSMI_HIP(hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts, offsets, k, st));
  KUPSMI_HIP(hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, counts + k, offsets + k, k, st));
  nc_kernel<true><<<grid, NL_THREADS, 0, st>>>(a, n_atoms, bad);
  KUPSMI_CHECK_LAUNCH();
  nc_total_kernel<<<1, 1, 0, st>>>(k, offsets, counts, bad, count_removed, count_added);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}
