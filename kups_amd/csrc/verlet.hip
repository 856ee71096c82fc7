// LJ + real-space Ewald over a Verlet list with a skin, rebuilt per system on
// the device when an atom has moved more than skin/2 since the last build
// (the neighbor-list reuse the reference's per-step rebuild lacks,
// core/neighborlist.py:985-1077 rebuilds every call).  Same energies, forces,
// virials and pair semantics as kupsmi_pair_forces; the state lives in a
// kupsmi_verlet handle.
//
// Per call, on the caller's stream:
//   clear   max displacement per system := 0
//   gather  per tile: sorted slot d <- the caller's atom sidx[d], moved by
//           the whole lattice vectors it was wrapped by since the build, so
//           the image shifts stored in the list stay valid; the squared
//           displacement since the build feeds the system's maximum;
//   bin     one workgroup per system that needs a rebuild (first call, or
//           maximum displacement > skin/2): geometry for rc + skin and the
//           counting sort of the cell-list path (cells >= (rc + skin)/2);
//   scan    the cells of those systems as one compact work list;
//   list    one wave per listed cell appends (j, shift, species) of each
//           stencil candidate within rc + skin (FP32 test with a margin) to
//           the lists of the cell's atoms, stored in groups of four entries,
//           entry-major (group k of atom d at k * n_atoms + d) so a wave's
//           loads of group k are contiguous;
//   pair    one lane per sorted atom walks its list with the next entry and
//           position in flight: FP64 pair vector, the pair terms of the
//           cell-list path; an atom whose list overflowed the capacity walks
//           the 125-cell stencil of the build instead;
//   reduce  fixed-order FP64 sums per system.
// Positions, cells (tril), charges, species, exclusion ids and sizes must stay
// fixed between rebuilds except for the positions; kupsmi_verlet_invalidate
// forces a rebuild of every system.
#include <utility>
#include <vector>
#include <type_traits>
#include "pair_common.hpp"

struct kupsmi_verlet {
  int device = -1;
  int n_sys = 0, n_atoms = 0, cap = 0, tiles = 0;
  void* mem = nullptr;
  size_t bytes = 0;
  // armed by kupsmi_verlet_record_pairs, taken by the next evaluation
  kupsmi::PairRecord rec;
  int rec_armed = 0;
  // set by kupsmi_verlet_freeze: reallocating the buffers is an error
  int frozen = 0;
  // kupsmi_verlet_profile: HIP events around the main pair pass of each call
  int profile = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  // kupsmi_verlet_gate: an event the next evaluation waits for between its
  // list upkeep and its pair passes (taken by that evaluation)
  hipEvent_t gate = nullptr;
};

namespace kupsmi {

constexpr int VL_TILE = 256;
static_assert(VL_TILE == PC_TILE, "band_ctx locates the left atom by the pair tile");
constexpr int VL_SHIFT_MAX = 7;  // |image shift| per axis in a packed entry
constexpr int VL_BIN_THREADS = 256;

struct VlArgs {
  int n_sys, n_atoms, n_species, has_lj, has_coul, has_lj_excl, has_coul_excl, accumulate;
  int tiles, cap;
  double skin;
  const int32_t* sys_off;
  const float* pos;
  const float* q;
  const int32_t* species;
  const int32_t* lj_excl;
  const int32_t* cl_excl;
  const double* sigma;
  const double* epsilon;
  const double* lj_rc;
  const double* alpha;
  const double* cl_rc;
  const double* tril;
  // persistent state
  SysGeom* geom;        // geometry of the last build (rc + skin)
  float4* spos;         // current positions in the build frame, sorted; q in w
  float4* sref;         // positions at the last build, sorted
  int* sidx;
  int* scell;
  int* sspec;
  int* slj;
  int* scl;
  int* cstart;          // n_sys x (PC_NCMAX + 1)
  int* nlist;           // entries per sorted atom (> cap: overflow)
  uint32_t* list;       // cap x n_atoms
  int* valid;           // n_sys
  int* need;            // n_sys: rebuild in this call
  int* reprc;           // n_sys: the cutoff itself needs the image grid
  unsigned* maxd2;      // n_sys: float bits of the max squared displacement
  unsigned long long* counters;  // [0] system rebuilds, [1] overflowed atom lists
  double* partial;      // main pass, then the band pass (n_sys x tiles x PC_NPART each)
  int* tile_off;        // n_sys + 1: prefix of the systems' tile counts
  int* cell_off;        // n_sys + 1: prefix of the cell counts of the systems rebuilt in a call
  // per sorted atom: the number of list entries inside a cutoff band met by
  // the main pass (VL_NBAND_FULL: more than VL_NBAND, or met on the stencil
  // path), and up to VL_NBAND of those entries
  uint8_t* bandflag;
  uint32_t* bandbuf;
  // tiles with lanes the band pass must walk in full (compact list, its
  // length, and a flag per (system, tile) for the reduction)
  int* band_tiles;
  int* band_count;
  uint8_t* tile_full;
  float* forces;
  double* energy;
  double* virial;
  PairRecord rec;  // parity output (REC instantiation only)
};

// List entry: partner slot (15 bits), image shift per axis (4-bit two's
// complement each, bits 15, 19, 23), partner species (5 bits).
__device__ inline uint32_t vl_pack(int t, int sx, int sy, int sz, int sp) {
  return (uint32_t)t | (((uint32_t)sx & 15u) << 15) | (((uint32_t)sy & 15u) << 19) |
         (((uint32_t)sz & 15u) << 23) | ((uint32_t)sp << 27);
}

// Signed shift fields of an entry (one signed bit-field extract each).
__device__ inline void vl_shifts(uint32_t e, int& sx, int& sy, int& sz) {
  sx = (int)(e << 13) >> 28;
  sy = (int)(e << 9) >> 28;
  sz = (int)(e << 5) >> 28;
}

__global__ void vl_invalidate_kernel(int n_sys, int* valid) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n_sys) valid[s] = 0;
}

// Workgroups map to (system, tile) through the tile prefix, XCD-aware: the
// hardware deals consecutive workgroups round-robin to the VL_XCDS dies, so
// workgroup w takes tile (w % VL_XCDS) * per + w / VL_XCDS of the compact
// tile order, and each die walks one contiguous stretch of systems (their
// positions stay in that die's L2 while its CUs work on them).
constexpr int VL_XCDS = 8;
constexpr int VL_PREP_THREADS = 1024;

__device__ inline bool vl_unit_of(const int* off, int n_sys, int& s, int& tile) {
  const int T = off[n_sys];
  const int per = (T + VL_XCDS - 1) / VL_XCDS;
  const int w = blockIdx.x, k = w / VL_XCDS;
  const int L = (w % VL_XCDS) * per + k;
  if (k >= per || L >= T) return false;
  int lo = 0, hi = n_sys;  // the last system whose first tile is <= L
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= L) lo = mid; else hi = mid;
  }
  s = lo;
  tile = L - off[lo];
  return true;
}

__device__ inline bool vl_tile_of(const VlArgs& a, int& s, int& tile) {
  return vl_unit_of(a.tile_off, a.n_sys, s, tile);
}

// Grid of a launch mapped by vl_unit_of: the units number at most
// T = ceil(n_atoms / unit) + n_sys, and workgroup w = VL_XCDS k + x serves
// unit x per + k (k < per = ceil(T / VL_XCDS)), so the grid is VL_XCDS per.
inline unsigned vl_grid(int n_atoms, int n_sys, int unit) {
  const long long t = ((long long)n_atoms + unit - 1) / unit + n_sys;
  return (unsigned)(VL_XCDS * ((t + VL_XCDS - 1) / VL_XCDS));
}

// Exclusive prefix of width(s) over the n_sys systems into out[0..n_sys], by
// one workgroup of VL_PREP_THREADS lanes (each lane sums a chunk of systems,
// then a block scan of the chunk sums).
template <class Width>
__device__ void vl_block_prefix(int ns, int* out, Width width, int* part) {
  const int tid = threadIdx.x;
  const int chunk = (ns + VL_PREP_THREADS - 1) / VL_PREP_THREADS;
  const int b = tid * chunk, e = min(b + chunk, ns);
  int sum = 0;
  for (int k = b; k < e; ++k) sum += width(k);
  __syncthreads();
  part[tid] = sum;
  __syncthreads();
  for (int o = 1; o < VL_PREP_THREADS; o <<= 1) {
    const int v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  int run = part[tid] - sum;  // exclusive prefix of this chunk
  for (int k = b; k < e; ++k) {
    out[k] = run;
    run += width(k);
  }
  if (tid == VL_PREP_THREADS - 1) out[ns] = part[tid];
}

// One workgroup: clears the per-system displacement maxima and the band
// pass's tile count and writes the tile prefix (ceil(n_s / VL_TILE) per
// system).
__global__ __launch_bounds__(VL_PREP_THREADS) void vl_prep_kernel(VlArgs a) {
  __shared__ int part[VL_PREP_THREADS];
  for (int k = threadIdx.x; k < a.n_sys; k += VL_PREP_THREADS) a.maxd2[k] = 0u;
  if (threadIdx.x == 0) *a.band_count = 0;
  vl_block_prefix(a.n_sys, a.tile_off, [&](int k) {
    return (a.sys_off[k + 1] - a.sys_off[k] + VL_TILE - 1) / VL_TILE;
  }, part);
}

// One workgroup, after the binning: the work list of the list build, the
// prefix of the cell counts of the systems rebuilt in this call (0 for the
// others), so its workers walk only those cells.
__global__ __launch_bounds__(VL_PREP_THREADS) void vl_cell_scan_kernel(VlArgs a) {
  __shared__ int part[VL_PREP_THREADS];
  vl_block_prefix(a.n_sys, a.cell_off, [&](int k) {
    const int n = a.sys_off[k + 1] - a.sys_off[k];
    return a.need[k] && n > 0 && n <= PC_NMAX ? a.geom[k].ncell : 0;
  }, part);
}

__global__ __launch_bounds__(VL_TILE) void vl_gather_kernel(VlArgs a) {
  __shared__ float red[VL_TILE / 64];
  int s, tile;
  if (!vl_tile_of(a, s, tile) || !a.valid[s]) return;
  const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
  const int slot = tile * VL_TILE + threadIdx.x;
  const SysGeom& g = a.geom[s];
  float d2 = 0.f;
  if (slot < n) {
    const int d = i0 + slot;
    const float* r = a.pos + 3 * (size_t)(i0 + a.sidx[d]);
    const float4 rb = a.sref[d];
    // whole lattice vectors the atom was wrapped by since the build (its true
    // displacement is below skin/2, far below half a cell)
    double f0, f1, f2;
    tmul(g.Li, (double)r[0] - rb.x, (double)r[1] - rb.y, (double)r[2] - rb.z, f0, f1, f2);
    double s0, s1, s2;
    tmul(g.L, rint(f0), rint(f1), rint(f2), s0, s1, s2);
    const double ux = (double)r[0] - s0, uy = (double)r[1] - s1, uz = (double)r[2] - s2;
    const double dx = ux - rb.x, dy = uy - rb.y, dz = uz - rb.z;
    d2 = (float)(dx * dx + dy * dy + dz * dz);
    // a system whose cutoff needs the image grid takes the reference's
    // centred grid of shifts in the caller's frame (core/neighborlist.py:
    // 595-624), which moves with every wrap: rebuild once an atom was wrapped
    if (a.reprc[s] && (rint(f0) != 0.0 || rint(f1) != 0.0 || rint(f2) != 0.0)) d2 = INFINITY;
    a.spos[d] = make_float4((float)ux, (float)uy, (float)uz, rb.w);
  }
  for (int o = 32; o > 0; o >>= 1) d2 = fmaxf(d2, __shfl_xor(d2, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = red[0];
    for (int w = 1; w < VL_TILE / 64; ++w) m = fmaxf(m, red[w]);
    atomicMax(a.maxd2 + s, __float_as_uint(m));  // non-negative floats order as their bits
  }
}

struct VlBinLds {
  PcBinLds bin;
  SysGeom g;
  int need;
};

// Systems that need a rebuild (first call, or max displacement > skin/2):
// geometry for rc + skin and the counting sort, one workgroup per system.
// VL_BIN_THREADS lanes and at most 96 VGPRs per lane:
// small enough to share CUs with kernels on another stream, since only the
// few systems whose trigger fired are binned in a typical call.
__global__ __launch_bounds__(VL_BIN_THREADS) __attribute__((amdgpu_num_vgpr(96))) void vl_bin_kernel(
    VlArgs a) {
  __shared__ VlBinLds L;
  const int tid = threadIdx.x;
  const float half2 = (float)(0.25 * a.skin * a.skin * (1.0 - 1e-6));
  for (int s = blockIdx.x; s < a.n_sys; s += gridDim.x) {
    const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
    __syncthreads();
    if (tid == 0) {
      L.need = !a.valid[s] || __uint_as_float(a.maxd2[s]) > half2;
      a.need[s] = L.need;
      if (L.need) {
        double rc = 0.0;
        if (a.has_lj) rc = a.lj_rc[s];
        if (a.has_coul && a.cl_rc[s] > rc) rc = a.cl_rc[s];
        SysGeom g0;
        make_geom(a.tril + 6 * s, rc, PC_MAX_NC_AXIS, g0, 2);
        a.reprc[s] = g0.rep;
        make_geom(a.tril + 6 * s, rc + a.skin, PC_MAX_NC_AXIS, L.g, 2);
        a.geom[s] = L.g;
        a.valid[s] = 1;
        atomicAdd(a.counters, 1ull);
      }
    }
    __syncthreads();
    if (!L.need) continue;
    PcSorted o{a.spos, a.sidx, a.scell, a.sspec, a.slj, a.scl,
               a.cstart + (size_t)s * (PC_NCMAX + 1)};
    pc_bin_system(L.g, i0, n, a.pos, a.q, a.species, a.has_lj_excl ? a.lj_excl : nullptr,
                  a.has_coul_excl ? a.cl_excl : nullptr, o, L.bin);
  }
}

// Lists of the rebuilt systems, one wave per cell of the compact work list
// (cell_off).  The wave walks the cell's stencil once per group of up to 64
// of its atoms: a row (dz, dy) of 2 h0 + 1 cells is one contiguous slot
// range of the sorted system per image shift along x (one piece, two where
// it wraps).  For a batch of rows the lanes (one per row) write the pieces
// as descriptors with their offsets in the batch's candidate stream (all
// rows' cell starts loaded at once); the wave then copies the stream, 64
// positions at a time with each lane finding its piece among the few that
// overlap (one candidate load per lane), moved by the image shift and with
// the packed entries, into a window of VL_LC_W in LDS (the stencil's pieces
// back to back, so the chunks below are full whatever the row lengths).
// Each time the window fills, and at the end, the group's
// atoms take it VL_LC_AB at a time (positions and counts uniform) in chunks
// of 64 candidates; the accepted lanes of a chunk append their entries in
// lane order (ballot and prefix count), so an atom's entries come in the
// stencil order (dz, dy, dx) and slot order within a cell.  The test is FP32
// with a relative margin of 1e-5 on (rc + skin)^2: the list only has to hold
// every pair that can come within rc before the next rebuild (a few pairs
// just beyond rc + skin cost nothing), the pair pass decides the cutoff.
// Orthorhombic cells skip the rows and the cells of a row whose gap to the
// atom's cell exceeds rc + skin (margin 1e-4 L for a cell assignment a few
// ulp off a face).  Entries live in groups of four, groups entry-major over
// the atoms (entry e of atom d at word 4 ((e >> 2) n_atoms + d) + (e & 3));
// the last group is padded with the atom's own slot, masked off by the count.
// WIDE: 64-bit word offsets (large batches).
//
// The grid is persistent, VL_XCDS workgroups per round: die x (workgroups
// w with w % VL_XCDS == x) takes the x-th eighth of the work list, so each die
// walks a contiguous stretch of systems, and its waves take its cells in
// turn until the list ends (a call that rebuilds nothing costs one load per
// wave).  The build is latency-bound (a chain of ballot, prefix count and
// store per atom and chunk): it runs best at five waves per SIMD, so its
// registers and LDS are sized for that.
constexpr int VL_LC_THREADS = 256;  // four workers (waves) per workgroup
constexpr int VL_LC_AB = 8;         // atoms of a cell per pass over the window
constexpr int VL_LC_W = 512;        // candidates per window (8 KB of LDS per wave)
constexpr int VL_LC_WG_PER_CU = 4;  // persistent workgroups per CU (measured: 5 is no faster)
constexpr int VL_LC_ND = 128;       // stencil piece descriptors per batch of rows (2 KB per wave)
constexpr uint32_t VL_UNPACKABLE = 0xffffffffu;  // entry of an image shift beyond VL_SHIFT_MAX

// A stencil piece of the list build: first slot, offset in the batch's
// candidate stream, image shift (signed bytes x, y, z) and slot count.
struct VlPiece {
  int tb, pos, shifts, len;
};

// One chunk of 64 window candidates (this lane's: position xj, yj, zj and
// packed entry ent) against the block's nl atoms: entries appended in lane
// order at each atom's count (entry e of block atom k at word
// 4 (e >> 2) N + 4 k + (e & 3) of blk).  SELF: the chunk may hold the block's
// atoms themselves, atom k at lane offset own == k.  BAD: the chunk may hold
// entries of unpackable shifts; a pair on one sends the atom to the stencil.
template <bool WIDE, bool SELF, bool BAD>
__device__ __forceinline__ void vl_list_chunk(uint32_t* blk, unsigned N, int cap, int nl, int own,
                                              float xj, float yj, float zj, uint32_t ent,
                                              const float (&rx)[VL_LC_AB],
                                              const float (&ry)[VL_LC_AB],
                                              const float (&rz)[VL_LC_AB], float rl2,
                                              int (&cnt)[VL_LC_AB]) {
  // lanes holding unpackable entries (the whole chunk's, once)
  const uint64_t ubad = BAD ? __builtin_amdgcn_ballot_w64(ent == VL_UNPACKABLE) : 0ull;
#pragma unroll
  for (int k = 0; k < VL_LC_AB; ++k) {
    if (k < nl) {
      // the accepted lanes as a mask straight from the compare (d^2 < rl2)
      uint64_t m = __builtin_amdgcn_fcmpf(norm2_32(xj - rx[k], yj - ry[k], zj - rz[k]), rl2, 4);
      if (SELF) m &= ~__builtin_amdgcn_ballot_w64(own == k);
      if (m) {
        const int c0 = cnt[k];
        cnt[k] = c0 + (int)__popcll(m);
        if (BAD && (m & ubad)) {
          cnt[k] = max(cnt[k], cap + 1);
          continue;
        }
        // entries past cap (an overflowing atom) are not written
        const unsigned e = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, (uint32_t)c0));
        if (__builtin_amdgcn_inverse_ballot_w64(m) && e < (unsigned)cap) {
          if (WIDE) blk[4 * k + ((size_t)(e >> 2) * N << 2) + (e & 3)] = ent;
          else blk[4 * k + (size_t)(__umul24(e >> 2, N << 2) + (e & 3))] = ent;
        }
      }
    }
  }
}

template <bool WIDE>
__global__ __launch_bounds__(VL_LC_THREADS) void vl_list_kernel(VlArgs a) {
  __shared__ float4 vl_win[VL_LC_THREADS / 64][VL_LC_W];
  __shared__ VlPiece vl_desc[VL_LC_THREADS / 64][VL_LC_ND];
  const int T = a.cell_off[a.n_sys];
  if (T == 0) return;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  float4* win = vl_win[wv];
  VlPiece* desc = vl_desc[wv];
  const int die = blockIdx.x % VL_XCDS;
  const int wpd = (int)(gridDim.x / VL_XCDS) * (VL_LC_THREADS / 64);  // workers per die
  const int per = (T + VL_XCDS - 1) / VL_XCDS;
  const int u_end = min(T, (die + 1) * per);
  const unsigned N = (unsigned)a.n_atoms;
  const int cap = a.cap;  // a multiple of four
  int s = -1;
  for (int u = die * per + (int)(blockIdx.x / VL_XCDS) * (VL_LC_THREADS / 64) + wv; u < u_end;
       u += wpd) {
    // the system of cell u: the last one whose first cell is <= u (the
    // units of one worker ascend, so after the first search it only advances)
    if (s < 0) {
      int lo = 0, hi = a.n_sys;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (a.cell_off[mid] <= u) lo = mid; else hi = mid;
      }
      s = lo;
    } else {
      while (a.cell_off[s + 1] <= u) ++s;
    }
    const int c = u - a.cell_off[s];
    const int i0 = a.sys_off[s];
    const int* cs = a.cstart + (size_t)s * (PC_NCMAX + 1);
    const int a0 = cs[c], a1 = cs[c + 1];
    if (a0 == a1) continue;
    const SysGeom& g = a.geom[s];
    const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];
    const int h0 = g.h[0], h1 = g.h[1], h2 = g.h[2];
    const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
    const float Lf[6] = {g.Lf[0], g.Lf[1], g.Lf[2], g.Lf[3], g.Lf[4], g.Lf[5]};
    const float rl = (float)g.rc;  // rc + skin
    const float rl2 = rl * rl * (1.0f + 1e-5f);
    const bool prune = g.ortho;
    const float lm = rl * (1.0f + 1e-5f) + 1e-4f * fmaxf(Lf[0], fmaxf(Lf[2], Lf[5]));
    const float lim = prune ? lm * lm : INFINITY;
    // squared gap along an axis between cell k and cell k + d (of width w)
    const float wx = Lf[0] / (float)ncx, wy = Lf[2] / (float)ncy, wz = Lf[5] / (float)ncz;
    auto gap2 = [](int d, float w) {
      const float t = (float)max(abs(d) - 1, 0) * w;
      return t * t;
    };
    const float4* sp = a.spos + i0;
    const int* ssp = a.sspec + i0;
    const int nry = 2 * h1 + 1, nrows = (2 * h2 + 1) * nry;
    const int pmax = (2 * h0) / ncx + 2;  // pieces of constant x shift per row, at most
    if (pmax > VL_LC_ND) {
      // more image pieces per row than the descriptors hold: no lists, the
      // pair passes walk the stencil for these atoms
      for (int k = a0 + lane; k < a1; k += 64) {
        a.nlist[i0 + k] = cap + 1;
        a.sref[i0 + k] = sp[k];
        atomicAdd(a.counters + 1, 1ull);
      }
      continue;
    }
    const int rows_per_batch = min(64, VL_LC_ND / pmax);
    for (int ag = a0; ag < a1; ag += 64) {
      const int na = min(64, a1 - ag);
      int vcnt = 0;          // lane k: entries of atom ag + k so far
      int wn = 0;            // candidates in the window
      int G = 0;             // stream position of the window's first candidate
      int selfp = -(1 << 30);  // stream position of slot ag (its own cell, zero shift)
      bool bad = false;      // the window holds entries of unpackable shifts
      // the group's atoms against the window, VL_LC_AB at a time
      auto flush = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (int kb = 0; kb < na; kb += VL_LC_AB) {
          const int nl = min(VL_LC_AB, na - kb);
          float rx[VL_LC_AB], ry[VL_LC_AB], rz[VL_LC_AB];
          int cnt[VL_LC_AB];
#pragma unroll
          for (int k = 0; k < VL_LC_AB; ++k) {
            const float4 p = sp[ag + kb + min(k, nl - 1)];
            rx[k] = p.x; ry[k] = p.y; rz[k] = p.z;
            cnt[k] = __builtin_amdgcn_readlane(vcnt, kb + min(k, nl - 1));
          }
          uint32_t* blk = a.list + 4 * ((size_t)i0 + ag + kb);  // group 0 of the block's atom 0
          const int sb = selfp + kb - G;  // window offset of the block's atom 0 itself
          for (int c0 = 0; c0 < wn; c0 += 64) {
            const float4 q = win[min(c0 + lane, wn - 1)];
            const float xj = c0 + lane < wn ? q.x : 1e30f;  // lanes past the window: far
            const uint32_t ent = __float_as_uint(q.w);
            const int own = c0 + lane - sb;
            const bool self = sb + nl > c0 && sb < c0 + 64;
            if (bad) {
              vl_list_chunk<WIDE, true, true>(blk, N, cap, nl, own, xj, q.y, q.z, ent, rx, ry, rz,
                                              rl2, cnt);
            } else if (self) {
              vl_list_chunk<WIDE, true, false>(blk, N, cap, nl, own, xj, q.y, q.z, ent, rx, ry, rz,
                                               rl2, cnt);
            } else {
              vl_list_chunk<WIDE, false, false>(blk, N, cap, nl, own, xj, q.y, q.z, ent, rx, ry,
                                                rz, rl2, cnt);
            }
          }
#pragma unroll
          for (int k = 0; k < VL_LC_AB; ++k)
            if (k < nl) vcnt = lane == kb + k ? cnt[k] : vcnt;
        }
        __builtin_amdgcn_wave_barrier();
        G += wn;
        wn = 0;
        bad = false;
      };
      // rows (dz, dy) of the stencil in batches whose pieces fit the
      // descriptors; per batch, one lane per row writes its pieces (one
      // contiguous slot range per image shift along x, in dx order) with
      // their offsets in the batch's candidate stream (the cell starts of
      // all rows loaded together), then the window is filled 64 stream
      // positions at a time (one candidate load per lane)
      for (int r0 = 0; r0 < nrows; r0 += rows_per_batch) {
        const int r = r0 + lane;
        int npl = 0, zz = 0, yy = 0, sy = 0, sz = 0, dxa = 0, dxb = -1;
        bool own_row = false;
        if (lane < rows_per_batch && r < nrows) {
          const int dz = r / nry - h2, dy = r % nry - h1;
          own_row = dz == 0 && dy == 0;
          const float gyz = prune ? gap2(dz, wz) + gap2(dy, wy) : 0.0f;
          if (gyz <= lim) {
            zz = cz + dz;
            sz = zz >= 0 ? zz / ncz : -((-zz + ncz - 1) / ncz);
            zz -= sz * ncz;
            yy = cy + dy;
            sy = yy >= 0 ? yy / ncy : -((-yy + ncy - 1) / ncy);
            yy -= sy * ncy;
            dxa = -h0;
            dxb = h0;
            if (prune) {
              while (dxa < 0 && gyz + gap2(dxa, wx) > lim) ++dxa;
              while (dxb > 0 && gyz + gap2(dxb, wx) > lim) --dxb;
            }
            const int xa = cx + dxa, xb = cx + dxb;
            const int fa = xa >= 0 ? xa / ncx : -((-xa + ncx - 1) / ncx);
            const int fb = xb >= 0 ? xb / ncx : -((-xb + ncx - 1) / ncx);
            npl = fb - fa + 1;
          }
        }
        const int* rowp = cs + (zz * ncy + yy) * ncx;
        // the lane's pieces in dx order: pass 1 sums their lengths, pass 2
        // (cached loads) writes the descriptors
        auto pieces = [&](auto&& use) {
          int k = 0;
          for (int xs = cx + dxa; xs <= cx + dxb; ++k) {
            const int sx = xs >= 0 ? xs / ncx : -((-xs + ncx - 1) / ncx);
            const int xe = min(cx + dxb, (sx + 1) * ncx - 1);
            const int tb = rowp[xs - sx * ncx];
            use(k, tb, rowp[xe - sx * ncx + 1] - tb, sx);
            xs = xe + 1;
          }
        };
        int lsum = 0;
        pieces([&](int, int, int len, int) { lsum += len; });
        int incl = lsum, inp = npl;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const int y = __shfl_up(incl, o, 64), z = __shfl_up(inp, o, 64);
          if (lane >= o) { incl += y; inp += z; }
        }
        const int bn = __builtin_amdgcn_readlane(incl, 63);  // candidates of the batch
        const int np = __builtin_amdgcn_readlane(inp, 63);   // pieces of the batch
        int self_off = -1;
        {
          int pos = incl - lsum;
          const int k0 = inp - npl;
          pieces([&](int k, int tb, int len, int sx) {
            desc[k0 + k] = VlPiece{tb, pos, (sx & 255) | ((sy & 255) << 8) | ((sz & 255) << 16), len};
            // the group's atoms themselves: in their own cell at zero shift
            if (own_row && sx == 0 && sy == 0 && sz == 0 && tb <= ag && ag < tb + len)
              self_off = pos + (ag - tb);
            pos += len;
          });
        }
        const uint64_t has_self = __builtin_amdgcn_ballot_w64(self_off >= 0);
        if (has_self)
          selfp = G + wn + __builtin_amdgcn_readlane(self_off, (int)__builtin_ctzll(has_self));
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // fill the window: stream positions c .. c + 63 of the batch; lane j
        // takes the piece holding c + j (pieces kc .. of the batch overlap)
        int kc = 0;
        for (int c = 0; c < bn; c += 64) {
          const int n = min(64, bn - c);
          if (wn + n > VL_LC_W) flush();
          const int p = c + lane;
          int t = 0, psx = 0, psy = 0, psz = 0;
          for (int k = kc; k < np; ++k) {
            const VlPiece d = desc[k];
            if (d.pos >= c + 64) break;
            if (p >= d.pos && p < d.pos + d.len) {
              t = d.tb + (p - d.pos);
              psx = (int)(signed char)d.shifts;
              psy = (int)(signed char)(d.shifts >> 8);
              psz = (int)(signed char)(d.shifts >> 16);
            }
          }
          while (kc < np && desc[kc].pos + desc[kc].len <= c + 64) ++kc;
          const bool packable = psx >= -VL_SHIFT_MAX && psx <= VL_SHIFT_MAX && psy >= -VL_SHIFT_MAX &&
                                psy <= VL_SHIFT_MAX && psz >= -VL_SHIFT_MAX && psz <= VL_SHIFT_MAX;
          if (lane < n) {
            const float4 pj = sp[t];
            float Sx, Sy, Sz;
            shift_vec32(Lf, psx, psy, psz, Sx, Sy, Sz);
            const uint32_t ent = packable ? (uint32_t)t | vl_pack(0, psx, psy, psz, 0) |
                                                ((uint32_t)ssp[t] << 27)
                                          : VL_UNPACKABLE;
            win[wn + lane] = make_float4(pj.x + Sx, pj.y + Sy, pj.z + Sz, __uint_as_float(ent));
          }
          bad = bad || __builtin_amdgcn_ballot_w64(lane < n && !packable) != 0;
          wn += n;
        }
        __builtin_amdgcn_wave_barrier();
      }
      if (wn > 0) flush();
      // per atom of the group (lane k): pad the last group, count, reference
      if (lane < na) {
        const int cn = vcnt;
        const int slot = ag + lane;
        const size_t d = (size_t)i0 + slot;
        if ((cn & 3) && cn < cap) {
          const uint32_t pad = vl_pack(slot, 0, 0, 0, ssp[slot]);
          for (int e = cn; e & 3; ++e) a.list[(((size_t)(e >> 2) * N + d) << 2) + (e & 3)] = pad;
        }
        a.nlist[d] = cn;
        a.sref[d] = sp[slot];
        if (cn > cap) atomicAdd(a.counters + 1, 1ull);
      }
    }
  }
}

// relative half width of the FP32 cutoff band re-decided in FP64: the FP32
// pair vector of positions within 1.5 box lengths of the origin is good to
// ~4e-6 A per component, i.e. ~2e-4 A^2 in r^2 at rc = 12 A, well inside
// 4e-6 rc^2 = 6e-4 A^2
constexpr float VL_BAND32 = 4e-6f;
// band entries kept per atom for the band pass (about 0.4 % of the atoms of
// a 12 A melt meet one at all), and the flag of a lane the band pass walks
// in full
constexpr int VL_NBAND = 4;
constexpr int VL_NBAND_FULL = 255;
// band-entry word: the list entry's slot and shifts, plus which terms had
// the pair inside their band
constexpr uint32_t VL_MID_LJ = 1u << 27, VL_MID_CL = 1u << 28;

// LJ and real-space Ewald terms of one list entry, d = r_j + s@L - r_i in
// FP32.  Each cutoff test is folded into its 1/r factor (zero for a rejected
// pair), so the loop has no data-dependent branch; exclusion ids compare on
// the minimum image only (core/neighborlist.py:731-735).  Adds the pair
// energies to elj / ecl, sets kept (bit 0 LJ, bit 1 Coulomb) and returns
// u'(r) / r.  PASS 0 keeps the pairs below the band of a cutoff and sets
// mid (bit 0 LJ, bit 1 Coulomb) for a pair inside a band; PASS 1 keeps
// exactly the band pairs that the FP64 decision on the caller's positions
// keeps, taking the band bits from given_mid when the main pass stored them
// (the band pass is a launch of its own, so the hot loop makes no calls).
template <bool EXCL, int PASS>
__device__ inline float vl_entry_terms(const PairParams& pp, const float (&band)[4],
                                       const SysGeom& g, int tj, float4 ri, float4 rj, int sx,
                                       int sy, int sz, float r2, int si, int sj, int xlj_i,
                                       int xlj_j, int xcl_i, int xcl_j, float& elj, float& ecl,
                                       int& kept, int& mid, int given_mid = -1) {
  const float K = 14.399645351950548f;  // HARTREE * BOHR, eV A
  float gsum = 0.f;
  bool min_img = true;
  kept = 0;
  mid = 0;
  if (EXCL && g.rep && ((pp.has_lj_excl && xlj_j == xlj_i) || (pp.has_coul_excl && xcl_j == xcl_i)))
    min_img = band_min_image(g, tj, sx, sy, sz);
  // one reciprocal square root serves both terms
  const float ir0 = __builtin_amdgcn_rsqf(r2);
  if (pp.has_lj) {
    const bool lo = r2 < band[0];
    bool m = !lo && r2 < band[1];
    if (PASS == 1 && given_mid >= 0) m = (given_mid & 1) != 0;
    bool keep;
    if (PASS == 0) keep = lo;
    else keep = m && band_keep(g, tj, sx, sy, sz, pp.lj_rc2);
    if (EXCL && pp.has_lj_excl && min_img && xlj_j == xlj_i) keep = false;
    kept |= keep ? 1 : 0;
    mid |= m ? 1 : 0;
    const int p = si * pp.ns + sj;
    const float ir = keep ? ir0 : 0.0f;
    const float ir2 = ir * ir;
    float c6 = pp.sig2[p] * ir2;
    c6 = c6 * c6 * c6;
    elj += pp.e4[p] * c6 * (c6 - 1.0f);
    gsum += pp.e24[p] * c6 * (1.0f - 2.0f * c6) * ir2;
  }
  if (pp.has_coul) {
    const bool lo = r2 < band[2];
    bool m = !lo && r2 < band[3];
    if (PASS == 1 && given_mid >= 0) m = (given_mid & 2) != 0;
    bool keep;
    if (PASS == 0) keep = lo;
    else keep = m && band_keep(g, tj, sx, sy, sz, pp.cl_rc2);
    if (EXCL && pp.has_coul_excl && min_img && xcl_j == xcl_i) keep = false;
    kept |= keep ? 2 : 0;
    mid |= m ? 2 : 0;
    const float ir = keep ? ir0 : 0.0f;
    const float x = pp.alpha * (r2 * ir0);  // alpha r
    const float ex = __expf(-x * x);
    const float ec = erfcx_fit(x) * ex;  // erfc(alpha r)
    const float qq = ri.w * rj.w * K;
    ecl += qq * ec * ir;
    // dU/dr = qq (-2a/sqrt(pi) e^{-a^2 r^2} - erfc / r) / r; u'/r = dU/dr / r
    gsum += qq * (-pp.two_a_sqrtpi * ex - ec * ir0) * ir * ir0;
  }
  return gsum;
}

// The main pass over a lane's list (no exclusion ids, no record): groups of
// four entries, two groups per trip in ping-pong registers so that the next
// group's positions and the group after's list words are in flight while a
// group is evaluated (groups past the end clamp to the last one; entries past
// the count get r^2 = 3e38, above every band, and add nothing).  Partner
// positions are gathered with 32-bit offsets from the system's base.
// ORTHO: the image shift is s_a L_aa per axis (the general form's cross terms
// are exact zeros there, so both give the same bits).  SAME: the LJ and
// Coulomb cutoffs are equal and one band test serves both terms.
template <bool ORTHO, bool SAME, typename Acc>
__device__ inline void vl_walk_main(const VlArgs& a, const PairParams& pp, const float (&band)[4],
                                    const SysGeom& g, int i0, int d, int cnt, float4 ri, int si,
                                    float qiK, float& elj, float& ecl, int& nband,
                                    const Acc& accumulate) {
  const uint4* lp = reinterpret_cast<const uint4*>(a.list) + d;
  const size_t stride = (size_t)a.n_atoms;
  const int ng = (cnt + 3) >> 2;
  const char* sbase = reinterpret_cast<const char*>(a.spos + i0);
  auto group = [&](int q) -> uint4 { return lp[(size_t)min(q, ng - 1) * stride]; };
  auto pos = [&](uint32_t e) -> float4 {
    return *reinterpret_cast<const float4*>(sbase + ((e << 4) & 0x7fff0u));
  };
  auto positions = [&](uint4 e, float4 (&r)[4]) {
    r[0] = pos(e.x); r[1] = pos(e.y); r[2] = pos(e.z); r[3] = pos(e.w);
  };
  const float nl2e = -pp.alpha2 * 1.4426950408889634f;  // -alpha^2 log2(e)
  auto entry = [&](uint32_t e, float4 rj, bool valid) {
    int sx, sy, sz;
    vl_shifts(e, sx, sy, sz);
    float ex, ey, ez;
    if (ORTHO) {
      ex = fmaf((float)sx, g.Lf[0], rj.x) - ri.x;
      ey = fmaf((float)sy, g.Lf[2], rj.y) - ri.y;
      ez = fmaf((float)sz, g.Lf[5], rj.z) - ri.z;
    } else {
      float Sx, Sy, Sz;
      shift_vec32(g.Lf, sx, sy, sz, Sx, Sy, Sz);
      ex = (rj.x + Sx) - ri.x; ey = (rj.y + Sy) - ri.y; ez = (rj.z + Sz) - ri.z;
    }
    const float r2 = valid ? norm2_32(ex, ey, ez) : 3.0e38f;
    float gsum;
    int mid;
    if (SAME) {
      const bool lo = r2 < band[0];
      mid = !lo && r2 < band[1] ? 3 : 0;
      const float ir0 = __builtin_amdgcn_rsqf(r2);
      const float ir = lo ? ir0 : 0.0f;
      // LJ
      const int p = si * pp.ns + (int)(e >> 27);
      const float ir2 = ir * ir;
      float c6 = pp.sig2[p] * ir2;
      c6 = c6 * c6 * c6;
      elj += pp.e4[p] * c6 * (c6 - 1.0f);
      gsum = pp.e24[p] * c6 * (1.0f - 2.0f * c6) * ir2;
      // real-space Ewald: erfc(alpha r) = erfcx(alpha r) exp(-alpha^2 r^2)
      const float x = pp.alpha * (r2 * ir0);
      const float ex2 = __builtin_amdgcn_exp2f(r2 * nl2e);
      const float ec = erfcx_fit(x) * ex2;
      const float qq = qiK * rj.w;
      ecl += qq * ec * ir;
      gsum += qq * (-pp.two_a_sqrtpi * ex2 - ec * ir0) * ir * ir0;
    } else {
      int kept;
      gsum = vl_entry_terms<false, 0>(pp, band, g, (int)(e & 0x7fffu), ri, rj, sx, sy, sz, r2,
                                      si, (int)(e >> 27), 0, 0, 0, 0, elj, ecl, kept, mid);
    }
    if (__builtin_expect(mid != 0, 0)) {
      if (nband < VL_NBAND)
        a.bandbuf[(size_t)d * VL_NBAND + nband] =
            (e & 0x7ffffffu) | (mid & 1 ? VL_MID_LJ : 0u) | (mid & 2 ? VL_MID_CL : 0u);
      ++nband;
    }
    accumulate(gsum, ex, ey, ez);
  };
  auto eval = [&](uint4 e, const float4 (&r)[4], int q) {
    const int k = 4 * q;
    entry(e.x, r[0], k < cnt);
    entry(e.y, r[1], k + 1 < cnt);
    entry(e.z, r[2], k + 2 < cnt);
    entry(e.w, r[3], k + 3 < cnt);
  };
  if (cnt <= 0) return;
  uint4 ga = group(0), gb = group(1);
  float4 ra[4], rb[4];
  positions(ga, ra);
  for (int q = 0; q < ng; q += 2) {
    positions(gb, rb);
    const uint4 gc = group(q + 2);
    eval(ga, ra, q);
    if (q + 1 >= ng) break;
    positions(gc, ra);
    const uint4 gd = group(q + 3);
    eval(gb, rb, q + 1);
    ga = gc;
    gb = gd;
  }
}

// PASS 0: the pair terms of every list entry below the band of a cutoff, one
// lane per sorted atom.  A lane that met up to VL_NBAND entries inside a band
// re-decides them in FP64 on the caller's positions after its walk (the
// decision of the band pass, in the same launch); a lane that met more, or
// met band pairs on the stencil path, is left to PASS 1 and its tile goes on
// a compact list.  PASS 1 (the band pass): a persistent grid over that list,
// each flagged lane walking its list (or stencil) again and adding the band
// pairs the FP64 decision keeps.  Partials of PASS 1 go to the second half of
// a.partial (the reduction reads them for the listed tiles only).
template <bool EXCL, bool VIR, bool REC, int PASS>
__global__ __launch_bounds__(VL_TILE) void vl_pair_kernel(VlArgs a) {
  __shared__ SysGeom g;
  __shared__ float s_sig2[PC_NSPEC * PC_NSPEC], s_e4[PC_NSPEC * PC_NSPEC],
      s_e24[PC_NSPEC * PC_NSPEC];
  __shared__ double red[VL_TILE / 64][PC_NPART];
  // one tile: the main pass's own, or (PASS 1) each tile of the band list
  auto tile_work = [&](int s, int tile) {
  const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
  const int start = tile * VL_TILE;
  const int tid = threadIdx.x;
  double* part = a.partial + ((size_t)(PASS ? a.n_sys : 0) * a.tiles + (size_t)s * a.tiles +
                              tile) * PC_NPART;
  if (n > PC_NMAX) {
    if (tid < PC_NPART) part[tid] = 0.0;
    return;
  }
  if (tid == 0) {
    g = a.geom[s];
    band_ctx = BandCtx{a.spos + i0, a.pos + 3 * (size_t)i0, a.sidx + i0, start};
  }
  const int ns = a.n_species;
  for (int k = tid; k < ns * ns; k += VL_TILE) {
    if (a.has_lj) {
      const double sg = a.sigma[k];
      s_sig2[k] = (float)(sg * sg);
      s_e4[k] = (float)(4.0 * a.epsilon[k]);
      s_e24[k] = (float)(24.0 * a.epsilon[k]);
    }
  }
  __syncthreads();
  PairParams pp;
  pc_pair_params(pp, a.has_lj, a.has_coul, a.has_lj_excl, a.has_coul_excl, a.lj_rc, a.cl_rc,
                 a.alpha, s, ns, s_sig2, s_e4, s_e24);
  const float band[4] = {(float)pp.lj_rc2 * (1.0f - VL_BAND32), (float)pp.lj_rc2 * (1.0f + VL_BAND32),
                         (float)pp.cl_rc2 * (1.0f - VL_BAND32), (float)pp.cl_rc2 * (1.0f + VL_BAND32)};
  const int slot = start + tid;
  float ax = 0.f, ay = 0.f, az = 0.f, elj = 0.f, ecl = 0.f;
  float w00 = 0.f, w01 = 0.f, w02 = 0.f, w11 = 0.f, w12 = 0.f, w22 = 0.f;
  int nband = 0;  // band entries met (PASS 0)
  bool inband = false;  // met a band pair on the stencil path
  const int flag = PASS == 1 && slot < n ? a.bandflag[i0 + slot] : 0;
  if (slot < n && (PASS == 0 || flag)) {
    const int d = i0 + slot;
    const float4 ri = a.spos[d];
    const int si = a.sspec[d];
    const int xlj = EXCL && a.has_lj_excl ? a.slj[d] : 0;
    const int xcl = EXCL && a.has_coul_excl ? a.scl[d] : 0;
    const float qiK = ri.w * 14.399645351950548f;  // q_i HARTREE BOHR (eV A)
    auto accumulate = [&](float gsum, float ex, float ey, float ez) {
      ax += gsum * ex; ay += gsum * ey; az += gsum * ez;
      if (VIR) {
        const float hg = -0.5f * gsum;
        w00 += hg * ex * ex; w01 += hg * ex * ey; w02 += hg * ex * ez;
        w11 += hg * ey * ey; w12 += hg * ey * ez; w22 += hg * ez * ez;
      }
    };
    const int cnt = a.nlist[d];
    // the FP64 decision over the band entries the main pass stored
    auto stored_band = [&](int nb) {
      for (int k = 0; k < nb; ++k) {
        const uint32_t e = a.bandbuf[(size_t)d * VL_NBAND + k];
        const int t = (int)(e & 0x7fffu);
        int sx, sy, sz;
        vl_shifts(e, sx, sy, sz);
        const float4 rj = a.spos[i0 + t];
        float Sx, Sy, Sz;
        shift_vec32(g.Lf, sx, sy, sz, Sx, Sy, Sz);
        const float ex = (rj.x + Sx) - ri.x, ey = (rj.y + Sy) - ri.y, ez = (rj.z + Sz) - ri.z;
        const float r2 = norm2_32(ex, ey, ez);
        int kept, mid;
        const float gsum = vl_entry_terms<EXCL, 1>(
            pp, band, g, t, ri, rj, sx, sy, sz, r2, si, a.sspec[i0 + t], xlj,
            EXCL && a.has_lj_excl ? a.slj[i0 + t] : 0, xcl,
            EXCL && a.has_coul_excl ? a.scl[i0 + t] : 0, elj, ecl, kept, mid,
            (e & VL_MID_LJ ? 1 : 0) | (e & VL_MID_CL ? 2 : 0));
        if (REC && kept) band_record(a.rec, g, i0, t, sx, sy, sz, kept);
        accumulate(gsum, ex, ey, ez);
      }
    };
    if (PASS == 1 && flag <= VL_NBAND) {
      stored_band(flag);
    } else if (cnt <= a.cap && PASS == 0 && !REC && !EXCL) {
      // the hot walk: variants for orthorhombic cells (no cross terms in the
      // image shift) and for equal LJ and Coulomb cutoffs (one band test)
      const bool same = pp.has_lj && pp.has_coul && pp.lj_rc2 == pp.cl_rc2;
      auto walk = [&](auto ortho_t, auto same_t) {
        constexpr bool ORTHO = decltype(ortho_t)::value, SAME = decltype(same_t)::value;
        vl_walk_main<ORTHO, SAME>(a, pp, band, g, i0, d, cnt, ri, si, qiK, elj, ecl, nband,
                                  accumulate);
      };
      if (g.ortho) {
        if (same) walk(std::true_type{}, std::true_type{});
        else walk(std::true_type{}, std::false_type{});
      } else {
        if (same) walk(std::false_type{}, std::true_type{});
        else walk(std::false_type{}, std::false_type{});
      }
    } else if (cnt <= a.cap) {
      // one group of four entries per trip (one 16-byte load); while group q
      // is evaluated, the positions of group q + 1 and the list words of
      // group q + 2 are in flight (groups past the end are clamped to the
      // last one; entries past the count are masked off)
      const uint4* lp = reinterpret_cast<const uint4*>(a.list) + d;
      const size_t stride = (size_t)a.n_atoms;
      const int ng = (cnt + 3) >> 2;
      auto group = [&](int q) -> uint4 { return lp[(size_t)min(q, ng - 1) * stride]; };
      auto positions = [&](uint4 e, float4 (&r)[4]) {
        r[0] = a.spos[i0 + (int)(e.x & 0x7fffu)];
        r[1] = a.spos[i0 + (int)(e.y & 0x7fffu)];
        r[2] = a.spos[i0 + (int)(e.z & 0x7fffu)];
        r[3] = a.spos[i0 + (int)(e.w & 0x7fffu)];
      };
      uint4 ga = make_uint4(0, 0, 0, 0), gb = ga;
      float4 ra[4];
      if (cnt > 0) {
        ga = group(0);
        gb = group(1);
        positions(ga, ra);
      }
      for (int q = 0; q < ng; ++q) {
        float4 rb[4];
        positions(gb, rb);
        const uint4 gc = group(q + 2);
        const uint32_t ew[4] = {ga.x, ga.y, ga.z, ga.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t e = ew[u];
          const float4 rj = ra[u];
          const int t = (int)(e & 0x7fffu);
          int sx, sy, sz;
          vl_shifts(e, sx, sy, sz);
          float Sx, Sy, Sz;
          shift_vec32(g.Lf, sx, sy, sz, Sx, Sy, Sz);
          const float ex = (rj.x + Sx) - ri.x, ey = (rj.y + Sy) - ri.y, ez = (rj.z + Sz) - ri.z;
          // a masked-off entry gets a huge finite r^2 (above every cutoff
          // band; times the zero 1/r of a rejected pair it stays 0)
          const float r2 = 4 * q + u < cnt ? norm2_32(ex, ey, ez) : 3.0e38f;
          int kept, mid;
          const float gsum = vl_entry_terms<EXCL, PASS>(
              pp, band, g, t, ri, rj, sx, sy, sz, r2, si, (int)(e >> 27),
              xlj, EXCL && a.has_lj_excl ? a.slj[i0 + t] : 0, xcl,
              EXCL && a.has_coul_excl ? a.scl[i0 + t] : 0, elj, ecl, kept, mid);
          if (REC && kept) band_record(a.rec, g, i0, t, sx, sy, sz, kept);
          if (PASS == 0 && __builtin_expect(mid != 0, 0)) {
            if (nband < VL_NBAND)
              a.bandbuf[(size_t)d * VL_NBAND + nband] =
                  (e & 0x7ffffffu) | (mid & 1 ? VL_MID_LJ : 0u) | (mid & 2 ? VL_MID_CL : 0u);
            ++nband;
          }
          accumulate(gsum, ex, ey, ez);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) ra[u] = rb[u];
        ga = gb;
        gb = gc;
      }
    } else {
      // overflowed list: the build's 125-cell stencil (cells sized for rc + skin)
      auto visit = [&](int t, float4 rj, int sj, int sx, int sy, int sz, bool self) {
        // FP64 pair vector d = r_j - r_i + s @ L
        double Sx, Sy, Sz;
        shift_vec64(g.L, sx, sy, sz, Sx, Sy, Sz);
        const double exd = (double)rj.x - ((double)ri.x - Sx);
        const double eyd = (double)rj.y - ((double)ri.y - Sy);
        const double ezd = (double)rj.z - ((double)ri.z - Sz);
        const double r2d = norm2_64(exd, eyd, ezd);
        if (!(r2d < pp.all_hi)) return;
        const float ex = (float)exd, ey = (float)eyd, ez = (float)ezd, r2 = (float)r2d;
        int kept;
        const float gsum = pc_pair_terms<EXCL, PASS>(
            pp, g, t, ri, rj, sx, sy, sz, r2d, r2, self, si, sj, xlj,
            EXCL && a.has_lj_excl ? a.slj[i0 + t] : 0, xcl,
            EXCL && a.has_coul_excl ? a.scl[i0 + t] : 0, elj, ecl, kept, inband);
        if (REC && kept) band_record(a.rec, g, i0, t, sx, sy, sz, kept);
        accumulate(gsum, ex, ey, ez);
      };
      const int* cs = a.cstart + (size_t)s * (PC_NCMAX + 1);
      const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];
      const int c = a.scell[d];
      const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
      for (int dz = -g.h[2]; dz <= g.h[2]; ++dz) {
        int zz = cz + dz;
        const int sz = zz >= 0 ? zz / ncz : -((-zz + ncz - 1) / ncz);
        zz -= sz * ncz;
        for (int dy = -g.h[1]; dy <= g.h[1]; ++dy) {
          int yy = cy + dy;
          const int sy = yy >= 0 ? yy / ncy : -((-yy + ncy - 1) / ncy);
          yy -= sy * ncy;
          for (int dx = -g.h[0]; dx <= g.h[0]; ++dx) {
            int xx = cx + dx;
            const int sx = xx >= 0 ? xx / ncx : -((-xx + ncx - 1) / ncx);
            xx -= sx * ncx;
            const bool center = dx == 0 && dy == 0 && dz == 0;
            const int nb = (zz * ncy + yy) * ncx + xx;
            for (int t = cs[nb]; t < cs[nb + 1]; ++t)
              visit(t, a.spos[i0 + t], a.sspec[i0 + t], sx, sy, sz, center && t == slot);
          }
        }
      }
    }
    if (PASS == 0 && !inband && nband > 0 && nband <= VL_NBAND) {
      stored_band(nband);
      nband = 0;  // decided
    }
    const int gi = i0 + a.sidx[d];
    if (PASS == 1 || a.accumulate) {
      a.forces[3 * gi] += ax; a.forces[3 * gi + 1] += ay; a.forces[3 * gi + 2] += az;
    } else {
      a.forces[3 * gi] = ax; a.forces[3 * gi + 1] = ay; a.forces[3 * gi + 2] = az;
    }
  }
  if (PASS == 0) {
    // lanes left to the band pass, and their tile on its list
    const bool full = slot < n && (inband || nband > VL_NBAND);
    if (slot < n) a.bandflag[i0 + slot] = full ? VL_NBAND_FULL : 0;
    const bool any = __syncthreads_or(full);
    if (tid == 0) {
      a.tile_full[(size_t)s * a.tiles + tile] = any;
      if (any) a.band_tiles[atomicAdd(a.band_count, 1)] = s * a.tiles + tile;
    }
  }
  const double v[PC_NPART] = {0.5 * elj, 0.5 * ecl, w00, w01, w02, w11, w12, w22};
  pc_block_partials<VL_TILE>(v, red, part);
  };
  if (PASS == 0) {
    int s, tile;
    if (vl_tile_of(a, s, tile)) tile_work(s, tile);
  } else {
    const int nt = *a.band_count;
    for (int k = blockIdx.x; k < nt; k += gridDim.x) {
      const int lin = a.band_tiles[k];
      tile_work(lin / a.tiles, lin % a.tiles);
      __syncthreads();
    }
  }
}

// Per system: the main pass's tile partials and the band pass's (tiles on
// its list), the system's own tiles only, in a fixed order.
__global__ __launch_bounds__(64) void vl_reduce_kernel(VlArgs a) {
  // one wave per system: lane l sums the partials j = l, l + 64, ... of the
  // list (main pass tiles, then band pass tiles), then a fixed butterfly
  const int s = blockIdx.x, lane = threadIdx.x;
  const int nt = a.tile_off[s + 1] - a.tile_off[s];
  double acc[PC_NPART] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = lane; j < 2 * nt; j += 64) {
    const int h = j >= nt, t = j - h * nt;
    if (h == 1 && !a.tile_full[(size_t)s * a.tiles + t]) continue;  // not on the band list
    const double* p = a.partial + (((size_t)h * a.n_sys + s) * a.tiles + t) * PC_NPART;
    for (int k = 0; k < PC_NPART; ++k) acc[k] += p[k];
  }
  for (int k = 0; k < PC_NPART; ++k) acc[k] = wave_sum(acc[k]);
  if (lane != 0) return;
  if (a.energy) {
    double* e = a.energy + 2 * (size_t)s;
    if (a.accumulate) { e[0] += acc[0]; e[1] += acc[1]; } else { e[0] = acc[0]; e[1] = acc[1]; }
  }
  if (a.virial) {
    const double w[9] = {acc[2], acc[3], acc[4], acc[3], acc[5], acc[6], acc[4], acc[6], acc[7]};
    double* W = a.virial + 9 * (size_t)s;
    for (int k = 0; k < 9; ++k) W[k] = a.accumulate ? W[k] + w[k] : w[k];
  }
}

// Byte offsets of the handle's arrays.
struct VlLayout {
  size_t geom, spos, sref, sidx, scell, sspec, slj, scl, cstart, nlist, list, valid, need, reprc,
      maxd2,
      counters, partial, bandflag, bandbuf, tile_off, cell_off, band_tiles, band_count, tile_full,
      total;
  VlLayout(int n_sys, int n_atoms, int cap, int tiles) {
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
    const size_t na = (size_t)n_atoms, ns = (size_t)n_sys;
    geom = take(sizeof(SysGeom) * ns);
    spos = take(sizeof(float4) * na);
    sref = take(sizeof(float4) * na);
    sidx = take(sizeof(int) * na);
    scell = take(sizeof(int) * na);
    sspec = take(sizeof(int) * na);
    slj = take(sizeof(int) * na);
    scl = take(sizeof(int) * na);
    cstart = take(sizeof(int) * ns * (PC_NCMAX + 1));
    nlist = take(sizeof(int) * na);
    list = take(sizeof(uint32_t) * na * (size_t)cap);
    valid = take(sizeof(int) * ns);
    need = take(sizeof(int) * ns);
    reprc = take(sizeof(int) * ns);
    maxd2 = take(sizeof(unsigned) * ns);
    counters = take(sizeof(unsigned long long) * 2);
    partial = take(2 * sizeof(double) * ns * (size_t)tiles * PC_NPART);
    bandflag = take(na);
    bandbuf = take(sizeof(uint32_t) * na * VL_NBAND);
    tile_off = take(sizeof(int) * (ns + 1));
    cell_off = take(sizeof(int) * (ns + 1));
    band_tiles = take(sizeof(int) * ns * (size_t)tiles);
    band_count = take(sizeof(int));
    tile_full = take(ns * (size_t)tiles);
    total = off;
  }
};

__global__ void vl_reset_counters_kernel(unsigned long long* c) { c[0] = 0; c[1] = 0; }

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_verlet_create(kupsmi_verlet** out) {
  KUPSMI_REQUIRE(out != nullptr, "null output");
  *out = new kupsmi_verlet();
  KUPSMI_HIP(hipGetDevice(&(*out)->device));
  return KUPSMI_OK;
}

extern "C" int kupsmi_verlet_gate(kupsmi_verlet* v, void* event) {
  KUPSMI_REQUIRE(v != nullptr, "null handle");
  v->gate = (hipEvent_t)event;
  return KUPSMI_OK;
}

extern "C" int kupsmi_verlet_destroy(kupsmi_verlet* v) {
  if (!v) return KUPSMI_OK;
  for (auto& e : v->events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  if (v->mem) {
    int cur = 0;
    KUPSMI_HIP(hipGetDevice(&cur));
    KUPSMI_HIP(hipSetDevice(v->device));
    const hipError_t e = hipFree(v->mem);
    KUPSMI_HIP(hipSetDevice(cur));
    KUPSMI_HIP(e);
  }
  delete v;
  return KUPSMI_OK;
}

extern "C" int kupsmi_verlet_profile(kupsmi_verlet* v, int enable, double* total_ms,
                                     long long* calls) {
  KUPSMI_REQUIRE(v != nullptr, "null handle");
  double ms = 0.0;
  long long n = 0;
  for (auto& e : v->events) {
    KUPSMI_HIP(hipEventSynchronize(e.second));
    float t = 0.f;
    KUPSMI_HIP(hipEventElapsedTime(&t, e.first, e.second));
    ms += t;
    ++n;
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  v->events.clear();
  if (total_ms) *total_ms = ms;
  if (calls) *calls = n;
  v->profile = enable != 0;
  return KUPSMI_OK;
}

extern "C" int kupsmi_verlet_record_pairs(kupsmi_verlet* v, int64_t capacity, int32_t* ij,
                                         int32_t* shift, uint8_t* terms,
                                         unsigned long long* count) {
  KUPSMI_REQUIRE(v != nullptr, "null handle");
  KUPSMI_REQUIRE(capacity >= 0 && count != nullptr, "count pointer and capacity >= 0 required");
  KUPSMI_REQUIRE(capacity == 0 || (ij && shift), "ij and shift buffers required");
  v->rec = PairRecord{capacity, ij, shift, terms, count};
  v->rec_armed = 1;
  return KUPSMI_OK;
}

extern "C" int kupsmi_verlet_freeze(kupsmi_verlet* v, int frozen) {
  KUPSMI_REQUIRE(v != nullptr, "null handle");
  v->frozen = frozen != 0;
  return KUPSMI_OK;
}

extern "C" int kupsmi_verlet_invalidate(kupsmi_verlet* v, void* stream) {
  KUPSMI_REQUIRE(v != nullptr, "null handle");
  if (!v->mem || v->n_sys == 0) return KUPSMI_OK;
  VlLayout lay(v->n_sys, v->n_atoms, v->cap, v->tiles);
  int* valid = reinterpret_cast<int*>(static_cast<char*>(v->mem) + lay.valid);
  vl_invalidate_kernel<<<(v->n_sys + 255) / 256, 256, 0, (hipStream_t)stream>>>(v->n_sys, valid);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_verlet_counters(kupsmi_verlet* v, void* stream, unsigned long long* out) {
  KUPSMI_REQUIRE(v != nullptr && out != nullptr, "null argument");
  out[0] = out[1] = 0;
  if (!v->mem) return KUPSMI_OK;
  VlLayout lay(v->n_sys, v->n_atoms, v->cap, v->tiles);
  KUPSMI_HIP(hipMemcpyAsync(out, static_cast<char*>(v->mem) + lay.counters,
                            2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                            (hipStream_t)stream));
  KUPSMI_HIP(hipStreamSynchronize((hipStream_t)stream));
  return KUPSMI_OK;
}

extern "C" int kupsmi_pair_forces_verlet(
    kupsmi_verlet* v, void* stream, int n_sys, const int32_t* sys_off, int n_atoms,
    int max_atoms_per_sys, const float* pos, const float* charges, const int32_t* species,
    int n_species, const double* sigma, const double* epsilon, const double* lj_cutoff,
    const int32_t* lj_exclusion, const double* alpha, const double* coul_cutoff,
    const int32_t* coul_exclusion, const double* tril, double skin, int max_neighbors,
    int accumulate, float* forces, double* energy, double* virial) {
  KUPSMI_REQUIRE(v != nullptr, "null handle");
  KUPSMI_REQUIRE(n_sys >= 0 && n_atoms >= 0, "negative size");
  KUPSMI_REQUIRE(max_atoms_per_sys <= PC_NMAX, "systems of the pair path hold at most 16384 atoms");
  KUPSMI_REQUIRE(n_species >= 1 && n_species <= PC_NSPEC, "n_species must be in [1, 16]");
  KUPSMI_REQUIRE(skin >= 0.0, "skin must be non-negative");
  KUPSMI_REQUIRE(max_neighbors >= 1, "max_neighbors must be positive");
  const bool has_lj = sigma && epsilon && lj_cutoff;
  const bool has_cl = alpha && coul_cutoff && charges;
  KUPSMI_REQUIRE(has_lj || has_cl, "nothing to compute: pass LJ tables or Ewald parameters");
  KUPSMI_REQUIRE(forces != nullptr, "forces output required");
  if (n_sys == 0 || n_atoms == 0) return KUPSMI_OK;
  int dev = 0;
  KUPSMI_HIP(hipGetDevice(&dev));
  KUPSMI_REQUIRE(dev == v->device, "the handle belongs to another device");
  hipStream_t st = (hipStream_t)stream;
  int tiles = (max_atoms_per_sys + VL_TILE - 1) / VL_TILE;
  if (tiles < 1) tiles = 1;
  const int cap4 = (max_neighbors + 3) & ~3;  // lists are stored in groups of four
  if (!v->mem || v->n_sys != n_sys || v->n_atoms != n_atoms || v->cap != cap4 ||
      v->tiles < tiles) {
    KUPSMI_REQUIRE(!v->frozen,
                   "a frozen Verlet handle cannot be resized (recorded work uses its buffers)");
    if (v->mem) {
      KUPSMI_HIP(hipStreamSynchronize(st));
      KUPSMI_HIP(hipFree(v->mem));
      v->mem = nullptr;
    }
    VlLayout lay(n_sys, n_atoms, cap4, tiles);
    void* p = nullptr;
    KUPSMI_HIP(hipMalloc(&p, lay.total));
    v->mem = p; v->bytes = lay.total;
    v->n_sys = n_sys; v->n_atoms = n_atoms; v->cap = cap4; v->tiles = tiles;
    char* b = static_cast<char*>(p);
    vl_invalidate_kernel<<<(n_sys + 255) / 256, 256, 0, st>>>(n_sys, reinterpret_cast<int*>(b + lay.valid));
    KUPSMI_CHECK_LAUNCH();
    vl_reset_counters_kernel<<<1, 1, 0, st>>>(reinterpret_cast<unsigned long long*>(b + lay.counters));
    KUPSMI_CHECK_LAUNCH();
  }
  tiles = v->tiles;
  VlLayout lay(n_sys, n_atoms, v->cap, tiles);
  char* b = static_cast<char*>(v->mem);
  VlArgs a{};
  a.n_sys = n_sys; a.n_atoms = n_atoms; a.n_species = n_species; a.has_lj = has_lj;
  a.has_coul = has_cl; a.has_lj_excl = lj_exclusion != nullptr;
  a.has_coul_excl = coul_exclusion != nullptr; a.accumulate = accumulate;
  a.tiles = tiles; a.cap = v->cap; a.skin = skin;
  a.sys_off = sys_off; a.pos = pos; a.q = charges; a.species = species;
  a.lj_excl = lj_exclusion; a.cl_excl = coul_exclusion; a.sigma = sigma; a.epsilon = epsilon;
  a.lj_rc = lj_cutoff; a.alpha = alpha; a.cl_rc = coul_cutoff; a.tril = tril;
  a.geom = reinterpret_cast<SysGeom*>(b + lay.geom);
  a.spos = reinterpret_cast<float4*>(b + lay.spos);
  a.sref = reinterpret_cast<float4*>(b + lay.sref);
  a.sidx = reinterpret_cast<int*>(b + lay.sidx);
  a.scell = reinterpret_cast<int*>(b + lay.scell);
  a.sspec = reinterpret_cast<int*>(b + lay.sspec);
  a.slj = reinterpret_cast<int*>(b + lay.slj);
  a.scl = reinterpret_cast<int*>(b + lay.scl);
  a.cstart = reinterpret_cast<int*>(b + lay.cstart);
  a.nlist = reinterpret_cast<int*>(b + lay.nlist);
  a.list = reinterpret_cast<uint32_t*>(b + lay.list);
  a.valid = reinterpret_cast<int*>(b + lay.valid);
  a.need = reinterpret_cast<int*>(b + lay.need);
  a.reprc = reinterpret_cast<int*>(b + lay.reprc);
  a.maxd2 = reinterpret_cast<unsigned*>(b + lay.maxd2);
  a.counters = reinterpret_cast<unsigned long long*>(b + lay.counters);
  a.partial = reinterpret_cast<double*>(b + lay.partial);
  a.bandflag = reinterpret_cast<uint8_t*>(b + lay.bandflag);
  a.bandbuf = reinterpret_cast<uint32_t*>(b + lay.bandbuf);
  a.tile_off = reinterpret_cast<int*>(b + lay.tile_off);
  a.cell_off = reinterpret_cast<int*>(b + lay.cell_off);
  a.band_tiles = reinterpret_cast<int*>(b + lay.band_tiles);
  a.band_count = reinterpret_cast<int*>(b + lay.band_count);
  a.tile_full = reinterpret_cast<uint8_t*>(b + lay.tile_full);
  a.forces = forces; a.energy = energy; a.virial = virial;
  int ncu = 256;
  KUPSMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  vl_prep_kernel<<<1, VL_PREP_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  // one workgroup per (system, tile), compact: at most one partial tile per system
  const dim3 grid(vl_grid(n_atoms, n_sys, VL_TILE));
  vl_gather_kernel<<<grid, VL_TILE, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  vl_bin_kernel<<<n_sys < ncu ? n_sys : ncu, VL_BIN_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  vl_cell_scan_kernel<<<1, VL_PREP_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  {
    // persistent list build: VL_LC_WG_PER_CU workgroups per CU, a whole
    // number of rounds over the dies
    const unsigned wgs = (unsigned)(VL_XCDS * ((VL_LC_WG_PER_CU * ncu + VL_XCDS - 1) / VL_XCDS));
    // 32-bit entry offsets: 4 n_atoms a 24-bit factor, cap n_atoms words
    const bool narrow = 4ull * (unsigned)n_atoms < (1ull << 24) &&
                        (unsigned long long)v->cap * (unsigned)n_atoms < (1ull << 32);
    if (narrow) vl_list_kernel<false><<<wgs, VL_LC_THREADS, 0, st>>>(a);
    else vl_list_kernel<true><<<wgs, VL_LC_THREADS, 0, st>>>(a);
    KUPSMI_CHECK_LAUNCH();
  }
  if (v->gate) {
    KUPSMI_HIP(hipStreamWaitEvent(st, v->gate, 0));
    v->gate = nullptr;
  }
  const bool excl = a.has_lj_excl || a.has_coul_excl, vir = a.virial != nullptr;
  // main pass, then the band pass; with a pending pair record both run as the
  // recording instantiation (the same pair decisions)
  const bool rec = v->rec_armed != 0;
  if (rec) {
    a.rec = v->rec;
    v->rec_armed = 0;
    record_clear_kernel<<<1, 1, 0, st>>>(a.rec.count);
    KUPSMI_CHECK_LAUNCH();
  }
#define VL_PASS(P, G)                                                            \
  if (rec) {                                                                     \
    if (excl) vl_pair_kernel<true, true, true, P><<<G, VL_TILE, 0, st>>>(a);     \
    else vl_pair_kernel<false, true, true, P><<<G, VL_TILE, 0, st>>>(a);         \
  } else if (excl) {                                                             \
    if (vir) vl_pair_kernel<true, true, false, P><<<G, VL_TILE, 0, st>>>(a);     \
    else vl_pair_kernel<true, false, false, P><<<G, VL_TILE, 0, st>>>(a);        \
  } else {                                                                       \
    if (vir) vl_pair_kernel<false, true, false, P><<<G, VL_TILE, 0, st>>>(a);    \
    else vl_pair_kernel<false, false, false, P><<<G, VL_TILE, 0, st>>>(a);       \
  }                                                                              \
  KUPSMI_CHECK_LAUNCH();
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (v->profile) {
    KUPSMI_HIP(hipEventCreate(&ev0));
    KUPSMI_HIP(hipEventCreate(&ev1));
    KUPSMI_HIP(hipEventRecord(ev0, st));
  }
  VL_PASS(0, grid)
  if (v->profile) {
    KUPSMI_HIP(hipEventRecord(ev1, st));
    v->events.emplace_back(ev0, ev1);
  }
  // the band pass: persistent, over the tiles the main pass listed (usually
  // none: a few workgroups read the count and leave)
  const unsigned band_grid = (unsigned)min(grid.x, (unsigned)(2 * ncu));
  VL_PASS(1, band_grid)
#undef VL_PASS
  vl_reduce_kernel<<<n_sys, 64, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}
