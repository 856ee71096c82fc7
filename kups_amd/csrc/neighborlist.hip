// Neighbor-list export: the directed pair set of every system, decided in
// FP64 with the reference's fractional-coordinate formula
// (core/neighborlist.py:542-564, 589, 595-624, 731-735).
//
// The export exists for parity and for callers that consume kUPS Edges; the
// MD and GCMC kernels never materialise the list.  Each system is scanned
// densely (every ordered pair, like DenseNearestNeighborList, 902-982), one
// thread per left atom with the system's atoms streamed through L1/L2, in two
// passes: count per atom, exclusive scan, then write at the scanned offsets.
// Edge order is i-major, then j, then image offset (o0, o1, o2 lexicographic),
// so the output is deterministic.
#include <hipcub/hipcub.hpp>
#include "common.hpp"

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

#pragma clang fp contract(off)
__device__ inline void frac64(const double* li, const float* r, double& f0, double& f1,
                              double& f2) {
  const double x = r[0], y = r[1], z = r[2];
  f0 = x * li[0] + y * li[1] + z * li[3];
  f1 = y * li[2] + z * li[4];
  f2 = z * li[5];
}

__device__ inline double dist2_64(const double* L, double d0, double d1, double d2) {
  const double x = d0 * L[0] + d1 * L[1] + d2 * L[3];
  const double y = d1 * L[2] + d2 * L[4];
  const double z = d2 * L[5];
  return x * x + y * y + z * z;
}

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
      frac64(g.Li, a.pos + 3 * i, fi0, fi1, fi2);
      const int ei = a.excl[i];
      int64_t cnt = 0;
      int64_t out = WRITE ? a.offsets[i] : 0;
      for (int j = i0; j < i1; ++j) {
        double fj0, fj1, fj2;
        frac64(g.Li, a.pos + 3 * j, fj0, fj1, fj2);
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
