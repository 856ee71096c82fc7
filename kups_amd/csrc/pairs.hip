// General pair path: LJ + Ewald real-space energies, forces and virials of
// systems of any size up to 16384 atoms, state in HBM, one launch per
// evaluation.  Used by the Ewald MD step, by GCMC set-up and wherever a
// system does not fit the LDS-resident MD kernel.
//
//   bin_kernel   one workgroup per system: wrap, counting sort of the atoms
//                into cells at least rc/2 wide (stable: ties keep atom order),
//                sorted copies (x, y, z, q | species | exclusion ids) to HBM.
//   pair_kernel  one thread per sorted atom, 125-cell stencil read through
//                L1/L2 (neighbouring lanes read the same candidates), pair
//                vectors in FP64, potential arithmetic in FP32, FP64
//                re-decision of the cutoff band, per-block FP64 partial
//                energies and virials.
//   reduce_kernel  fixed-order FP64 sum of the partials of each system.
//
// Reference: potential/classical/lennard_jones.py:149-178 (LJ per directed
// edge, sum / 2), potential/classical/ewald.py:305-325 (real-space sum
// K/2 sum q_i q_j erfc(alpha r)/r [r < rc]), neighbor semantics of
// core/neighborlist.py:692-786 (strict cutoff, image grid for narrow cells,
// exclusion on the minimum image only, 731-735).
#include <type_traits>
#include "pair_common.hpp"

namespace kupsmi {

struct PairArgs {
  int n_sys, n_species, has_lj, has_coul, has_lj_excl, has_coul_excl, accumulate;
  int tiles;
  const int32_t* sys_off;
  const float* pos;
  const float* q;
  const int32_t* species;
  const int32_t* lj_excl;
  const int32_t* cl_excl;
  const double* sigma;
  const double* epsilon;
  const double* lj_rc;
  const double* lj_rtr;  // per-system LJ switch start (NULL: no switch)
  const double* alpha;
  const double* cl_rc;
  const double* tril;
  const SysGeom* geom;
  // sorted copies
  float4* spos;
  int* sidx;
  int* scell;
  int* sspec;
  int* slj;
  int* scl;
  int* cstart;
  double* partial;      // main pass, then the band pass (n_sys x tiles x PC_NPART each)
  uint8_t* bandflag;    // per sorted atom: the main pass met a pair in a cutoff band
  float* forces;
  double* energy;
  double* virial;
  PairRecord rec;  // parity output (REC instantiation only)
};

__global__ void pc_geom_kernel(PairArgs a) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n_sys) return;
  double rc = 0.0;
  if (a.has_lj) rc = a.lj_rc[s];
  if (a.has_coul && a.cl_rc[s] > rc) rc = a.cl_rc[s];
  make_geom(a.tril + 6 * s, rc, PC_MAX_NC_AXIS, const_cast<SysGeom*>(a.geom)[s], 2);
}

__global__ __launch_bounds__(PC_BIN_THREADS) void pc_bin_kernel(PairArgs a) {
  __shared__ PcBinLds L;
  __shared__ SysGeom g;
  const int tid = threadIdx.x;
  for (int s = blockIdx.x; s < a.n_sys; s += gridDim.x) {
    const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
    __syncthreads();
    if (tid == 0) g = a.geom[s];
    __syncthreads();
    PcSorted o{a.spos, a.sidx, a.scell, a.sspec, a.slj, a.scl, a.cstart + (size_t)s * (PC_NCMAX + 1)};
    pc_bin_system(g, i0, n, a.pos, a.q, a.species, a.has_lj_excl ? a.lj_excl : nullptr,
                  a.has_coul_excl ? a.cl_excl : nullptr, o, L);
  }
}

// PASS 0: the pair terms of every candidate below the band of a cutoff, one
// lane per sorted atom; PASS 1 (the band pass): lanes whose PASS 0 met a band
// pair add the band pairs the FP64 decision on the caller's positions keeps
// (a launch of its own, so the hot loop makes no calls).  Partials of PASS 1
// go to the second half of a.partial.
template <bool EXCL, bool REC, int PASS>
__global__ __launch_bounds__(PC_TILE) void pc_pair_kernel(PairArgs a) {
  __shared__ SysGeom g;
  __shared__ float s_sig2[PC_NSPEC * PC_NSPEC], s_e4[PC_NSPEC * PC_NSPEC],
      s_e24[PC_NSPEC * PC_NSPEC];
  __shared__ double red[PC_TILE / 64][PC_NPART];
  const int s = blockIdx.y;
  const int i0 = a.sys_off[s], n = a.sys_off[s + 1] - i0;
  const int start = blockIdx.x * PC_TILE;
  const int tid = threadIdx.x;
  double* part = a.partial + ((size_t)(PASS ? a.n_sys : 0) * a.tiles + (size_t)s * a.tiles +
                              blockIdx.x) * PC_NPART;
  if (start >= n || n > PC_NMAX) {
    if (tid < PC_NPART) part[tid] = 0.0;
    return;
  }
  if (PASS == 1) {
    if (!__syncthreads_or(start + tid < n && a.bandflag[i0 + start + tid])) {
      if (tid < PC_NPART) part[tid] = 0.0;
      return;
    }
  }
  if (tid == 0) {
    g = a.geom[s];
    band_ctx = BandCtx{a.spos + i0, a.pos + 3 * (size_t)i0, a.sidx + i0,
                       (int)blockIdx.x * PC_TILE};
  }
  const int ns = a.n_species;
  for (int k = tid; k < ns * ns; k += PC_TILE) {
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
                 a.alpha, s, ns, s_sig2, s_e4, s_e24, a.lj_rtr);
  const double all_hi = pp.all_hi;
  const int* cs = a.cstart + (size_t)s * (PC_NCMAX + 1);
  const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2];

  const int slot = start + tid;
  float ax = 0.f, ay = 0.f, az = 0.f, elj = 0.f, ecl = 0.f;
  float w00 = 0.f, w01 = 0.f, w02 = 0.f, w11 = 0.f, w12 = 0.f, w22 = 0.f;
  bool inband = false;
  if (slot < n && (PASS == 0 || a.bandflag[i0 + slot])) {
    const float4 ri = a.spos[i0 + slot];
    const int c = a.scell[i0 + slot];
    const int si = a.sspec[i0 + slot];
    const int xlj = EXCL && a.has_lj_excl ? a.slj[i0 + slot] : 0;
    const int xcl = EXCL && a.has_coul_excl ? a.scl[i0 + slot] : 0;
    const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
    {
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
            // pair vectors are formed in FP64 (same VALU rate as unpacked FP32
            // on gfx950): |r| up to half a box would otherwise cost ~1e-5
            // relative
            double Sx, Sy, Sz;
            shift_vec64(g.L, sx, sy, sz, Sx, Sy, Sz);
            const double qx = (double)ri.x - Sx, qy = (double)ri.y - Sy, qz = (double)ri.z - Sz;
            const bool center = dx == 0 && dy == 0 && dz == 0;
            const int nb = (zz * ncy + yy) * ncx + xx;
            const int tb = cs[nb], te = cs[nb + 1];
            for (int t = tb; t < te; ++t) {
              const float4 rj = a.spos[i0 + t];
              const double exd = (double)rj.x - qx, eyd = (double)rj.y - qy,
                           ezd = (double)rj.z - qz;
              const double r2d = norm2_64(exd, eyd, ezd);
              if (!(r2d < all_hi)) continue;
              const float ex = (float)exd, ey = (float)eyd, ez = (float)ezd, r2 = (float)r2d;
              // particle exclusion: the zero-shift self pair
              const bool self = center && t == slot;
              int kept;
              const float gsum = pc_pair_terms<EXCL, PASS>(
                  pp, g, t, ri, rj, sx, sy, sz, r2d, r2, self, si, a.sspec[i0 + t], xlj,
                  EXCL && a.has_lj_excl ? a.slj[i0 + t] : 0, xcl,
                  EXCL && a.has_coul_excl ? a.scl[i0 + t] : 0, elj, ecl, kept, inband);
              if (REC && kept) band_record(a.rec, g, i0, t, sx, sy, sz, kept);
              ax += gsum * ex; ay += gsum * ey; az += gsum * ez;
              const float hg = -0.5f * gsum;
              w00 += hg * ex * ex; w01 += hg * ex * ey; w02 += hg * ex * ez;
              w11 += hg * ey * ey; w12 += hg * ey * ez; w22 += hg * ez * ez;
            }
          }
        }
      }
    }
    const int gi = i0 + a.sidx[i0 + slot];
    if (PASS == 1 || a.accumulate) {
      a.forces[3 * gi] += ax; a.forces[3 * gi + 1] += ay; a.forces[3 * gi + 2] += az;
    } else {
      a.forces[3 * gi] = ax; a.forces[3 * gi + 1] = ay; a.forces[3 * gi + 2] = az;
    }
  }
  if (PASS == 0 && slot < n) a.bandflag[i0 + slot] = inband;
  const double v[PC_NPART] = {0.5 * elj, 0.5 * ecl, w00, w01, w02, w11, w12, w22};
  pc_block_partials<PC_TILE>(v, red, part);
}

__global__ void pc_reduce_kernel(PairArgs a) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= a.n_sys) return;
  pc_reduce_system(s, a.tiles, a.partial, a.accumulate, a.energy, a.virial,
                   a.partial + (size_t)a.n_sys * a.tiles * PC_NPART);
}

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_pair_forces(kupsmi_ctx* ctx, void* stream, int n_sys, const int32_t* sys_off,
                                  int n_atoms, int max_atoms_per_sys, const float* pos,
                                  const float* charges, const int32_t* species, int n_species,
                                  const double* sigma, const double* epsilon,
                                  const double* lj_cutoff, const int32_t* lj_exclusion,
                                  const double* alpha, const double* coul_cutoff,
                                  const int32_t* coul_exclusion, const double* tril,
                                  int accumulate, float* forces, double* energy, double* virial,
                                  const double* lj_truncation) {
  KUPSMI_REQUIRE(ctx != nullptr, "null context");
  // a pending pair record is taken by this call
  const PairRecord rec = ctx->rec_armed ? ctx->rec : PairRecord{};
  ctx->rec_armed = 0;
  KUPSMI_REQUIRE(n_sys >= 0 && n_atoms >= 0, "negative size");
  KUPSMI_REQUIRE(max_atoms_per_sys <= PC_NMAX, "systems of the pair path hold at most 16384 atoms");
  KUPSMI_REQUIRE(n_species >= 1 && n_species <= PC_NSPEC, "n_species must be in [1, 16]");
  const bool has_lj = sigma && epsilon && lj_cutoff;
  const bool has_cl = alpha && coul_cutoff && charges;
  KUPSMI_REQUIRE(has_lj || has_cl, "nothing to compute: pass LJ tables or Ewald parameters");
  KUPSMI_REQUIRE(forces != nullptr, "forces output required");
  if (n_sys == 0 || n_atoms == 0) return KUPSMI_OK;
  hipStream_t st = (hipStream_t)stream;
  PairArgs a{};
  a.n_sys = n_sys; a.n_species = n_species; a.has_lj = has_lj; a.has_coul = has_cl;
  a.has_lj_excl = lj_exclusion != nullptr; a.has_coul_excl = coul_exclusion != nullptr;
  a.accumulate = accumulate;
  a.tiles = (max_atoms_per_sys + PC_TILE - 1) / PC_TILE;
  if (a.tiles < 1) a.tiles = 1;
  a.sys_off = sys_off; a.pos = pos; a.q = charges; a.species = species;
  a.lj_excl = lj_exclusion; a.cl_excl = coul_exclusion; a.sigma = sigma; a.epsilon = epsilon;
  a.lj_rtr = lj_truncation;
  a.lj_rc = lj_cutoff; a.alpha = alpha; a.cl_rc = coul_cutoff; a.tril = tril;
  a.forces = forces; a.energy = energy; a.virial = virial;
  // scratch layout (slot 3)
  const size_t na = (size_t)n_atoms;
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
  const size_t o_geom = take(sizeof(SysGeom) * n_sys);
  const size_t o_spos = take(sizeof(float4) * na);
  const size_t o_sidx = take(sizeof(int) * na);
  const size_t o_scell = take(sizeof(int) * na);
  const size_t o_sspec = take(sizeof(int) * na);
  const size_t o_slj = take(sizeof(int) * na);
  const size_t o_scl = take(sizeof(int) * na);
  const size_t o_cs = take(sizeof(int) * (size_t)n_sys * (PC_NCMAX + 1));
  const size_t o_part = take(2 * sizeof(double) * (size_t)n_sys * a.tiles * PC_NPART);
  const size_t o_flag = take(na);
  void* buf = nullptr;
  int rc = scratch(ctx, 3, off, &buf);
  if (rc) return rc;
  char* b = static_cast<char*>(buf);
  a.geom = reinterpret_cast<SysGeom*>(b + o_geom);
  a.spos = reinterpret_cast<float4*>(b + o_spos);
  a.sidx = reinterpret_cast<int*>(b + o_sidx);
  a.scell = reinterpret_cast<int*>(b + o_scell);
  a.sspec = reinterpret_cast<int*>(b + o_sspec);
  a.slj = reinterpret_cast<int*>(b + o_slj);
  a.scl = reinterpret_cast<int*>(b + o_scl);
  a.cstart = reinterpret_cast<int*>(b + o_cs);
  a.partial = reinterpret_cast<double*>(b + o_part);
  a.bandflag = reinterpret_cast<uint8_t*>(b + o_flag);
  pc_geom_kernel<<<(n_sys + 255) / 256, 256, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  int dev = 0, ncu = 256;
  KUPSMI_HIP(hipGetDevice(&dev));
  KUPSMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  pc_bin_kernel<<<n_sys < ncu ? n_sys : ncu, PC_BIN_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  dim3 grid(a.tiles, n_sys);
  const bool excl = a.has_lj_excl || a.has_coul_excl;
  if (rec.count) {
    a.rec = rec;
    record_clear_kernel<<<1, 1, 0, st>>>(rec.count);
    KUPSMI_CHECK_LAUNCH();
  }
  // main pass, then the band pass
#define PC_PASS(P)                                                               \
  if (rec.count) {                                                               \
    if (excl) pc_pair_kernel<true, true, P><<<grid, PC_TILE, 0, st>>>(a);        \
    else pc_pair_kernel<false, true, P><<<grid, PC_TILE, 0, st>>>(a);            \
  } else if (excl) {                                                             \
    pc_pair_kernel<true, false, P><<<grid, PC_TILE, 0, st>>>(a);                 \
  } else {                                                                       \
    pc_pair_kernel<false, false, P><<<grid, PC_TILE, 0, st>>>(a);                \
  }                                                                              \
  KUPSMI_CHECK_LAUNCH();
  PC_PASS(0)
  PC_PASS(1)
#undef PC_PASS
  pc_reduce_kernel<<<(n_sys + 255) / 256, 256, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}
