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

namespace kupsmi {

constexpr int MD_THREADS = 1024;
constexpr int MD_NMAX = KUPSMI_LDS_MAX_ATOMS;
constexpr int MD_SLOTS = MD_NMAX / MD_THREADS;
constexpr int MD_MAX_NC_AXIS = 10;  // <= 1000 cells per system
constexpr int MD_NCMAX = 1024;
constexpr int MD_NSPEC = 8;
// relative half width of the FP32 cutoff band re-decided in FP64
constexpr float CUT_BAND = 4e-6f;

struct MdArgs {
  int integrator, n_steps, n_sys, n_species, want_virial;
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
  double* pot;
  double* kin;
  double* virial;
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
__global__ void geom_kernel(int n_sys, const double* tril, const double* cutoff, int max_nc,
                            SysGeom* geom) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n_sys) make_geom(tril + 6 * s, cutoff[s], max_nc, geom[s]);
}

// Exclusive scan of cnt[0..n) into start[0..n] (n <= 1024) by the block.
__device__ inline void block_scan_cells(const int* cnt, int* start, int n, int* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int own = tid < n ? cnt[tid] : 0;
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
  if (tid < n) start[tid] = ex;
  if (tid == n - 1) start[n] = ex + own;
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
  double fi0, fi1, fi2, fj0, fj1, fj2, d0, d1, d2;
  tmul(g.Li, (double)ri.x, (double)ri.y, (double)ri.z, fi0, fi1, fi2);
  tmul(g.Li, (double)rj.x, (double)rj.y, (double)rj.z, fj0, fj1, fj2);
  tmul(g.L, fi0 - fj0 - sx, fi1 - fj1 - sy, fi2 - fj2 - sz, d0, d1, d2);
  return d0 * d0 + d1 * d1 + d2 * d2 < g.rc2;
}

__global__ __launch_bounds__(MD_THREADS) void md_lj_kernel(MdArgs a) {
  __shared__ float4 sp[MD_NMAX];   // sorted x,y,z ; w = bits(species | cell << 8)
  __shared__ float4 pb[MD_NMAX];   // sorted px,py,pz ; w = bits(local id)
  __shared__ short keys[MD_NMAX];  // previous slot of each provisional entry
  __shared__ int cstart[MD_NCMAX + 1];
  __shared__ int ccount[MD_NCMAX];
  __shared__ int wsum[MD_THREADS / 64];
  __shared__ double red[MD_THREADS / 64];
  __shared__ float s_sig2[MD_NSPEC * MD_NSPEC], s_e4[MD_NSPEC * MD_NSPEC],
      s_e24[MD_NSPEC * MD_NSPEC];
  __shared__ SysGeom g;
  __shared__ int s_q;

  const int tid = threadIdx.x;
  const int ns = a.n_species;
  for (int k = tid; k < ns * ns; k += MD_THREADS) {
    double sg = a.sigma[k];
    s_sig2[k] = (float)(sg * sg);
    s_e4[k] = (float)(4.0 * a.epsilon[k]);
    s_e24[k] = (float)(24.0 * a.epsilon[k]);
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
    const int ncx = g.nc[0], ncy = g.nc[1], ncz = g.nc[2], ncell = g.ncell;
    const float dtf = a.n_steps > 0 ? (float)a.dt[s] : 0.0f;
    const float hdt = 0.5f * dtf;
    float c1 = 1.0f, c2 = 0.0f;
    if (a.integrator == 1 && a.n_steps > 0) {
      double e1 = exp(-a.gamma[s] * a.dt[s]);
      c1 = (float)e1;
      c2 = (float)sqrt(a.kT[s] * (1.0 - e1 * e1));
    }

    const int total = a.n_steps > 0 ? a.n_steps : 1;
    for (int st = 0; st < total; ++st) {
      const bool last = st == total - 1;
      // every force loop of the previous step has finished reading sp[]
      __syncthreads();
      // ---- [B(dt/2) at st = 0] A(...) [O A] in place, wrap, cell of every atom
      int cel[MD_SLOTS];
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
          if (st == 0) {
            p.x += a.forces[3 * gi] * hdt;
            p.y += a.forces[3 * gi + 1] * hdt;
            p.z += a.forces[3 * gi + 2] * hdt;
          }
          const float ta = a.integrator == 1 ? hdt : dtf;
          r.x += p.x * im * ta; r.y += p.y * im * ta; r.z += p.z * im * ta;
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
          }
          pb[slot] = p;
        }
        float f0, f1, f2;
        wrap_shift(g.L, g.Lif, r.x, r.y, r.z, f0, f1, f2);
        cel[k] = (cell_index(f2, ncz) * ncy + cell_index(f1, ncy)) * ncx + cell_index(f0, ncx);
        r.w = __int_as_float((__float_as_int(r.w) & 0xff) | (cel[k] << 8));
        sp[slot] = r;
      }
      // ---- stable counting sort into cells (rank by previous slot)
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
      {
        int dst[MD_SLOTS];
        float4 tr[MD_SLOTS], tp[MD_SLOTS];
#pragma unroll
        for (int k = 0; k < MD_SLOTS; ++k) {
          const int slot = tid + k * MD_THREADS;
          dst[k] = 0;
          if (slot < n) {
            const int b = cstart[cel[k]], e = cstart[cel[k] + 1];
            int rank = 0;
            for (int t = b; t < e; ++t) rank += keys[t] < slot;
            dst[k] = b + rank;
            tr[k] = sp[slot];
            tp[k] = pb[slot];
          }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MD_SLOTS; ++k) {
          if (tid + k * MD_THREADS < n) {
            sp[dst[k]] = tr[k];
            pb[dst[k]] = tp[k];
          }
        }
      }
      __syncthreads();
      // ---- forces: every slot walks its stencil of (cell, image) pairs,
      // then the kick and the kinetic energy of that slot
      float epot = 0.0f, ekin = 0.0f;
      float w00 = 0.f, w01 = 0.f, w02 = 0.f, w11 = 0.f, w12 = 0.f, w22 = 0.f;
      const bool do_w = last && a.want_virial;
      const int h0 = g.h[0], h1 = g.h[1], h2 = g.h[2];
#pragma unroll 1
      for (int k = 0; k < MD_SLOTS; ++k) {
        const int slot = tid + k * MD_THREADS;
        if (slot >= n) continue;
        const float4 ri = sp[slot];
        const int wi = __float_as_int(ri.w);
        const int c = wi >> 8;
        const int si = wi & 0xff;
        const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
        float ax = 0.f, ay = 0.f, az = 0.f, ei = 0.f;
        for (int dz = -h2; dz <= h2; ++dz) {
          int zz = cz + dz;
          const int sz = zz >= 0 ? zz / ncz : -((-zz + ncz - 1) / ncz);
          zz -= sz * ncz;
          for (int dy = -h1; dy <= h1; ++dy) {
            int yy = cy + dy;
            const int sy = yy >= 0 ? yy / ncy : -((-yy + ncy - 1) / ncy);
            yy -= sy * ncy;
            for (int dx = -h0; dx <= h0; ++dx) {
              int xx = cx + dx;
              const int sx = xx >= 0 ? xx / ncx : -((-xx + ncx - 1) / ncx);
              xx -= sx * ncx;
              // image shift s @ L; edge vector d = r_j + s@L - r_i
              float Sx, Sy, Sz;
              tmul(g.Lf, (float)sx, (float)sy, (float)sz, Sx, Sy, Sz);
              const float qx = ri.x - Sx, qy = ri.y - Sy, qz = ri.z - Sz;
              const bool center = dx == 0 && dy == 0 && dz == 0;
              const int nb = (zz * ncy + yy) * ncx + xx;
              const int tb = cstart[nb], te = cstart[nb + 1];
              for (int t = tb; t < te; ++t) {
                const float4 rj = sp[t];
                const float ex = rj.x - qx, ey = rj.y - qy, ez = rj.z - qz;
                const float r2 = ex * ex + ey * ey + ez * ez;
                if (!(r2 < rc2_hi) || (center && t == slot)) continue;
                if (r2 > rc2_lo && !band_keep(g, ri, rj, sx, sy, sz)) continue;
                const int p = si * ns + (__float_as_int(rj.w) & 0xff);
                const float ir2 = 1.0f / r2;
                float c6 = s_sig2[p] * ir2;
                c6 = c6 * c6 * c6;
                ei += s_e4[p] * c6 * (c6 - 1.0f);
                const float g2 = s_e24[p] * c6 * (1.0f - 2.0f * c6) * ir2;  // 2 u'
                ax += g2 * ex; ay += g2 * ey; az += g2 * ez;
                if (do_w) {
                  const float hg = -0.5f * g2;
                  w00 += hg * ex * ex; w01 += hg * ex * ey; w02 += hg * ex * ez;
                  w11 += hg * ey * ey; w12 += hg * ey * ez; w22 += hg * ez * ez;
                }
              }
            }
          }
        }
        epot += 0.5f * ei;
        // closing B(dt/2) (+ the next step's opening B(dt/2)), kinetic energy
        float4 p = pb[slot];
        const int gi = i0 + __float_as_int(p.w);
        if (a.n_steps > 0 || a.kin) {
          const float im = 1.0f / a.mass[gi];
          const float vx = p.x + ax * hdt, vy = p.y + ay * hdt, vz = p.z + az * hdt;
          ekin += 0.5f * (vx * vx + vy * vy + vz * vz) * im;
        }
        if (a.n_steps > 0) {
          const float kick = last ? hdt : dtf;
          p.x += ax * kick; p.y += ay * kick; p.z += az * kick;
          pb[slot] = p;
        }
        if (last) {
          a.forces[3 * gi] = ax; a.forces[3 * gi + 1] = ay; a.forces[3 * gi + 2] = az;
        }
      }
      if (a.pot) {
        double e = block_sum((double)epot, red);
        if (tid == 0) a.pot[(size_t)s * total + st] = e;
      }
      if (a.kin) {
        double e = block_sum((double)ekin, red);
        if (tid == 0) a.kin[(size_t)s * total + st] = e;
      }
      if (do_w && a.virial) {
        const float wv[6] = {w00, w01, w02, w11, w12, w22};
        double out[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) out[c] = block_sum((double)wv[c], red);
        if (tid == 0) {
          double* W = a.virial + 9 * (size_t)s;
          W[0] = out[0]; W[1] = out[1]; W[2] = out[2];
          W[3] = out[1]; W[4] = out[3]; W[5] = out[4];
          W[6] = out[2]; W[7] = out[4]; W[8] = out[5];
        }
      }
    }
    // ---- store the system back in its original atom order (forces were
    // written by the last force pass)
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
  KUPSMI_REQUIRE(a.integrator == 0 || a.integrator == 1, "integrator must be 0 (VV) or 1 (BAOAB)");
  KUPSMI_REQUIRE(a.n_steps >= 0, "n_steps < 0");
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
  geom_kernel<<<(a.n_sys + 255) / 256, 256, 0, stream>>>(a.n_sys, a.tril, a.cutoff, MD_MAX_NC_AXIS, geom);
  KUPSMI_CHECK_LAUNCH();
  a.order = order;
  a.queue = queue;
  a.geom = geom;
  int dev = 0, ncu = 256;
  KUPSMI_HIP(hipGetDevice(&dev));
  KUPSMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int grid = a.n_sys < ncu ? a.n_sys : ncu;
  md_lj_kernel<<<grid, MD_THREADS, 0, stream>>>(a);
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
                            const double* cutoff, const double* dt, const double* kT,
                            const double* gamma, double* pot_energy, double* kin_energy,
                            double* virial) {
  KUPSMI_REQUIRE(n_atoms >= 0, "n_atoms < 0");
  KUPSMI_REQUIRE(n_steps == 0 || (dt != nullptr), "dt required");
  KUPSMI_REQUIRE(integrator != 1 || n_steps == 0 || (kT && gamma), "kT and gamma required for BAOAB");
  MdArgs a{};
  a.integrator = integrator; a.n_steps = n_steps; a.n_sys = n_sys; a.n_species = n_species;
  a.want_virial = virial != nullptr;
  a.step0 = step0; a.seed = seed;
  a.sys_off = sys_off; a.pos = pos; a.mom = mom; a.forces = forces; a.mass = mass;
  a.species = species; a.sigma = sigma; a.epsilon = epsilon; a.tril = tril; a.cutoff = cutoff;
  a.dt = dt; a.kT = kT; a.gamma = gamma; a.pot = pot_energy; a.kin = kin_energy; a.virial = virial;
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
