// Elementwise integrator kernels of the general (HBM-resident) MD path, the
// A/B/O steps of md/integrators.py:162-402 around a force evaluation done by
// kupsmi_pair_forces + kupsmi_ewald_recip:
//   phase 0: B(dt/2) A(dt)                     (velocity Verlet, 273-317)
//            B(dt/2) A(dt/2) O A(dt/2)         (BAOAB, 405-453)
//            and the fractional wrap of the positions (104-141);
//   phase 1: B(dt/2) and the per-system kinetic energy (md/observables.py:17-34).
// The O step uses the same Philox4x32-10 stream as kupsmi_md_lj.
#include "common.hpp"

namespace kupsmi {

constexpr int IG_THREADS = 256;

__global__ __launch_bounds__(IG_THREADS) void ig_pre_kernel(
    int integrator, int n_sys, const int32_t* sys_off, float* pos, float* mom,
    const float* forces, const float* mass, const double* tril, const double* dt,
    const double* kT, const double* gamma, uint64_t seed, int64_t step) {
  const int s = blockIdx.y;
  const int i0 = sys_off[s], n = sys_off[s + 1] - i0;
  const int i = blockIdx.x * IG_THREADS + threadIdx.x;
  if (i >= n) return;
  const int gi = i0 + i;
  const double* t = tril + 6 * s;
  double li[6];
  tril_inverse(t, li);
  float Lf[6], Lif[6];
  for (int k = 0; k < 6; ++k) { Lf[k] = (float)t[k]; Lif[k] = (float)li[k]; }
  const float d = (float)dt[s], h = 0.5f * d;
  const float im = 1.0f / mass[gi];
  float px = mom[3 * gi] + forces[3 * gi] * h;
  float py = mom[3 * gi + 1] + forces[3 * gi + 1] * h;
  float pz = mom[3 * gi + 2] + forces[3 * gi + 2] * h;
  float rx = pos[3 * gi], ry = pos[3 * gi + 1], rz = pos[3 * gi + 2];
  const float ta = integrator == 1 ? h : d;
  rx += px * im * ta; ry += py * im * ta; rz += pz * im * ta;
  float f0, f1, f2;
  if (integrator == 1) {
    wrap_shift(t, Lif, rx, ry, rz, f0, f1, f2);
    const double e1 = exp(-gamma[s] * dt[s]);
    const float c1 = (float)e1, c2 = (float)sqrt(kT[s] * (1.0 - e1 * e1));
    float e0, ea, eb;
    normals3(seed, (uint32_t)gi, step, e0, ea, eb);
    const float sm = c2 * sqrtf(mass[gi]);
    px = c1 * px + sm * e0; py = c1 * py + sm * ea; pz = c1 * pz + sm * eb;
    rx += px * im * h; ry += py * im * h; rz += pz * im * h;
  }
  wrap_shift(t, Lif, rx, ry, rz, f0, f1, f2);
  pos[3 * gi] = rx; pos[3 * gi + 1] = ry; pos[3 * gi + 2] = rz;
  mom[3 * gi] = px; mom[3 * gi + 1] = py; mom[3 * gi + 2] = pz;
}

// Closing B(dt/2) and the kinetic energy, one workgroup per system (fixed
// reduction order).
__global__ __launch_bounds__(IG_THREADS) void ig_post_kernel(
    int n_sys, const int32_t* sys_off, float* mom, const float* forces, const float* mass,
    const double* dt, int kick, double* kin) {
  __shared__ double red[IG_THREADS / 64];
  const int s = blockIdx.x;
  const int i0 = sys_off[s], n = sys_off[s + 1] - i0;
  const float h = kick ? 0.5f * (float)dt[s] : 0.0f;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += IG_THREADS) {
    const int gi = i0 + i;
    float px = mom[3 * gi] + forces[3 * gi] * h;
    float py = mom[3 * gi + 1] + forces[3 * gi + 1] * h;
    float pz = mom[3 * gi + 2] + forces[3 * gi + 2] * h;
    if (kick) { mom[3 * gi] = px; mom[3 * gi + 1] = py; mom[3 * gi + 2] = pz; }
    acc += 0.5 * (double)((px * px + py * py + pz * pz) / mass[gi]);
  }
  acc = wave_sum(acc);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0 && kin) {
    double e = 0.0;
    for (int w = 0; w < IG_THREADS / 64; ++w) e += red[w];
    kin[s] = e;
  }
}

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_integrate(void* stream, int phase, int integrator, int n_sys,
                                const int32_t* sys_off, int max_atoms_per_sys, float* pos,
                                float* mom, const float* forces, const float* mass,
                                const double* tril, const double* dt, const double* kT,
                                const double* gamma, uint64_t seed, int64_t step,
                                double* kin_energy) {
  KUPSMI_REQUIRE(phase == 0 || phase == 1 || phase == 2, "phase must be 0, 1 or 2");
  KUPSMI_REQUIRE(integrator == 0 || integrator == 1, "integrator must be 0 (VV) or 1 (BAOAB)");
  KUPSMI_REQUIRE(n_sys >= 0 && max_atoms_per_sys >= 0, "negative size");
  KUPSMI_REQUIRE(integrator == 0 || phase != 0 || (kT && gamma), "BAOAB needs kT and gamma");
  if (n_sys == 0) return KUPSMI_OK;
  hipStream_t st = (hipStream_t)stream;
  if (phase == 0) {
    if (max_atoms_per_sys == 0) return KUPSMI_OK;
    dim3 g((max_atoms_per_sys + IG_THREADS - 1) / IG_THREADS, n_sys);
    ig_pre_kernel<<<g, IG_THREADS, 0, st>>>(integrator, n_sys, sys_off, pos, mom, forces, mass,
                                            tril, dt, kT, gamma, seed, step);
  } else {
    // phase 1: kick + kinetic energy; phase 2: kinetic energy only
    ig_post_kernel<<<n_sys, IG_THREADS, 0, st>>>(n_sys, sys_off, mom, forces, mass, dt,
                                                 phase == 1, kin_energy);
  }
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}
