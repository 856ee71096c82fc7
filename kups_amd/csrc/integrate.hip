// Elementwise integrator kernels of the general (HBM-resident) MD path, the
// A/B/O steps of md/integrators.py:162-402 around a force evaluation done by
// kupsmi_pair_forces + kupsmi_ewald_recip:
//   phase 0: B(dt/2) A(dt)                     (velocity Verlet, 273-317)
//            B(dt/2) A(dt/2) O A(dt/2)         (BAOAB, 405-453)
//            and the fractional wrap of the positions (104-141);
//   phase 1: B(dt/2) and the per-system kinetic energy (md/observables.py:17-34).
// The O step uses the same Philox4x32-10 stream as kupsmi_md_lj.
// kupsmi_thermostat_barostat adds the per-system CSVR rescaling
// (md/integrators.py:470-635) and the stochastic cell rescaling
// (657-811) with the random streams and formulas of kupsmi_md_lj's
// integrators 2 and 3, so the HBM path runs the same NVT / NPT chains.
#include "common.hpp"

namespace kupsmi {

constexpr int IG_THREADS = 256;

__global__ __launch_bounds__(IG_THREADS) void ig_pre_kernel(
    int integrator, int n_sys, const int32_t* sys_off, float* pos, float* mom,
    const float* forces, const float* mass, const double* tril, const double* dt,
    const double* kT, const double* gamma, uint64_t seed, int64_t step, const int64_t* step_dev) {
  // the system's constants once per workgroup (the inverse cell takes four
  // FP64 divisions, the O step an FP64 exp and sqrt)
  __shared__ float s_Lif[6], s_c1, s_c2;
  const int s = blockIdx.y;
  const double* t = tril + 6 * s;
  if (threadIdx.x == 0) {
    double li[6];
    tril_inverse(t, li);
    for (int k = 0; k < 6; ++k) s_Lif[k] = (float)li[k];
    if (integrator == 1) {
      const double e1 = exp(-gamma[s] * dt[s]);
      s_c1 = (float)e1;
      s_c2 = (float)sqrt(kT[s] * (1.0 - e1 * e1));
    }
  }
  __syncthreads();
  const int i0 = sys_off[s], n = sys_off[s + 1] - i0;
  const int i = blockIdx.x * IG_THREADS + threadIdx.x;
  if (i >= n) return;
  const int gi = i0 + i;
  float Lif[6];
  for (int k = 0; k < 6; ++k) Lif[k] = s_Lif[k];
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
    const float c1 = s_c1, c2 = s_c2;
    float e0, ea, eb;
    normals3(seed, (uint32_t)gi, step_dev ? *step_dev : step, e0, ea, eb);
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
// One workgroup of IG_POST_THREADS per system (the kinetic energy is one
// fixed-order FP64 sum per system); wide, so that enough waves are in flight
// for this memory-bound pass.
constexpr int IG_POST_THREADS = 1024;

__global__ __launch_bounds__(IG_POST_THREADS) void ig_post_kernel(
    int n_sys, const int32_t* sys_off, float* mom, const float* forces, const float* mass,
    const double* dt, int kick, double* kin) {
  __shared__ double red[IG_POST_THREADS / 64];
  const int s = blockIdx.x;
  const int i0 = sys_off[s], n = sys_off[s + 1] - i0;
  const float h = kick ? 0.5f * (float)dt[s] : 0.0f;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += IG_POST_THREADS) {
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
    for (int w = 0; w < IG_POST_THREADS / 64; ++w) e += red[w];
    kin[s] = e;
  }
}

// CSVR: K of each system (FP64 sum), the Bussi factor alpha (one R1 for the
// batch, R2 per system) and p <- alpha p.  One workgroup per system.
__global__ __launch_bounds__(IG_THREADS) void ig_csvr_kernel(
    const int32_t* sys_off, float* mom, const float* mass, const double* dt, const double* kT,
    const double* tau, uint64_t seed, int64_t step, const int64_t* step_dev) {
  __shared__ double red[IG_THREADS / 64];
  __shared__ float s_alpha;
  const int s = blockIdx.x;
  const int i0 = sys_off[s], n = sys_off[s + 1] - i0;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += IG_THREADS) {
    const int gi = i0 + i;
    const float px = mom[3 * gi], py = mom[3 * gi + 1], pz = mom[3 * gi + 2];
    acc += 0.5 * (double)((px * px + py * py + pz * pz) / mass[gi]);
  }
  acc = wave_sum(acc);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double K = 0.0;
    for (int w = 0; w < IG_THREADS / 64; ++w) K += red[w];
    s_alpha = (float)csvr_alpha(K, n, kT[s], dt[s], tau[s], seed, (uint32_t)s,
                                step_dev ? *step_dev : step);
  }
  __syncthreads();
  const float alpha = s_alpha;
  for (int i = threadIdx.x; i < n; i += IG_THREADS) {
    const int gi = i0 + i;
    mom[3 * gi] *= alpha; mom[3 * gi + 1] *= alpha; mom[3 * gi + 2] *= alpha;
  }
}

// Stochastic cell rescaling (Bernetti and Bussi 2020): P = 2K / 3V + tr W / 3V,
// d eps = (dt / tau_P) beta (P - P0) + sqrt(2 kT beta dt / (tau_P V)) R with one
// R ~ N(0, 1) for the batch, mu = exp(d eps / 3) clipped to [mu_min, 1 / mu_min];
// the cell (FP64, in place) and the positions are scaled by mu, momenta kept.
__global__ __launch_bounds__(IG_THREADS) void ig_scr_kernel(
    const int32_t* sys_off, float* pos, double* tril, const double* dt, const double* kT,
    const double* baro, const double* virial, const double* kin, uint64_t seed, int64_t step,
    const int64_t* step_dev) {
  __shared__ double s_mu;
  const int s = blockIdx.x;
  const int i0 = sys_off[s], n = sys_off[s + 1] - i0;
  if (threadIdx.x == 0) {
    double* t = tril + 6 * (size_t)s;
    const double* bp = baro + 4 * (size_t)s;
    const double* W = virial + 9 * (size_t)s;
    const double V = fabs(t[0] * t[2] * t[5]);
    const double P = 2.0 * kin[s] / (3.0 * V) + (W[0] + W[4] + W[8]) / (3.0 * V);
    uint32_t c4[4];
    csvr_draw(seed, 0xFFFFFFFEu, step_dev ? *step_dev : step, SCR_TAG, c4);
    const double R = csvr_bm(c4[0], c4[1]);
    const double de = dt[s] / bp[1] * bp[2] * (P - bp[0]) +
                      sqrt(2.0 * kT[s] * bp[2] * dt[s] / (bp[1] * V)) * R;
    const double mu = fmin(fmax(exp(de / 3.0), bp[3]), 1.0 / bp[3]);
    for (int c = 0; c < 6; ++c) t[c] *= mu;
    s_mu = mu;
  }
  __syncthreads();
  const float mu = (float)s_mu;
  for (int i = threadIdx.x; i < n; i += IG_THREADS) {
    const int gi = i0 + i;
    pos[3 * gi] *= mu; pos[3 * gi + 1] *= mu; pos[3 * gi + 2] *= mu;
  }
}

// One thread: *step += n (the device-side step counter of captured steps).
__global__ void ig_step_kernel(int64_t* step, int64_t n) { *step += n; }

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_thermostat_barostat(void* stream, int phase, int n_sys,
                                          const int32_t* sys_off, float* pos, float* mom,
                                          const float* mass, double* tril, const double* dt,
                                          const double* kT, const double* tau,
                                          const double* baro, const double* virial,
                                          const double* kin_energy, uint64_t seed,
                                          int64_t step, const int64_t* step_dev) {
  KUPSMI_REQUIRE(phase == 0 || phase == 1, "phase must be 0 (CSVR) or 1 (cell rescaling)");
  KUPSMI_REQUIRE(n_sys >= 0, "negative size");
  KUPSMI_REQUIRE(sys_off && dt && kT, "system offsets, time steps and kT required");
  KUPSMI_REQUIRE(phase != 0 || (mom && mass && tau), "CSVR needs momenta, masses and tau");
  KUPSMI_REQUIRE(phase != 1 || (pos && tril && baro && virial && kin_energy),
                 "cell rescaling needs positions, cell, barostat, virial and kinetic energy");
  if (n_sys == 0) return KUPSMI_OK;
  hipStream_t st = (hipStream_t)stream;
  if (phase == 0)
    ig_csvr_kernel<<<n_sys, IG_THREADS, 0, st>>>(sys_off, mom, mass, dt, kT, tau, seed, step,
                                                 step_dev);
  else
    ig_scr_kernel<<<n_sys, IG_THREADS, 0, st>>>(sys_off, pos, tril, dt, kT, baro, virial,
                                                kin_energy, seed, step, step_dev);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_step_advance(void* stream, int64_t* step_dev, int64_t n) {
  KUPSMI_REQUIRE(step_dev != nullptr, "null step counter");
  ig_step_kernel<<<1, 1, 0, (hipStream_t)stream>>>(step_dev, n);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

extern "C" int kupsmi_integrate(void* stream, int phase, int integrator, int n_sys,
                                const int32_t* sys_off, int max_atoms_per_sys, float* pos,
                                float* mom, const float* forces, const float* mass,
                                const double* tril, const double* dt, const double* kT,
                                const double* gamma, uint64_t seed, int64_t step,
                                double* kin_energy, const int64_t* step_dev) {
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
                                            tril, dt, kT, gamma, seed, step, step_dev);
  } else {
    // phase 1: kick + kinetic energy; phase 2: kinetic energy only
    ig_post_kernel<<<n_sys, IG_POST_THREADS, 0, st>>>(n_sys, sys_off, mom, forces, mass, dt,
                                                 phase == 1, kin_energy);
  }
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}
