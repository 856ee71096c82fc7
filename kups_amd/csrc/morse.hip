// Morse bond energy, forces and virial over an explicit bond list
// (potential/classical/morse.py:142-168): for bond (i, j, s) with
// d = r_j - r_i + s @ L, E = D [1 - e^{-alpha (r - r0)}]^2, parameters by the
// species pair of (i, j), every bond counted once.  One lane per bond; the
// bond vector and the energy in FP64 (bonds are few), forces added to FP32
// with hardware atomics, per-system FP64 energy and virial with FP64 atomics
// (bond terms of one system are summed in arrival order).  Bonds whose
// indices fall outside [0, n_atoms) are skipped; the host wrapper rejects them.
#include "common.hpp"

namespace kupsmi {

constexpr int MO_THREADS = 256;
constexpr int MO_NSPEC = 16;

struct MorseArgs {
  int n_sys, n_atoms, n_species;
  long long n_bonds;
  const int32_t* sys_off;
  const float* pos;
  const int32_t* species;
  const double* D;
  const double* alpha;
  const double* r0;
  const int32_t* bonds;   // n_bonds x 2
  const int32_t* shifts;  // n_bonds x 3 (NULL: no image shifts)
  const double* tril;
  float* forces;
  double* energy;
  double* virial;
};

__device__ inline int mo_system(const int32_t* off, int n_sys, int i) {
  int lo = 0, hi = n_sys - 1;  // last s with off[s] <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(MO_THREADS) void morse_kernel(MorseArgs a) {
  const long long b = (long long)blockIdx.x * MO_THREADS + threadIdx.x;
  if (b >= a.n_bonds) return;
  const int i = a.bonds[2 * b], j = a.bonds[2 * b + 1];
  if (i < 0 || j < 0 || i >= a.n_atoms || j >= a.n_atoms) return;
  const int s = mo_system(a.sys_off, a.n_sys, i);
  const double* L = a.tril + 6 * s;
  double sx = 0.0, sy = 0.0, sz = 0.0;
  if (a.shifts)
    tmul(L, (double)a.shifts[3 * b], (double)a.shifts[3 * b + 1], (double)a.shifts[3 * b + 2],
         sx, sy, sz);
  const double dx = (double)a.pos[3 * j] - (double)a.pos[3 * i] + sx;
  const double dy = (double)a.pos[3 * j + 1] - (double)a.pos[3 * i + 1] + sy;
  const double dz = (double)a.pos[3 * j + 2] - (double)a.pos[3 * i + 2] + sz;
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  const int p = a.species[i] * a.n_species + a.species[j];
  const double al = a.alpha[p], Dp = a.D[p];
  const double ex = exp(-al * (r - a.r0[p]));
  const double u = Dp * (1.0 - ex) * (1.0 - ex);
  const double g = 2.0 * Dp * (1.0 - ex) * al * ex / r;  // dE/dr / r
  // dE/dr_i = -g d, dE/dr_j = +g d; forces are -dE/dr
  atomicAdd(a.forces + 3 * i, (float)(g * dx));
  atomicAdd(a.forces + 3 * i + 1, (float)(g * dy));
  atomicAdd(a.forces + 3 * i + 2, (float)(g * dz));
  atomicAdd(a.forces + 3 * j, (float)(-g * dx));
  atomicAdd(a.forces + 3 * j + 1, (float)(-g * dy));
  atomicAdd(a.forces + 3 * j + 2, (float)(-g * dz));
  atomicAdd(a.energy + s, u);
  if (a.virial) {
    const double d[3] = {dx, dy, dz};
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) atomicAdd(a.virial + 9 * s + 3 * k + l, -g * d[k] * d[l]);
  }
}

__global__ void morse_clear_kernel(int n_atoms, int n_sys, float* forces, double* energy,
                                   double* virial) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * n_atoms) forces[t] = 0.f;
  if (t < n_sys) energy[t] = 0.0;
  if (virial && t < 9 * n_sys) virial[t] = 0.0;
}

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_morse_bonds(void* stream, int n_sys, const int32_t* sys_off, int n_atoms,
                                  const float* pos, const int32_t* species, int n_species,
                                  const double* D, const double* alpha, const double* r0,
                                  long long n_bonds, const int32_t* bonds, const int32_t* shifts,
                                  const double* tril, int accumulate, float* forces,
                                  double* energy, double* virial) {
  KUPSMI_REQUIRE(n_sys >= 1 && n_atoms >= 0 && n_bonds >= 0, "bad sizes");
  KUPSMI_REQUIRE(n_species >= 1 && n_species <= MO_NSPEC, "n_species must be in [1, 16]");
  KUPSMI_REQUIRE(forces && energy, "forces and energy are required");
  KUPSMI_REQUIRE(n_bonds == 0 || (bonds && pos && species && tril && D && alpha && r0 && sys_off),
                 "null input");
  hipStream_t st = (hipStream_t)stream;
  if (!accumulate) {
    const int nclear = 3 * n_atoms > 9 * n_sys ? 3 * n_atoms : 9 * n_sys;
    morse_clear_kernel<<<(nclear + 255) / 256, 256, 0, st>>>(n_atoms, n_sys, forces, energy,
                                                            virial);
    KUPSMI_CHECK_LAUNCH();
  }
  if (n_bonds == 0) return KUPSMI_OK;
  MorseArgs a{};
  a.n_sys = n_sys; a.n_atoms = n_atoms; a.n_species = n_species; a.n_bonds = n_bonds;
  a.sys_off = sys_off; a.pos = pos; a.species = species; a.D = D; a.alpha = alpha; a.r0 = r0;
  a.bonds = bonds; a.shifts = shifts; a.tril = tril; a.forces = forces; a.energy = energy;
  a.virial = virial;
  const long long grid = (n_bonds + MO_THREADS - 1) / MO_THREADS;
  KUPSMI_REQUIRE(grid < (1ll << 31), "too many bonds");
  morse_kernel<<<(unsigned)grid, MO_THREADS, 0, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}
