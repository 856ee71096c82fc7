// Rigid-molecule GCMC of a batch of chains, every chain resident on one CU.
//
// One 512-thread workgroup owns one chain at a time: the host framework and
// the guest molecules live in LDS, the chain's structure factors S(k) live in
// registers (k = tid + q * 512, KPT per thread), and the scheduler state
// (step widths, acceptance histories, counters, species counts) lives in LDS
// for the whole launch, so a trial touches HBM only for a learned-proposal
// lookup.  Per trial:
//   1. wave 0: the four Philox draws on lanes 0..3 and the three FP64 logs of
//      the acceptance (ln u, ln(N+1), ln N) on lanes 1..3, in one pass; lane 0
//      then forms the proposal (translation, rotation, reinsertion,
//      insertion, deletion) from LDS data only;
//   2. all lanes: FP64 axis phase tables e^{2 pi i n f} of the changed sites;
//      LJ + Ewald real-space change, one lane per partner (host atoms and the
//      other guests), FP32 pair terms;
//   3. reciprocal-space change K sum_k P(k) (|S + dS|^2 - |S|^2) over the
//      register-resident k list, FP64 (the host's |S| is large, so FP32 site
//      phases would leave ~1e-4 eV per move in the cached S and the tracked
//      energy);
//   4. one combined FP64 block reduction; lane 0 adds the molecule constant
//      and the analytic LJ tail change (O(1): 2 dN (T m).c + m.T m), forms the
//      FP64 log acceptance ratio and runs the Metropolis test and the step-
//      width schedule;
//   5. on accept every lane updates its S(k) registers, wave 0 the guest table.
//
// Reference: application/simulations/mcmc_rigid.py:281-307; mcmc/moves.py
// (122-295 displacement moves, 573-835 insertion/deletion, 1078-1230 mixing);
// mcmc/probability.py:60-225 (Boltzmann and fugacity log ratios);
// core/propagator.py:244-344 (one move type for all systems, accept per
// system); core/parameter_scheduler.py:161-208; potential/classical/
// ewald.py:442-486 (S' = S + dS), lennard_jones.py:276-313 (tail).
// Random numbers: Philox4x32-10 counters (chain, step, tag), the stream the
// FP64 oracle draws too.
#include "common.hpp"

namespace kupsmi {

// per-chain statistics: trials[4], accepted[4], insertions, deletions, trials
// that evaluated sites, site evaluations
constexpr int GC_NSTATS = 12;

constexpr int GC_THREADS = 512;
constexpr int GC_WAVES = GC_THREADS / 64;
constexpr int GC_MSITE = 4;      // sites per molecule
constexpr int GC_NM = 12;        // max |n| per axis of the k list
constexpr int GC_NSPEC = 16;
constexpr int GC_NMOTIF = 4;     // guest species (motifs) of a mixture
constexpr int GC_KPT_MAX = 8;    // k vectors per thread (nk <= 4096)
constexpr uint32_t GC_TAG_BATCH = 0x42415443u, GC_TAG_A = 0x47434d41u,
                   GC_TAG_B = 0x47434d42u, GC_TAG_C = 0x47434d43u;

struct GcArgs {
  int n_chain, n_steps, gmax, m, n_motif, nspec, nk, nm, n_trace, max_host;
  int mnsite[GC_NMOTIF];     // sites of each motif (<= m)
  int64_t step0;
  const int64_t* step_dev;  // when set, the first trial's counter is read here (graph replays)
  uint64_t seed;
  double rc_lj, rc_cl;
  double w_move[4], upper[4];
  const int32_t* host_off;
  const float4* host;        // x, y, z, q
  const int32_t* host_spec;
  const double* tril;
  const float* motif_pos;    // n_motif x m x 3, centred
  const float* motif_q;      // n_motif x m
  const int32_t* motif_spec; // n_motif x m
  const double* sigma;       // nspec x nspec
  const double* epsilon;
  const double* alpha;       // per chain
  const double* tail;        // n_chain x nspec x nspec
  const double* mol_const;   // n_chain x n_motif: self + exclusion energy of one molecule
  const double* kT;
  const double* log_activity;  // n_chain x n_motif
  const int32_t* kn;         // n_chain x nk packed (n0, n1 + 64, n2 + 64)
  const double* pk;          // n_chain x nk, P(k) (0 for padding)
  // state (in/out)
  double* guest;             // n_chain x gmax x m x 3 (FP64: rigid bodies stay rigid)
  int32_t* n_guest;
  int32_t* guest_motif;      // n_chain x gmax: motif of each guest slot
  int32_t* counts;           // n_chain x nspec
  double* sk;                // n_chain x nk x 2
  double* energy;
  float* width;              // n_chain x 4
  uint32_t* hist;            // n_chain x 4 x 4
  int32_t* hist_idx;         // n_chain x 4
  long long* stats;          // n_chain x 10: trials[4], accepted[4], ins, del
  double* trace;             // n_trace x n_steps x 4
  // learned insertion proposal (NULL: uniform) and insertion-sample recording
  const float* prop_cdf;     // n_chain x nv^3
  const int32_t* prop_guide; // n_chain x nv^3: search start for u in [j, j+1) / nv^3
  int prop_nv;
  int* sample_count;
  int sample_cap;
  int32_t* sample_row;       // chain * nv^3 + voxel of the insertion offset
  float* sample_y;           // -dE / kT of the insertion trial
};

__device__ inline double2 gc_zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline void gc_draw(uint64_t seed, uint32_t a, int64_t t, uint32_t tag, uint32_t o[4]) {
  o[0] = a; o[1] = (uint32_t)t; o[2] = (uint32_t)((uint64_t)t >> 32); o[3] = tag;
  philox4x32_10(o, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__device__ inline uint32_t gc_readlane(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ inline double gc_readlane(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const uint32_t lo = gc_readlane((uint32_t)b, l), hi = gc_readlane((uint32_t)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Shoemake quaternion, power w, rotation matrix (core/utils/quaternion.py).
// The quaternion comes from FP32 transcendentals of the FP32 uniforms; the
// matrix is formed in FP64 from the quaternion normalised in FP64, so it is
// orthogonal to FP64 rounding and repeated rotations do not deform a molecule.
__device__ void gc_rotation(const float u[3], float w, double R[3][3]) {
  float s2, c2, s3, c3;
  sincospif(2.0f * u[1], &s2, &c2);
  sincospif(2.0f * u[2], &s3, &c3);
  float q0 = sqrtf(1.0f - u[0]) * s2, q1 = sqrtf(1.0f - u[0]) * c2;
  float q2 = sqrtf(u[0]) * s3, q3 = sqrtf(u[0]) * c3;
  if (q0 < 0.f) { q0 = -q0; q1 = -q1; q2 = -q2; q3 = -q3; }
  const float ang = 2.0f * acosf(fminf(q0, 1.0f));
  const float sn = sqrtf(q1 * q1 + q2 * q2 + q3 * q3);
  float ax = 1.f, ay = 0.f, az = 0.f;
  if (sn > 1e-10f) { ax = q1 / sn; ay = q2 / sn; az = q3 / sn; }
  float sh, ch;
  sincosf(0.5f * ang * w, &sh, &ch);
  double W = ch, X = (double)ax * sh, Y = (double)ay * sh, Z = (double)az * sh;
  const double inv = 1.0 / sqrt(W * W + X * X + Y * Y + Z * Z);
  W *= inv; X *= inv; Y *= inv; Z *= inv;
  const double x2 = X * X, y2 = Y * Y, z2 = Z * Z, w2 = W * W;
  R[0][0] = x2 - y2 - z2 + w2; R[0][1] = 2 * (X * Y - Z * W); R[0][2] = 2 * (X * Z + Y * W);
  R[1][0] = 2 * (X * Y + Z * W); R[1][1] = -x2 + y2 - z2 + w2; R[1][2] = 2 * (Y * Z - X * W);
  R[2][0] = 2 * (X * Z - Y * W); R[2][1] = 2 * (Y * Z + X * W); R[2][2] = -x2 - y2 + z2 + w2;
}

struct GcCell {
  double L[6], Li[6];
  float Lf[6], Lif[6];
};

// Wrap into the primary cell by whole lattice vectors, n = floor(f + 1/2) (FP64).
__device__ inline void gc_wrap(const GcCell& c, double* r) {
  double f0, f1, f2;
  tmul(c.Li, r[0], r[1], r[2], f0, f1, f2);
  const double n0 = floor(f0 + 0.5), n1 = floor(f1 + 0.5), n2 = floor(f2 + 0.5);
  double s0, s1, s2;
  tmul(c.L, n0, n1, n2, s0, s1, s2);
  r[0] -= s0; r[1] -= s1; r[2] -= s2;
}

// minimum-image vector b - a (FP32 for the pair loop, FP64 for proposals)
__device__ inline void gc_minimg(const GcCell& c, const float* a, const float* b, float* d) {
  float f0, f1, f2;
  tmul(c.Lif, b[0] - a[0], b[1] - a[1], b[2] - a[2], f0, f1, f2);
  f0 -= rintf(f0); f1 -= rintf(f1); f2 -= rintf(f2);
  tmul(c.Lf, f0, f1, f2, d[0], d[1], d[2]);
}
__device__ inline void gc_minimg(const GcCell& c, const double* a, const double* b, double* d) {
  double f0, f1, f2;
  tmul(c.Li, b[0] - a[0], b[1] - a[1], b[2] - a[2], f0, f1, f2);
  f0 -= rint(f0); f1 -= rint(f1); f2 -= rint(f2);
  tmul(c.L, f0, f1, f2, d[0], d[1], d[2]);
}

// Voxel of a fractional insertion offset u in [0, 1)^3.
__device__ inline int gc_voxel(float u0, float u1, float u2, int nv) {
  int a = (int)(u0 * nv), b = (int)(u1 * nv), c = (int)(u2 * nv);
  a = a < 0 ? 0 : (a >= nv ? nv - 1 : a);
  b = b < 0 ? 0 : (b >= nv ? nv - 1 : b);
  c = c < 0 ? 0 : (c >= nv ? nv - 1 : c);
  return (a * nv + b) * nv + c;
}

// q(v) nvox of the learned table (the exact probability the sampler uses).
__device__ inline double gc_qn(const float* cdf, int v, int nvox) {
  const double lo = v > 0 ? (double)cdf[v - 1] : 0.0;
  return ((double)cdf[v] - lo) * nvox;
}

// First v with u < cdf[v] (the last voxel if none): start at the guide entry
// of u's bin and walk; cdf is non-decreasing, so the walk ends at the same v
// as a binary search, after ~1-2 loads.
__device__ inline int gc_sample_cdf(const float* cdf, const int32_t* guide, int nvox, float u) {
  int j = (int)(u * (float)nvox);
  j = j < 0 ? 0 : (j >= nvox ? nvox - 1 : j);
  int v = guide[j];
  v = v < 0 ? 0 : (v >= nvox ? nvox - 1 : v);
  while (v > 0 && u < cdf[v - 1]) --v;
  while (v < nvox - 1 && u >= cdf[v]) ++v;
  return v;
}

// Scheduler and bookkeeping state of the chain being run (LDS).
struct GcChain {
  double kT;
  // per motif t: ln(a_t V), molecule constant, 2 (T m_t).c without the 2, and
  // m_u T m_t (T the symmetrised tail matrix, m_t the species counts of one
  // molecule of t, c the chain's species counts)
  double lnfV[GC_NMOTIF], molc[GC_NMOTIF], tcm[GC_NMOTIF], mtm[GC_NMOTIF][GC_NMOTIF];
  int nmol[GC_NMOTIF];
  float wid[4];
  uint32_t hist[16];
  int hist_idx[4];
  int cnt[GC_NSPEC];
  long long stats[GC_NSTATS];
};

// The trial formed by lane 0 and read by the workgroup.
struct GcTrial {
  int move, deleting, n_old, n_new, slot, voxel, accept, motif;
  double log_ratio;  // proposal log ratio ln(q(x|x') / q(x'|x)) of the move
};

// Proposal of one trial (lane 0 of wave 0): move type, sites before/after
// the move, written to LDS as FP32 (x, y, z, q) for the pair loop and FP64 for
// the phase tables and the guest table.  bb/ra/rb/rc are the batch, A, B and
// C draws; g64 is the FP64 guest table, gmot the motif of each guest slot.
// Translation, rotation and reinsertion pick a guest of any motif; an
// insertion or a deletion first picks the motif uniformly (per chain, from
// ra[2]; mcmc/moves.py:600-602, 711-712), a deletion then the ra[0] % N_t-th
// guest of that motif (no-op when there is none).
__device__ inline void gc_propose(const GcArgs& a, const GcCell& cell, const GcChain& cs, int ch,
                                  int ng, const double* g64, const uint8_t* gmot,
                                  const float4 (*motifs)[GC_MSITE], const int* mnsite,
                                  const uint32_t bb[4],
                                  const uint32_t ra[4], const uint32_t rb[4], const uint32_t rc[4],
                                  float4* site_old, float4* site_new, double (*site64)[3],
                                  GcTrial& tr) {
  const int M = a.m;
  const double wsum = a.w_move[0] + a.w_move[1] + a.w_move[2] + a.w_move[3];
  const double um = (double)u01(bb[0]) * wsum;
  double acc = 0;
  int move = 3;
  for (int k = 0; k < 4; ++k) {
    acc += a.w_move[k];
    if (um < acc) { move = k; break; }
  }
  const int deleting = (bb[1] >> 31) & 1;
  int n_old = 0, n_new = 0, slot = -1, voxel = -1;
  int t = move == 3 ? (int)(ra[2] % (uint32_t)a.n_motif) : 0;
  double log_ratio = 0.0;
  const double* old = g64;
  if (move < 3 && ng > 0) {
    slot = (int)(ra[0] % (uint32_t)ng);
    t = gmot[slot];
  } else if (move == 3 && deleting && cs.nmol[t] > 0) {
    int k = (int)(ra[0] % (uint32_t)cs.nmol[t]);
    for (int g = 0; g < ng; ++g)
      if (gmot[g] == t && k-- == 0) { slot = g; break; }
  }
  const float4* motif = motifs[t];
  const int Mt = mnsite[t];
  if (slot >= 0) {
    n_old = Mt;
    old = g64 + slot * M * 3;
    for (int s = 0; s < Mt; ++s) {
      const double* p = old + 3 * s;
      site_old[s] = make_float4((float)p[0], (float)p[1], (float)p[2], motif[s].w);
      site64[s][0] = p[0]; site64[s][1] = p[1]; site64[s][2] = p[2];
    }
  }
  double (*snew)[3] = site64 + n_old;
  // geometric centre of the old molecule (minimum image relative to its last site)
  double com[3] = {0.0, 0.0, 0.0};
  if (n_old && (move == 1 || move == 2 || (move == 3 && a.prop_cdf))) {
    const double* last = old + 3 * (Mt - 1);
    double cen[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < Mt; ++s) {
      double rel[3];
      gc_minimg(cell, last, old + 3 * s, rel);
      for (int x = 0; x < 3; ++x) cen[x] += rel[x] / Mt;
    }
    for (int x = 0; x < 3; ++x) com[x] = last[x] + cen[x];
  }
  if (move == 3 && deleting && n_old && a.prop_cdf) {
    // reverse move = learned insertion at the molecule's centre
    double f0, f1, f2;
    tmul(cell.Li, com[0], com[1], com[2], f0, f1, f2);
    f0 -= floor(f0); f1 -= floor(f1); f2 -= floor(f2);
    const int nv = a.prop_nv, nvox = nv * nv * nv;
    const int v = gc_voxel((float)f0, (float)f1, (float)f2, nv);
    log_ratio = log(gc_qn(a.prop_cdf + (size_t)ch * nvox, v, nvox));
  }
  if (move == 0 && n_old) {
    const float u0 = u01(rc[0]), u1 = u01(rc[1]), u2 = u01(rc[2]), u3 = u01(rc[3]);
    const float r0 = sqrtf(-2.0f * logf(u0)), r1 = sqrtf(-2.0f * logf(u2));
    float s1, c1, s3, c3;
    sincospif(2.0f * u1, &s1, &c1);
    sincospif(2.0f * u3, &s3, &c3);
    const double w = cs.wid[0];
    const double g3[3] = {(double)(r0 * c1) * w, (double)(r0 * s1) * w, (double)(r1 * c3) * w};
    for (int s = 0; s < Mt; ++s)
      for (int x = 0; x < 3; ++x) snew[s][x] = old[3 * s + x] + g3[x];
    n_new = Mt;
  } else if ((move == 1 || move == 2) && n_old) {
    const float uq[3] = {u01(rb[0]), u01(rb[1]), u01(rb[2])};
    double R[3][3];
    gc_rotation(uq, move == 1 ? cs.wid[1] : 1.0f, R);
    gc_wrap(cell, com);
    double off[3] = {0.0, 0.0, 0.0};
    if (move == 2)
      tmul(cell.L, (double)u01(rc[0]), (double)u01(rc[1]), (double)u01(rc[2]), off[0], off[1], off[2]);
    for (int s = 0; s < Mt; ++s) {
      double rr[3];
      gc_minimg(cell, com, old + 3 * s, rr);
      for (int x = 0; x < 3; ++x)
        snew[s][x] = com[x] + rr[0] * R[0][x] + rr[1] * R[1][x] + rr[2] * R[2][x] + off[x];
    }
    n_new = Mt;
  } else if (move == 3 && !deleting) {
    const float uq[3] = {u01(rb[0]), u01(rb[1]), u01(rb[2])};
    double R[3][3];
    gc_rotation(uq, 1.0f, R);
    double fu[3] = {u01(rc[0]), u01(rc[1]), u01(rc[2])};
    if (a.prop_cdf) {
      // voxel ~ q, uniform offset inside it
      const int nv = a.prop_nv, nvox = nv * nv * nv;
      const float* cdf = a.prop_cdf + (size_t)ch * nvox;
      const int v = gc_sample_cdf(cdf, a.prop_guide + (size_t)ch * nvox, nvox, u01(rb[3]));
      const int i0 = v / (nv * nv), i1 = (v / nv) % nv, i2 = v % nv;
      fu[0] = (i0 + fu[0]) / nv;
      fu[1] = (i1 + fu[1]) / nv;
      fu[2] = (i2 + fu[2]) / nv;
      voxel = v;
      log_ratio = -log(gc_qn(cdf, v, nvox));
    } else if (a.prop_nv > 0) {
      voxel = gc_voxel((float)fu[0], (float)fu[1], (float)fu[2], a.prop_nv);
    }
    double off[3];
    tmul(cell.L, fu[0], fu[1], fu[2], off[0], off[1], off[2]);
    if (ng < a.gmax) {
      for (int s = 0; s < Mt; ++s) {
        const double tp[3] = {motif[s].x, motif[s].y, motif[s].z};
        for (int x = 0; x < 3; ++x)
          snew[s][x] = tp[0] * R[0][x] + tp[1] * R[1][x] + tp[2] * R[2][x] + off[x];
      }
      n_new = Mt;
    }
  }
  for (int s = 0; s < n_new; ++s) {
    gc_wrap(cell, snew[s]);
    site_new[s] = make_float4((float)snew[s][0], (float)snew[s][1], (float)snew[s][2], motif[s].w);
  }
  tr.move = move; tr.deleting = deleting; tr.n_old = n_old; tr.n_new = n_new; tr.slot = slot;
  tr.voxel = voxel; tr.log_ratio = log_ratio; tr.motif = t;
}

template <int KPT>
__global__ __launch_bounds__(GC_THREADS) void gcmc_kernel(GcArgs a) {
  extern __shared__ float4 lds[];
  __shared__ GcCell cell;
  __shared__ GcChain cs;
  __shared__ GcTrial tr;
  __shared__ float s_sig2[GC_NSPEC * GC_NSPEC], s_e4[GC_NSPEC * GC_NSPEC];
  __shared__ float4 s_motif[GC_NMOTIF][GC_MSITE];
  __shared__ int s_mspec[GC_NMOTIF][GC_MSITE];
  __shared__ int s_mn[GC_NMOTIF];  // sites per motif (indexed at run time: LDS, not arguments)
  __shared__ float4 site_old[GC_MSITE], site_new[GC_MSITE];
  __shared__ double site64[2 * GC_MSITE][3];  // FP64 old then new sites
  __shared__ double2 tab[2 * GC_MSITE][3][GC_NM + 1];
  __shared__ double red[2][GC_WAVES];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int M = a.m, ns = a.nspec;
  float4* host = lds;
  float4* gst = lds + a.max_host;                            // guests (x, y, z, q) FP32
  double* g64 = reinterpret_cast<double*>(gst + a.gmax * M);  // guests FP64 (state)
  uint8_t* hspec = reinterpret_cast<uint8_t*>(g64 + a.gmax * M * 3);
  uint8_t* gmot = hspec + ((a.max_host + 15) & ~15);        // motif of each guest slot
  for (int k = tid; k < ns * ns; k += GC_THREADS) {
    const double sg = a.sigma[k];
    s_sig2[k] = (float)(sg * sg);
    s_e4[k] = (float)(4.0 * a.epsilon[k]);
  }
  if (tid < GC_NMOTIF) s_mn[tid] = a.mnsite[tid];
  if (tid < a.n_motif * M) {
    const int t = tid / M, s = tid % M;
    s_motif[t][s] = make_float4(a.motif_pos[3 * tid], a.motif_pos[3 * tid + 1],
                                a.motif_pos[3 * tid + 2], a.motif_q[tid]);
    s_mspec[t][s] = a.motif_spec[tid];
  }
  const float rclj2 = (float)(a.rc_lj * a.rc_lj), rccl2 = (float)(a.rc_cl * a.rc_cl);
  const float K = 14.399645351950548f;

  for (int ch = blockIdx.x; ch < a.n_chain; ch += gridDim.x) {
    __syncthreads();
    const int h0 = a.host_off[ch], nh = a.host_off[ch + 1] - h0;
    if (tid == 0) {
      const double* t = a.tril + 6 * ch;
      tril_inverse(t, cell.Li);
      for (int k = 0; k < 6; ++k) {
        cell.L[k] = t[k]; cell.Lf[k] = (float)t[k]; cell.Lif[k] = (float)cell.Li[k];
      }
      cs.kT = a.kT[ch];
      const double lnV = log(fabs(t[0] * t[2] * t[5]));
      for (int u = 0; u < a.n_motif; ++u) {
        cs.lnfV[u] = a.log_activity[(size_t)ch * a.n_motif + u] + lnV;
        cs.molc[u] = a.mol_const[(size_t)ch * a.n_motif + u];
        cs.nmol[u] = 0;
      }
      const int ng0 = a.n_guest[ch];
      for (int g = 0; g < ng0; ++g) cs.nmol[a.guest_motif[(size_t)ch * a.gmax + g]] += 1;
      // LJ tail: E(c) = c.T.c; with m_u the species counts of one molecule of
      // motif u and T symmetrised, E(c + dn m_u) - E(c) = 2 dn (T m_u).c + m_u.T.m_u
      const double* T = a.tail + (size_t)ch * ns * ns;
      for (int x = 0; x < ns; ++x) cs.cnt[x] = a.counts[(size_t)ch * ns + x];
      for (int u = 0; u < a.n_motif; ++u) {
        // (T m_u)_y = sum over the sites s of u of T_sym[spec(s), y]
        double tc = 0.0;
        for (int y = 0; y < ns; ++y) {
          double tv = 0.0;
          for (int s = 0; s < s_mn[u]; ++s) {
            const int x = s_mspec[u][s];
            tv += 0.5 * (T[x * ns + y] + T[y * ns + x]);
          }
          tc += tv * cs.cnt[y];
        }
        cs.tcm[u] = tc;
        for (int v = 0; v < a.n_motif; ++v) {
          double mm = 0.0;
          for (int s = 0; s < s_mn[u]; ++s)
            for (int r = 0; r < s_mn[v]; ++r) {
              const int x = s_mspec[u][s], y = s_mspec[v][r];
              mm += 0.5 * (T[x * ns + y] + T[y * ns + x]);
            }
          cs.mtm[u][v] = mm;
        }
      }
      for (int k = 0; k < 4; ++k) {
        cs.wid[k] = a.width[(size_t)ch * 4 + k];
        cs.hist_idx[k] = a.hist_idx[(size_t)ch * 4 + k];
      }
      for (int k = 0; k < 16; ++k) cs.hist[k] = a.hist[(size_t)ch * 16 + k];
      for (int k = 0; k < GC_NSTATS; ++k) cs.stats[k] = a.stats[(size_t)ch * GC_NSTATS + k];
    }
    for (int j = tid; j < nh; j += GC_THREADS) {
      host[j] = a.host[h0 + j];
      hspec[j] = (uint8_t)a.host_spec[h0 + j];
    }
    double* gpos = a.guest + (size_t)ch * a.gmax * M * 3;
    int32_t* gmo = a.guest_motif + (size_t)ch * a.gmax;
    int ng = a.n_guest[ch];
    for (int g = tid; g < ng; g += GC_THREADS) gmot[g] = (uint8_t)gmo[g];
    for (int j = tid; j < ng * M; j += GC_THREADS) {
      const double x = gpos[3 * j], y = gpos[3 * j + 1], z = gpos[3 * j + 2];
      g64[3 * j] = x; g64[3 * j + 1] = y; g64[3 * j + 2] = z;
      gst[j] = make_float4((float)x, (float)y, (float)z, a.motif_q[gmo[j / M] * M + j % M]);
    }
    // this lane's k vectors: packed n, P(k), S(k)
    double* S = a.sk + (size_t)ch * a.nk * 2;
    int rkn[KPT];
    double rP[KPT];
    double2 rS[KPT];
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
      const int k = tid + q * GC_THREADS;
      rkn[q] = 0; rP[q] = 0.0; rS[q] = make_double2(0.0, 0.0);
      if (k < a.nk) {
        rkn[q] = a.kn[(size_t)ch * a.nk + k];
        rP[q] = a.pk[(size_t)ch * a.nk + k];
        rS[q] = make_double2(S[2 * k], S[2 * k + 1]);
      }
    }
    const float alpha = (float)a.alpha[ch];
    double energy = a.energy[ch];
    __syncthreads();

    const int64_t step0 = a.step_dev ? a.step_dev[0] : a.step0;
    for (int st = 0; st < a.n_steps; ++st) {
      const int64_t t = step0 + st;
      // ---- 1. draws and logs (wave 0), proposal (lane 0)
      double lu = 0.0, ln_np1 = 0.0, ln_n = 0.0;
      if (wv == 0) {
        const int l = lane < 4 ? lane : 3;
        const uint32_t tag = l == 0 ? GC_TAG_BATCH : (l == 1 ? GC_TAG_A : (l == 2 ? GC_TAG_B : GC_TAG_C));
        uint32_t w[4];
        gc_draw(a.seed, l == 0 ? 0xFFFFFFFFu : (uint32_t)ch, t, tag, w);
        // ln u, and ln(N_t + 1), ln N_t of the exchange motif t (from ra[2])
        const int nt = cs.nmol[gc_readlane(w[2], 1) % (uint32_t)a.n_motif];
        const double x = lane == 1 ? (double)u01(w[1])
                                   : (lane == 2 ? (double)(nt + 1) : (double)(nt > 0 ? nt : 1));
        const double lx = log(x);
        uint32_t bb[4], ra[4], rb[4], rc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          bb[i] = gc_readlane(w[i], 0);
          ra[i] = gc_readlane(w[i], 1);
          rb[i] = gc_readlane(w[i], 2);
          rc[i] = gc_readlane(w[i], 3);
        }
        lu = gc_readlane(lx, 1);
        ln_np1 = gc_readlane(lx, 2);
        ln_n = gc_readlane(lx, 3);
        if (lane == 0)
          gc_propose(a, cell, cs, ch, ng, g64, gmot, s_motif, s_mn, bb, ra, rb, rc, site_old,
                     site_new, site64, tr);
      }
      __syncthreads();
      const int n_old = tr.n_old, n_new = tr.n_new, slot = tr.slot, tm = tr.motif;
      const int nsite = n_old + n_new;
      // ---- 2a. FP64 axis phase tables of the changed sites, n = 0..nm
      if (tid < 3 * nsite) {
        const int s = tid / 3, ax = tid % 3;
        double f0, f1, f2;
        tmul(cell.Li, site64[s][0], site64[s][1], site64[s][2], f0, f1, f2);
        double f = ax == 0 ? f0 : (ax == 1 ? f1 : f2);
        f -= rint(f);
        double2 u;
        sincospi(2.0 * f, &u.y, &u.x);
        double2 p = make_double2(1.0, 0.0);
        tab[s][ax][0] = p;
        for (int n = 1; n <= a.nm; ++n) {
          p = gc_zmul(p, u);
          tab[s][ax][n] = p;
        }
      }
      // ---- 2b. pair energy change: one lane per partner
      float e_new = 0.f, e_old = 0.f;
      if (nsite) {
        const int ntot = nh + ng * M;
        for (int j = tid; j < ntot; j += GC_THREADS) {
          float4 rj;
          int sj;
          if (j < nh) {
            rj = host[j];
            sj = hspec[j];
          } else {
            const int gj = j - nh, g = gj / M, sg = gj - g * M;
            const int gm = gmot[g];
            if (g == slot || sg >= s_mn[gm]) continue;
            rj = gst[gj];
            sj = s_mspec[gm][sg];
          }
          const float rjv[3] = {rj.x, rj.y, rj.z};
          for (int s = 0; s < nsite; ++s) {
            const bool is_new = s >= n_old;
            const float4 ri = is_new ? site_new[s - n_old] : site_old[s];
            const int si = s_mspec[tm][is_new ? s - n_old : s];
            const float riv[3] = {ri.x, ri.y, ri.z};
            float d[3];
            gc_minimg(cell, riv, rjv, d);
            const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            // each cutoff test folded into its 1/r factor (zero outside)
            const int p = si * ns + sj;
            const float ir2 = r2 < rclj2 ? __builtin_amdgcn_rcpf(r2) : 0.0f;
            float c6 = s_sig2[p] * ir2;
            c6 = c6 * c6 * c6;
            float e = s_e4[p] * c6 * (c6 - 1.0f);
            const float ir = __builtin_amdgcn_rsqf(r2);
            const float irc = r2 < rccl2 ? ir : 0.0f;
            e += K * ri.w * rj.w * erfc_cheb(alpha * (r2 * irc)) * irc;
            if (is_new) e_new += e; else e_old += e;
          }
        }
      }
      __syncthreads();  // tables ready
      // ---- 3. reciprocal-space change over this lane's k vectors (FP64)
      double2 dS[KPT];
      double elr = 0.0;
#pragma unroll
      for (int q = 0; q < KPT; ++q) {
        dS[q] = make_double2(0.0, 0.0);
        if (!nsite || rP[q] == 0.0) continue;
        const int code = rkn[q];
        const int n0 = code & 0xff, n1 = ((code >> 8) & 0xff) - 64, n2 = ((code >> 16) & 0xff) - 64;
        const int a1 = n1 < 0 ? -n1 : n1, a2 = n2 < 0 ? -n2 : n2;
        double dr = 0.0, di = 0.0;
        for (int s = 0; s < nsite; ++s) {
          double2 e1 = tab[s][1][a1], e2 = tab[s][2][a2];
          if (n1 < 0) e1.y = -e1.y;
          if (n2 < 0) e2.y = -e2.y;
          const double2 e = gc_zmul(gc_zmul(tab[s][0][n0], e1), e2);
          const double q_ = s < n_old ? -(double)site_old[s].w : (double)site_new[s - n_old].w;
          dr += q_ * e.x;
          di += q_ * e.y;
        }
        dS[q] = make_double2(dr, di);
        elr += rP[q] * (2.0 * (rS[q].x * dr + rS[q].y * di) + dr * dr + di * di);
      }
      // ---- 4. one block reduction, acceptance on lane 0
      {
        const double vp = wave_sum((double)e_new - (double)e_old), vl = wave_sum(elr);
        if (lane == 0) { red[0][wv] = vp; red[1][wv] = vl; }
      }
      __syncthreads();
      if (tid == 0) {
        double dpair = 0.0, dlr = 0.0;
        for (int w = 0; w < GC_WAVES; ++w) { dpair += red[0][w]; dlr += red[1][w]; }
        const int move = tr.move;
        double dE = dpair + dlr * 14.399645351950548;
        double logp = 0.0;
        int dn = 0;
        if (move == 3) dn = n_new ? 1 : (n_old ? -1 : 0);
        if (dn) {
          dE += dn * cs.molc[tm] + 2.0 * dn * cs.tcm[tm] + cs.mtm[tm][tm];
          logp += dn > 0 ? cs.lnfV[tm] - ln_np1 : -cs.lnfV[tm] + ln_n;
        }
        logp += -dE / cs.kT + tr.log_ratio;
        if (a.sample_count && move == 3 && n_new && tr.voxel >= 0) {
          const int at = atomicAdd(a.sample_count, 1);
          if (at < a.sample_cap) {
            a.sample_row[at] = ch * (a.prop_nv * a.prop_nv * a.prop_nv) + tr.voxel;
            a.sample_y[at] = (float)(-dE / cs.kT);
          }
        }
        const int accept = logp > lu;
        if (ch < a.n_trace) {
          double* trc = a.trace + ((size_t)ch * a.n_steps + st) * 4;
          trc[0] = move * 2 + (move == 3 ? tr.deleting : 0);
          trc[1] = dE; trc[2] = logp; trc[3] = accept;
        }
        cs.stats[move] += 1;
        // trials that evaluated any site, and the site evaluations (old and
        // new configurations) they made: the work behind the trial count
        if (nsite) {
          cs.stats[10] += 1;
          cs.stats[11] += nsite;
        }
        if (accept) {
          if (nsite) cs.stats[4 + move] += 1;  // moves that changed the configuration
          energy += dE;
          if (dn) {
            for (int s = 0; s < s_mn[tm]; ++s) cs.cnt[s_mspec[tm][s]] += dn;
            for (int u = 0; u < a.n_motif; ++u) cs.tcm[u] += dn * cs.mtm[u][tm];
            cs.nmol[tm] += dn;
            cs.stats[dn > 0 ? 8 : 9] += 1;
          }
        }
        // acceptance-target schedule (parameter_scheduler.py:161-208)
        uint32_t* hb = cs.hist + move * 4;
        int idx = cs.hist_idx[move];
        if (accept) hb[idx >> 5] |= (1u << (idx & 31)); else hb[idx >> 5] &= ~(1u << (idx & 31));
        idx = (idx + 1) % 100;
        cs.hist_idx[move] = idx;
        if (idx == 0) {
          int ones = 0;
          for (int w = 0; w < 4; ++w) ones += __popc(hb[w] & (w == 3 ? 0xFu : 0xFFFFFFFFu));
          const double avg = ones / 100.0;
          if (fabs(avg - 0.5) > 0.05) {
            double v = cs.wid[move];
            v = avg < 0.5 ? v / 1.1 : v * 1.1;
            if (v < 0) v = 0;
            if (v > a.upper[move]) v = a.upper[move];
            cs.wid[move] = (float)v;
          }
        }
        tr.accept = accept;
      }
      __syncthreads();
      // ---- 5. apply an accepted move
      if (tr.accept && nsite) {
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
          rS[q].x += dS[q].x;
          rS[q].y += dS[q].y;
        }
        // guest table: replace the moved molecule, append an inserted one, or
        // move the last molecule into a deleted one's slot
        const int dst = n_old ? slot : ng;
        if (tid < M) {
          const int d = dst * M + tid;
          if (n_new) {
            if (tid < n_new) {
              gst[d] = site_new[tid];
              for (int x = 0; x < 3; ++x) g64[3 * d + x] = site64[n_old + tid][x];
            }
          } else {
            const int src = (ng - 1) * M + tid;
            gst[d] = gst[src];
            for (int x = 0; x < 3; ++x) g64[3 * d + x] = g64[3 * src + x];
          }
        }
        if (tid == 0) {
          if (n_new && !n_old) gmot[ng] = (uint8_t)tm;
          else if (!n_new) gmot[slot] = gmot[ng - 1];
        }
        ng += n_new && !n_old ? 1 : (!n_new ? -1 : 0);
      }
      __syncthreads();
    }
    // ---- store the chain
#pragma unroll
    for (int q = 0; q < KPT; ++q) {
      const int k = tid + q * GC_THREADS;
      if (k < a.nk) { S[2 * k] = rS[q].x; S[2 * k + 1] = rS[q].y; }
    }
    for (int j = tid; j < ng * M * 3; j += GC_THREADS) gpos[j] = g64[j];
    for (int g = tid; g < ng; g += GC_THREADS) gmo[g] = gmot[g];
    if (tid == 0) {
      a.n_guest[ch] = ng;
      a.energy[ch] = energy;
      for (int x = 0; x < ns; ++x) a.counts[(size_t)ch * ns + x] = cs.cnt[x];
      for (int k = 0; k < 4; ++k) {
        a.width[(size_t)ch * 4 + k] = cs.wid[k];
        a.hist_idx[(size_t)ch * 4 + k] = cs.hist_idx[k];
      }
      for (int k = 0; k < 16; ++k) a.hist[(size_t)ch * 16 + k] = cs.hist[k];
      for (int k = 0; k < GC_NSTATS; ++k) a.stats[(size_t)ch * GC_NSTATS + k] = cs.stats[k];
    }
  }
}

template <int KPT>
int gcmc_launch(const GcArgs& a, size_t lds, hipStream_t st) {
  int dev = 0, ncu = 256, per_cu = 1;
  KUPSMI_HIP(hipGetDevice(&dev));
  KUPSMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  KUPSMI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, gcmc_kernel<KPT>, GC_THREADS, lds));
  KUPSMI_REQUIRE(per_cu >= 1, "the GCMC kernel does not fit on a CU with this LDS size");
  const int resident = per_cu * ncu;
  const int grid = a.n_chain < resident ? a.n_chain : resident;
  gcmc_kernel<KPT><<<grid, GC_THREADS, lds, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_gcmc_run(void* stream, int n_chain, int n_steps, int64_t step0,
                               uint64_t seed, int max_host, const int32_t* host_off,
                               const float* host_xyzq, const int32_t* host_species,
                               const double* tril, int gmax, int m, int n_motif,
                               const int32_t* motif_nsite, const float* motif_pos,
                               const float* motif_q, const int32_t* motif_species, int n_species,
                               const double* sigma, const double* epsilon, double rc_lj,
                               double rc_coul, const double* alpha, const double* tail,
                               const double* mol_const, const double* kT,
                               const double* log_activity, int nk, int nm, const int32_t* kn,
                               const double* pk, const double* move_weights,
                               const double* width_upper, double* guest, int32_t* n_guest,
                               int32_t* guest_motif, int32_t* counts,
                               double* structure_factor, double* energy,
                               float* width, uint32_t* hist, int32_t* hist_idx,
                               long long* stats, int n_trace, double* trace,
                               const float* prop_cdf, const int32_t* prop_guide, int prop_nv,
                               int* sample_count, int sample_cap, int32_t* sample_row,
                               float* sample_y, const int64_t* step_dev) {
  KUPSMI_REQUIRE(n_chain >= 0 && n_steps >= 0, "negative size");
  KUPSMI_REQUIRE(m >= 1 && m <= GC_MSITE, "molecules must have 1..4 sites");
  KUPSMI_REQUIRE(n_motif >= 1 && n_motif <= GC_NMOTIF, "1..4 guest motifs");
  KUPSMI_REQUIRE(motif_nsite != nullptr, "sites per motif required (host array)");
  for (int t = 0; t < n_motif; ++t)
    KUPSMI_REQUIRE(motif_nsite[t] >= 1 && motif_nsite[t] <= m, "a motif has 1..m sites");
  KUPSMI_REQUIRE(guest_motif != nullptr, "guest motif table required");
  KUPSMI_REQUIRE(n_species >= 1 && n_species <= GC_NSPEC, "n_species must be in [1, 16]");
  KUPSMI_REQUIRE(nk >= 0 && nk <= GC_KPT_MAX * GC_THREADS, "at most 4096 k vectors per chain");
  KUPSMI_REQUIRE(nm >= 0 && nm <= GC_NM, "k list half width must be <= 12");
  KUPSMI_REQUIRE(gmax >= 1 && max_host >= 0, "bad buffer sizes");
  const size_t lds = sizeof(float4) * ((size_t)max_host + (size_t)gmax * m) +
                     sizeof(double) * 3 * (size_t)gmax * m + (((size_t)max_host + 15) & ~(size_t)15) +
                     (size_t)gmax;
  KUPSMI_REQUIRE(lds <= 120 * 1024, "host + guest buffer exceeds the LDS budget of a chain");
  if (n_chain == 0 || n_steps == 0) return KUPSMI_OK;
  GcArgs a{};
  a.n_chain = n_chain; a.n_steps = n_steps; a.gmax = gmax; a.m = m; a.nspec = n_species;
  a.n_motif = n_motif;
  for (int t = 0; t < GC_NMOTIF; ++t) a.mnsite[t] = t < n_motif ? motif_nsite[t] : 0;
  a.nk = nk; a.nm = nm; a.n_trace = n_trace; a.max_host = max_host;
  a.step0 = step0; a.step_dev = step_dev; a.seed = seed; a.rc_lj = rc_lj; a.rc_cl = rc_coul;
  for (int k = 0; k < 4; ++k) {
    a.w_move[k] = move_weights ? move_weights[k] : 1.0;
    a.upper[k] = width_upper ? width_upper[k] : 1e30;
  }
  a.host_off = host_off; a.host = reinterpret_cast<const float4*>(host_xyzq);
  a.host_spec = host_species; a.tril = tril; a.motif_pos = motif_pos; a.motif_q = motif_q;
  a.motif_spec = motif_species; a.sigma = sigma; a.epsilon = epsilon; a.alpha = alpha;
  a.tail = tail; a.mol_const = mol_const; a.kT = kT; a.log_activity = log_activity;
  a.kn = kn; a.pk = pk; a.guest = guest; a.n_guest = n_guest; a.guest_motif = guest_motif;
  a.counts = counts;
  a.sk = structure_factor; a.energy = energy; a.width = width; a.hist = hist;
  a.hist_idx = hist_idx; a.stats = stats; a.trace = trace;
  KUPSMI_REQUIRE(!prop_cdf || (prop_nv >= 1 && prop_guide),
                 "a proposal table needs its voxel count and guide table");
  KUPSMI_REQUIRE(!sample_count || (prop_nv >= 1 && sample_row && sample_y),
                 "insertion samples need the voxel grid and output buffers");
  a.prop_cdf = prop_cdf; a.prop_guide = prop_guide; a.prop_nv = prop_nv;
  a.sample_count = sample_count; a.sample_cap = sample_cap; a.sample_row = sample_row;
  a.sample_y = sample_y;
  hipStream_t st = (hipStream_t)stream;
  if (nk <= 2 * GC_THREADS) return gcmc_launch<2>(a, lds, st);
  if (nk <= 4 * GC_THREADS) return gcmc_launch<4>(a, lds, st);
  return gcmc_launch<8>(a, lds, st);
}
