// Rigid-molecule GCMC of a batch of chains, every chain resident in LDS.
//
// One 512-thread workgroup owns one chain (host framework + guest molecules
// in LDS, the structure-factor cache S(k) in HBM/L2) and runs n_steps trial
// moves without leaving the CU.  Per trial:
//   1. move type and insert/delete coin shared by the whole batch, per-chain
//      proposal (translation, rotation, reinsertion, insertion, deletion);
//   2. Delta E = sum over the changed sites of LJ + Ewald real-space pairs
//      with the host and the other guests (one lane per partner),
//      + K sum_k P(k) (|S + dS|^2 - |S|^2) over the chain's k list (one lane
//      per k, phases from per-site axis tables) + self/exclusion constant of
//      an inserted/deleted molecule + analytic LJ tail change, FP64 sums;
//   3. FP64 log acceptance ratio (Boltzmann + fugacity), Metropolis test,
//      update of the guest table, S(k) and the energy on accept;
//   4. acceptance-target step-width schedule of the chosen move.
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

constexpr int GC_THREADS = 512;
constexpr int GC_MSITE = 4;      // sites per molecule
constexpr int GC_NM = 12;        // max |n| per axis of the k list
constexpr int GC_W = 2 * GC_NM + 1;
constexpr int GC_KPT = 8;        // k vectors per thread (nk <= 4096)
constexpr int GC_NSPEC = 16;
constexpr uint32_t GC_TAG_BATCH = 0x42415443u, GC_TAG_A = 0x47434d41u,
                   GC_TAG_B = 0x47434d42u, GC_TAG_C = 0x47434d43u;

struct GcArgs {
  int n_chain, n_steps, gmax, m, nspec, nk, nm, n_trace, max_host;
  int64_t step0;
  uint64_t seed;
  double rc_lj, rc_cl;
  double w_move[4], upper[4];
  const int32_t* host_off;
  const float4* host;        // x, y, z, q
  const int32_t* host_spec;
  const double* tril;
  const float* motif_pos;    // m x 3, centred
  const float* motif_q;
  const int32_t* motif_spec;
  const double* sigma;       // nspec x nspec
  const double* epsilon;
  const double* alpha;       // per chain
  const double* tail;        // n_chain x nspec x nspec
  const double* mol_const;   // per chain: self + exclusion energy of one molecule
  const double* kT;
  const double* log_activity;
  const int32_t* kn;         // n_chain x nk packed (n0, n1 + 64, n2 + 64)
  const double* pk;          // n_chain x nk, P(k) (0 for padding)
  // state (in/out)
  float* guest;              // n_chain x gmax x m x 3
  int32_t* n_guest;
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
  int prop_nv;
  int* sample_count;
  int sample_cap;
  int32_t* sample_row;       // chain * nv^3 + voxel of the insertion offset
  float* sample_y;           // -dE / kT of the insertion trial
};

__device__ inline float2 gc_cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline void gc_draw(uint64_t seed, uint32_t a, int64_t t, uint32_t tag, uint32_t o[4]) {
  o[0] = a; o[1] = (uint32_t)t; o[2] = (uint32_t)((uint64_t)t >> 32); o[3] = tag;
  philox4x32_10(o, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Shoemake quaternion, power w, rotation matrix (core/utils/quaternion.py).
__device__ void gc_rotation(const float u[3], float w, float R[3][3]) {
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
  const float W = ch, X = ax * sh, Y = ay * sh, Z = az * sh;
  const float x2 = X * X, y2 = Y * Y, z2 = Z * Z, w2 = W * W;
  R[0][0] = x2 - y2 - z2 + w2; R[0][1] = 2 * (X * Y - Z * W); R[0][2] = 2 * (X * Z + Y * W);
  R[1][0] = 2 * (X * Y + Z * W); R[1][1] = -x2 + y2 - z2 + w2; R[1][2] = 2 * (Y * Z - X * W);
  R[2][0] = 2 * (X * Z - Y * W); R[2][1] = 2 * (Y * Z + X * W); R[2][2] = -x2 - y2 + z2 + w2;
}

struct GcCell {
  double L[6];
  float Lf[6], Lif[6];
};

__device__ inline void gc_wrap(const GcCell& c, float* r) {
  float f0, f1, f2;
  wrap_shift(c.L, c.Lif, r[0], r[1], r[2], f0, f1, f2);
}

// minimum-image vector b - a
__device__ inline void gc_minimg(const GcCell& c, const float* a, const float* b, float* d) {
  float f0, f1, f2;
  tmul(c.Lif, b[0] - a[0], b[1] - a[1], b[2] - a[2], f0, f1, f2);
  f0 -= rintf(f0); f1 -= rintf(f1); f2 -= rintf(f2);
  tmul(c.Lf, f0, f1, f2, d[0], d[1], d[2]);
}

__device__ inline double gc_block_sum(double v, double* red) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < GC_THREADS / 64; ++w) t += red[w];
  return t;  // every thread holds the sum
}

// Proposal of one trial (thread 0): move type, sites before/after the move.
// Out of line: it runs on one lane and its registers must not weigh on the
// pair and k loops of the whole workgroup.
struct GcProposal {
  int move, deleting, n_old, n_new, slot, voxel;
  uint32_t accept_bits;
  double log_ratio;  // proposal log ratio ln(q(x|x') / q(x'|x)) of the move
};

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

__device__ __noinline__ GcProposal gc_propose(const GcArgs& a, const GcCell& cell, int ch, int64_t t,
                                              int ng, const float4* gst, const float* wid,
                                              float4* site_old, float4* site_new) {
  const int M = a.m;
  uint32_t bb[4], ra[4], rb[4], rcn[4];
  gc_draw(a.seed, 0xFFFFFFFFu, t, GC_TAG_BATCH, bb);
  const double wsum = a.w_move[0] + a.w_move[1] + a.w_move[2] + a.w_move[3];
  const double um = (double)u01(bb[0]) * wsum;
  double acc = 0;
  int move = 3;
  for (int k = 0; k < 4; ++k) {
    acc += a.w_move[k];
    if (um < acc) { move = k; break; }
  }
  const int deleting = (bb[1] >> 31) & 1;
  gc_draw(a.seed, (uint32_t)ch, t, GC_TAG_A, ra);
  gc_draw(a.seed, (uint32_t)ch, t, GC_TAG_B, rb);
  gc_draw(a.seed, (uint32_t)ch, t, GC_TAG_C, rcn);
  int n_old = 0, n_new = 0, slot = -1, voxel = -1;
  double log_ratio = 0.0;
  float oldp[GC_MSITE][3], newp[GC_MSITE][3];
  if (move < 3 || deleting) {
    if (ng > 0) {
      slot = (int)(ra[0] % (uint32_t)ng);
      n_old = M;
      for (int s = 0; s < M; ++s) {
        const float4 g = gst[slot * M + s];
        oldp[s][0] = g.x; oldp[s][1] = g.y; oldp[s][2] = g.z;
      }
    }
  }
  if (move == 3 && deleting && n_old && a.prop_cdf) {
    // reverse move = learned insertion at the molecule's centre
    float rel[3], cen[3] = {0.f, 0.f, 0.f};
    for (int s = 0; s < M; ++s) {
      gc_minimg(cell, oldp[M - 1], oldp[s], rel);
      for (int x = 0; x < 3; ++x) cen[x] += rel[x] / M;
    }
    float f0, f1, f2;
    tmul(cell.Lif, oldp[M - 1][0] + cen[0], oldp[M - 1][1] + cen[1], oldp[M - 1][2] + cen[2],
         f0, f1, f2);
    f0 -= floorf(f0); f1 -= floorf(f1); f2 -= floorf(f2);
    const int nv = a.prop_nv, nvox = nv * nv * nv;
    const int v = gc_voxel(f0, f1, f2, nv);
    log_ratio = log(gc_qn(a.prop_cdf + (size_t)ch * nvox, v, nvox));
  }
  if (move == 0 && n_old) {
    const float u0 = u01(rcn[0]), u1 = u01(rcn[1]), u2 = u01(rcn[2]), u3 = u01(rcn[3]);
    const float r0 = sqrtf(-2.0f * logf(u0)), r1 = sqrtf(-2.0f * logf(u2));
    float s1, c1, s3, c3;
    sincospif(2.0f * u1, &s1, &c1);
    sincospif(2.0f * u3, &s3, &c3);
    const float g3[3] = {r0 * c1, r0 * s1, r1 * c3};
    for (int s = 0; s < M; ++s) {
      for (int x = 0; x < 3; ++x) newp[s][x] = oldp[s][x] + g3[x] * wid[0];
      gc_wrap(cell, newp[s]);
    }
    n_new = M;
  } else if ((move == 1 || move == 2) && n_old) {
    const float uq[3] = {u01(rb[0]), u01(rb[1]), u01(rb[2])};
    float R[3][3];
    gc_rotation(uq, move == 1 ? wid[1] : 1.0f, R);
    float rel[3], cen[3] = {0.f, 0.f, 0.f};
    for (int s = 0; s < M; ++s) {
      gc_minimg(cell, oldp[M - 1], oldp[s], rel);
      for (int x = 0; x < 3; ++x) cen[x] += rel[x] / M;
    }
    float com[3] = {oldp[M - 1][0] + cen[0], oldp[M - 1][1] + cen[1], oldp[M - 1][2] + cen[2]};
    gc_wrap(cell, com);
    float off[3] = {0.f, 0.f, 0.f};
    if (move == 2) tmul(cell.Lf, u01(rcn[0]), u01(rcn[1]), u01(rcn[2]), off[0], off[1], off[2]);
    for (int s = 0; s < M; ++s) {
      float rr[3];
      gc_minimg(cell, com, oldp[s], rr);
      for (int x = 0; x < 3; ++x)
        newp[s][x] = com[x] + rr[0] * R[0][x] + rr[1] * R[1][x] + rr[2] * R[2][x] + off[x];
      gc_wrap(cell, newp[s]);
    }
    n_new = M;
  } else if (move == 3 && !deleting) {
    const float uq[3] = {u01(rb[0]), u01(rb[1]), u01(rb[2])};
    float R[3][3];
    gc_rotation(uq, 1.0f, R);
    float fu[3] = {u01(rcn[0]), u01(rcn[1]), u01(rcn[2])};
    if (a.prop_cdf) {
      // voxel ~ q (first v with u < cdf[v]), uniform offset inside it
      const int nv = a.prop_nv, nvox = nv * nv * nv;
      const float* cdf = a.prop_cdf + (size_t)ch * nvox;
      const float uv = u01(rb[3]);
      int lo = 0, hi = nvox - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (uv < cdf[mid]) hi = mid; else lo = mid + 1;
      }
      const int i0 = lo / (nv * nv), i1 = (lo / nv) % nv, i2 = lo % nv;
      fu[0] = (i0 + fu[0]) / nv;
      fu[1] = (i1 + fu[1]) / nv;
      fu[2] = (i2 + fu[2]) / nv;
      voxel = lo;
      log_ratio = -log(gc_qn(cdf, lo, nvox));
    } else if (a.prop_nv > 0) {
      voxel = gc_voxel(fu[0], fu[1], fu[2], a.prop_nv);
    }
    float off[3];
    tmul(cell.Lf, fu[0], fu[1], fu[2], off[0], off[1], off[2]);
    if (ng < a.gmax) {
      for (int s = 0; s < M; ++s) {
        const float* tp = a.motif_pos + 3 * s;
        for (int x = 0; x < 3; ++x)
          newp[s][x] = tp[0] * R[0][x] + tp[1] * R[1][x] + tp[2] * R[2][x] + off[x];
        gc_wrap(cell, newp[s]);
      }
      n_new = M;
    }
  }
  for (int s = 0; s < M; ++s) {
    const float q = a.motif_q[s];
    if (n_old) site_old[s] = make_float4(oldp[s][0], oldp[s][1], oldp[s][2], q);
    if (n_new) site_new[s] = make_float4(newp[s][0], newp[s][1], newp[s][2], q);
  }
  GcProposal pr;
  pr.move = move; pr.deleting = deleting; pr.n_old = n_old; pr.n_new = n_new; pr.slot = slot;
  pr.accept_bits = ra[1];
  pr.voxel = voxel;
  pr.log_ratio = log_ratio;
  return pr;
}

__global__ __launch_bounds__(GC_THREADS) void gcmc_kernel(GcArgs a) {
  extern __shared__ float4 lds[];
  __shared__ GcCell cell;
  __shared__ float s_sig2[GC_NSPEC * GC_NSPEC], s_e4[GC_NSPEC * GC_NSPEC];
  __shared__ float4 site_old[GC_MSITE], site_new[GC_MSITE];
  __shared__ float2 tab[2 * GC_MSITE][3][GC_W];
  __shared__ double red[GC_THREADS / 64];
  __shared__ int s_move, s_del, s_nold, s_nnew, s_slot, s_accept, s_voxel;
  __shared__ double s_dE, s_lratio;
  const int tid = threadIdx.x;
  const int M = a.m, ns = a.nspec;
  float4* host = lds;
  float4* gst = lds + a.max_host;
  int* hspec = reinterpret_cast<int*>(gst + a.gmax * M);
  for (int k = tid; k < ns * ns; k += GC_THREADS) {
    const double sg = a.sigma[k];
    s_sig2[k] = (float)(sg * sg);
    s_e4[k] = (float)(4.0 * a.epsilon[k]);
  }
  const float rclj2 = (float)(a.rc_lj * a.rc_lj), rccl = (float)a.rc_cl;
  const float K = 14.399645351950548f;

  for (int ch = blockIdx.x; ch < a.n_chain; ch += gridDim.x) {
    __syncthreads();
    const int h0 = a.host_off[ch], nh = a.host_off[ch + 1] - h0;
    if (tid == 0) {
      const double* t = a.tril + 6 * ch;
      double li[6];
      tril_inverse(t, li);
      for (int k = 0; k < 6; ++k) { cell.L[k] = t[k]; cell.Lf[k] = (float)t[k]; cell.Lif[k] = (float)li[k]; }
    }
    for (int j = tid; j < nh; j += GC_THREADS) {
      host[j] = a.host[h0 + j];
      hspec[j] = a.host_spec[h0 + j];
    }
    float* gpos = a.guest + (size_t)ch * a.gmax * M * 3;
    int ng = a.n_guest[ch];
    for (int j = tid; j < ng * M; j += GC_THREADS)
      gst[j] = make_float4(gpos[3 * j], gpos[3 * j + 1], gpos[3 * j + 2], a.motif_q[j % M]);
    const float alpha = (float)a.alpha[ch];
    const float two_a_sqrtpi = 2.0f * alpha * 0.5641895835477563f;
    (void)two_a_sqrtpi;
    const double kT = a.kT[ch];
    const int32_t* kn = a.kn + (size_t)ch * a.nk;
    const double* pk = a.pk + (size_t)ch * a.nk;
    double* S = a.sk + (size_t)ch * a.nk * 2;
    const double* T = a.tail + (size_t)ch * ns * ns;
    int32_t* cnt = a.counts + (size_t)ch * ns;
    float* wid = a.width + (size_t)ch * 4;
    long long* stt = a.stats + (size_t)ch * 10;
    double energy = a.energy[ch];
    __syncthreads();

    for (int st = 0; st < a.n_steps; ++st) {
      const int64_t t = a.step0 + st;
      // ---- proposal (thread 0)
      if (tid == 0) {
        const GcProposal pr = gc_propose(a, cell, ch, t, ng, gst, wid, site_old, site_new);
        s_move = pr.move; s_del = pr.deleting; s_nold = pr.n_old; s_nnew = pr.n_new;
        s_slot = pr.slot;
        s_voxel = pr.voxel;
        s_lratio = pr.log_ratio;
        s_accept = (int)pr.accept_bits;  // accept uniform bits, decided after the reduction
      }
      __syncthreads();
      const int n_old = s_nold, n_new = s_nnew, slot = s_slot;
      const int nsite = n_old + n_new;
      // ---- axis phase tables of the changed sites: e^{2 pi i n f_a}
      if (tid < 3 * nsite) {
        const int s = tid / 3, ax = tid % 3;
        const float4 r = s < n_old ? site_old[s] : site_new[s - n_old];
        float f0, f1, f2;
        tmul(cell.Lif, r.x, r.y, r.z, f0, f1, f2);
        const float f = ax == 0 ? f0 : (ax == 1 ? f1 : f2);
        float2 u;
        sincospif(2.0f * (f - rintf(f)), &u.y, &u.x);
        float2 p = make_float2(1.f, 0.f);
        tab[s][ax][GC_NM] = p;
        for (int n = 1; n <= a.nm; ++n) {
          p = gc_cmul(p, u);
          tab[s][ax][GC_NM + n] = p;
          tab[s][ax][GC_NM - n] = make_float2(p.x, -p.y);
        }
      }
      // ---- pair energy change: one lane per partner
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
            const int gj = j - nh;
            if (gj / M == slot) continue;
            rj = gst[gj];
            sj = a.motif_spec[gj % M];
          }
          const float rjv[3] = {rj.x, rj.y, rj.z};
          for (int s = 0; s < nsite; ++s) {
            const bool is_new = s >= n_old;
            const float4 ri = is_new ? site_new[s - n_old] : site_old[s];
            const int si = a.motif_spec[is_new ? s - n_old : s];
            const float riv[3] = {ri.x, ri.y, ri.z};
            float d[3];
            gc_minimg(cell, riv, rjv, d);
            const float r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            float e = 0.f;
            if (r2 < rclj2) {
              const int p = si * ns + sj;
              float c6 = s_sig2[p] / r2;
              c6 = c6 * c6 * c6;
              e += s_e4[p] * c6 * (c6 - 1.0f);
            }
            const float r = sqrtf(r2);
            if (r < rccl) e += K * ri.w * rj.w * erfcf(alpha * r) / r;
            if (is_new) e_new += e; else e_old += e;
          }
        }
      }
      __syncthreads();  // tables ready
      // ---- reciprocal-space change over this thread's k vectors
      float2 dS[GC_KPT];
      double elr = 0.0;
#pragma unroll
      for (int q = 0; q < GC_KPT; ++q) {
        const int k = tid + q * GC_THREADS;
        dS[q] = make_float2(0.f, 0.f);
        if (k >= a.nk || !nsite) continue;
        const double P = pk[k];
        if (P == 0.0) continue;
        const int code = kn[k];
        const int n0 = code & 0xff, n1 = ((code >> 8) & 0xff) - 64, n2 = ((code >> 16) & 0xff) - 64;
        float dr = 0.f, di = 0.f;
        for (int s = 0; s < nsite; ++s) {
          const float2 e = gc_cmul(gc_cmul(tab[s][0][GC_NM + n0], tab[s][1][GC_NM + n1]),
                                   tab[s][2][GC_NM + n2]);
          const float q_ = s < n_old ? -site_old[s].w : site_new[s - n_old].w;
          dr += q_ * e.x;
          di += q_ * e.y;
        }
        dS[q] = make_float2(dr, di);
        const double sr = S[2 * k], si = S[2 * k + 1];
        elr += P * (2.0 * (sr * dr + si * di) + (double)dr * dr + (double)di * di);
      }
      const double dpair = gc_block_sum((double)e_new - (double)e_old, red);
      const double dlr = gc_block_sum(elr, red) * 14.399645351950548;
      // ---- acceptance (thread 0)
      if (tid == 0) {
        const int move = s_move;
        double dE = dpair + dlr;
        double logp = 0.0;
        int dn = 0;
        if (move == 3) dn = n_new ? 1 : (n_old ? -1 : 0);
        if (dn) {
          dE += dn * a.mol_const[ch];
          double et = 0.0;
          int np_[GC_NSPEC];
          for (int x = 0; x < ns; ++x) np_[x] = cnt[x];
          for (int s = 0; s < M; ++s) np_[a.motif_spec[s]] += dn;
          for (int x = 0; x < ns; ++x)
            for (int y = 0; y < ns; ++y)
              et += T[x * ns + y] * ((double)np_[x] * np_[y] - (double)cnt[x] * cnt[y]);
          dE += et;
          const double vol = fabs(cell.L[0] * cell.L[2] * cell.L[5]);
          const double lnfV = a.log_activity[ch] + log(vol);
          logp += dn > 0 ? lnfV - log((double)(ng + 1)) : -lnfV + log((double)ng);
        }
        logp += -dE / kT + s_lratio;
        if (a.sample_count && move == 3 && n_new && s_voxel >= 0) {
          const int at = atomicAdd(a.sample_count, 1);
          if (at < a.sample_cap) {
            a.sample_row[at] = ch * (a.prop_nv * a.prop_nv * a.prop_nv) + s_voxel;
            a.sample_y[at] = (float)(-dE / kT);
          }
        }
        const double lu = log((double)u01((uint32_t)s_accept));
        const int accept = logp > lu;
        if (ch < a.n_trace) {
          double* tr = a.trace + ((size_t)ch * a.n_steps + st) * 4;
          tr[0] = move * 2 + (move == 3 ? s_del : 0); tr[1] = dE; tr[2] = logp; tr[3] = accept;
        }
        stt[move] += 1;
        if (accept) {
          stt[4 + move] += 1;
          energy += dE;
          if (dn) for (int s = 0; s < M; ++s) cnt[a.motif_spec[s]] += dn;
        }
        // acceptance-target schedule (parameter_scheduler.py:161-208)
        uint32_t* hb = a.hist + ((size_t)ch * 4 + move) * 4;
        int32_t* hi = a.hist_idx + (size_t)ch * 4 + move;
        int idx = *hi;
        if (accept) hb[idx >> 5] |= (1u << (idx & 31)); else hb[idx >> 5] &= ~(1u << (idx & 31));
        idx = (idx + 1) % 100;
        *hi = idx;
        if (idx == 0) {
          int ones = 0;
          for (int w = 0; w < 4; ++w) ones += __popc(hb[w] & (w == 3 ? 0xFu : 0xFFFFFFFFu));
          const double avg = ones / 100.0;
          if (fabs(avg - 0.5) > 0.05) {
            double v = wid[move];
            v = avg < 0.5 ? v / 1.1 : v * 1.1;
            if (v < 0) v = 0;
            if (v > a.upper[move]) v = a.upper[move];
            wid[move] = (float)v;
          }
        }
        s_accept = accept;
        s_dE = dE;
      }
      __syncthreads();
      if (s_accept) {
#pragma unroll
        for (int q = 0; q < GC_KPT; ++q) {
          const int k = tid + q * GC_THREADS;
          if (k < a.nk && nsite) {
            S[2 * k] += (double)dS[q].x;
            S[2 * k + 1] += (double)dS[q].y;
          }
        }
        if (n_new && n_old) {
          if (tid < M) gst[slot * M + tid] = site_new[tid];
        } else if (n_new) {
          if (tid < M) gst[ng * M + tid] = site_new[tid];
          ng += 1;
        } else if (n_old) {
          if (tid < M) gst[slot * M + tid] = gst[(ng - 1) * M + tid];
          ng -= 1;
        }
        if (tid == 0 && s_move == 3) {
          if (n_new) stt[8] += 1;
          else if (n_old) stt[9] += 1;
        }
      }
      __syncthreads();
    }
    // ---- store the chain
    for (int j = tid; j < ng * M; j += GC_THREADS) {
      const float4 g = gst[j];
      gpos[3 * j] = g.x; gpos[3 * j + 1] = g.y; gpos[3 * j + 2] = g.z;
    }
    if (tid == 0) {
      a.n_guest[ch] = ng;
      a.energy[ch] = energy;
    }
  }
}

}  // namespace kupsmi

using namespace kupsmi;

extern "C" int kupsmi_gcmc_run(void* stream, int n_chain, int n_steps, int64_t step0,
                               uint64_t seed, int max_host, const int32_t* host_off,
                               const float* host_xyzq, const int32_t* host_species,
                               const double* tril, int gmax, int m, const float* motif_pos,
                               const float* motif_q, const int32_t* motif_species, int n_species,
                               const double* sigma, const double* epsilon, double rc_lj,
                               double rc_coul, const double* alpha, const double* tail,
                               const double* mol_const, const double* kT,
                               const double* log_activity, int nk, int nm, const int32_t* kn,
                               const double* pk, const double* move_weights,
                               const double* width_upper, float* guest, int32_t* n_guest,
                               int32_t* counts, double* structure_factor, double* energy,
                               float* width, uint32_t* hist, int32_t* hist_idx,
                               long long* stats, int n_trace, double* trace,
                               const float* prop_cdf, int prop_nv, int* sample_count,
                               int sample_cap, int32_t* sample_row, float* sample_y) {
  KUPSMI_REQUIRE(n_chain >= 0 && n_steps >= 0, "negative size");
  KUPSMI_REQUIRE(m >= 1 && m <= GC_MSITE, "molecules must have 1..4 sites");
  KUPSMI_REQUIRE(n_species >= 1 && n_species <= GC_NSPEC, "n_species must be in [1, 16]");
  KUPSMI_REQUIRE(nk >= 0 && nk <= GC_KPT * GC_THREADS, "at most 4096 k vectors per chain");
  KUPSMI_REQUIRE(nm >= 0 && nm <= GC_NM, "k list half width must be <= 12");
  KUPSMI_REQUIRE(gmax >= 1 && max_host >= 0, "bad buffer sizes");
  const size_t lds = sizeof(float4) * ((size_t)max_host + (size_t)gmax * m) +
                     sizeof(int) * (size_t)max_host;
  KUPSMI_REQUIRE(lds <= 120 * 1024, "host + guest buffer exceeds the LDS budget of a chain");
  if (n_chain == 0 || n_steps == 0) return KUPSMI_OK;
  GcArgs a{};
  a.n_chain = n_chain; a.n_steps = n_steps; a.gmax = gmax; a.m = m; a.nspec = n_species;
  a.nk = nk; a.nm = nm; a.n_trace = n_trace; a.max_host = max_host;
  a.step0 = step0; a.seed = seed; a.rc_lj = rc_lj; a.rc_cl = rc_coul;
  for (int k = 0; k < 4; ++k) {
    a.w_move[k] = move_weights ? move_weights[k] : 1.0;
    a.upper[k] = width_upper ? width_upper[k] : 1e30;
  }
  a.host_off = host_off; a.host = reinterpret_cast<const float4*>(host_xyzq);
  a.host_spec = host_species; a.tril = tril; a.motif_pos = motif_pos; a.motif_q = motif_q;
  a.motif_spec = motif_species; a.sigma = sigma; a.epsilon = epsilon; a.alpha = alpha;
  a.tail = tail; a.mol_const = mol_const; a.kT = kT; a.log_activity = log_activity;
  a.kn = kn; a.pk = pk; a.guest = guest; a.n_guest = n_guest; a.counts = counts;
  a.sk = structure_factor; a.energy = energy; a.width = width; a.hist = hist;
  a.hist_idx = hist_idx; a.stats = stats; a.trace = trace;
  KUPSMI_REQUIRE(!prop_cdf || prop_nv >= 1, "a proposal table needs its voxel count");
  KUPSMI_REQUIRE(!sample_count || (prop_nv >= 1 && sample_row && sample_y),
                 "insertion samples need the voxel grid and output buffers");
  a.prop_cdf = prop_cdf; a.prop_nv = prop_nv; a.sample_count = sample_count;
  a.sample_cap = sample_cap; a.sample_row = sample_row; a.sample_y = sample_y;
  hipStream_t st = (hipStream_t)stream;
  int dev = 0, ncu = 256;
  KUPSMI_HIP(hipGetDevice(&dev));
  KUPSMI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid = n_chain < 2 * ncu ? n_chain : 2 * ncu;
  gcmc_kernel<<<grid, GC_THREADS, lds, st>>>(a);
  KUPSMI_CHECK_LAUNCH();
  return KUPSMI_OK;
}
