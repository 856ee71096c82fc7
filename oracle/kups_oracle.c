/*
 * kups_oracle.c -- FP64 CPU restatement of kUPS's batched particle-interaction
 * hot path.  TEST INFRASTRUCTURE ONLY: this file is the parity checker for the
 * HIP library (kups_amd/csrc) and the CPU baseline timed next to it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * Nothing in the product path (kups_amd) may call it.
 *
 * Every function cites the reference file:line whose algorithm it restates
 * (paths relative to the kUPS source root, src/kups/...).
 *
 * Conventions shared with the reference:
 *   - unit cell: lower-triangular lattice L with rows a,b,c, stored as
 *     tril = [L00, L10, L11, L20, L21, L22] (core/unitcell.py:251-300);
 *     Cartesian r = f @ L, fractional f = r @ Linv
 *     (core/utils/math.py:177-239, side="right").
 *   - an edge (i, j, s) has difference vector d = r_j - r_i + s @ L
 *     (core/neighborlist.py:164-184).
 *   - energy unit eV, length Angstrom, mass amu; Coulomb prefactor
 *     K = HARTREE*BOHR (potential/classical/ewald.py:82).
 *
 * Parity pins (what checks each function, see tests/test_oracle.py):
 *   orc_neighborlist  -- reference test cases of test/core/test_neighborlist.py
 *                        (exact cutoff, PBC, self-images, minimum-image
 *                        exclusion, small/triclinic cells) + brute force over
 *                        an explicit image box.
 *   orc_lj            -- closed form 4eps((s/r)^12-(s/r)^6), finite
 *                        differences, Newton's third law.
 *   orc_morse_bonds   -- closed form, dissociation limit, finite differences.
 *   orc_ewald_*       -- NaCl Madelung constant, two-charge identities,
 *                        alpha independence of the total, finite differences.
 *   orc_md_*          -- exact one-step formulas, VV energy drift,
 *                        time reversibility.
 *   orc_gcmc_*        -- exact log-ratio formulas, incremental == full.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* HARTREE*BOHR in eV*Angstrom from the ASE/CODATA-2014 constants that
 * core/constants.py:26-55 uses: e/(4 pi eps0) * 1e10. */
static double coulomb_k(void) {
  const double c = 299792458.0, mu0 = 4e-7 * M_PI;
  const double e = 1.6021766208e-19;
  const double eps0 = 1.0 / (mu0 * c * c);
  const double hbar = 6.626070040e-34 / (2 * M_PI);
  const double me = 9.10938356e-31;
  const double bohr = 4e10 * M_PI * eps0 * hbar * hbar / me / e / e;
  const double hartree = me * e * e * e / 16 / M_PI / M_PI / eps0 / eps0 / hbar / hbar;
  return hartree * bohr;
}
double orc_coulomb_k(void) { return coulomb_k(); }

/* ---------------------------------------------------------------- cells -- */
typedef struct {
  double L[3][3];    /* rows = lattice vectors */
  double Li[3][3];   /* lower-triangular inverse */
  double perp[3];    /* perpendicular widths */
  double vol;
} cell_t;

/* core/unitcell.py:275-300 (lattice from tril), core/utils/math.py:135-167
 * (triangular inverse), core/unitcell.py:174-195 (perpendicular lengths). */
static void cell_from_tril(const double* t, cell_t* c) {
  memset(c, 0, sizeof(*c));
  c->L[0][0] = t[0];
  c->L[1][0] = t[1]; c->L[1][1] = t[2];
  c->L[2][0] = t[3]; c->L[2][1] = t[4]; c->L[2][2] = t[5];
  double a = t[0], b = t[1], cc = t[2], d = t[3], e = t[4], f = t[5];
  c->Li[0][0] = 1.0 / a;
  c->Li[1][0] = -b / (a * cc); c->Li[1][1] = 1.0 / cc;
  c->Li[2][0] = (b * e - cc * d) / (a * cc * f); c->Li[2][1] = -e / (cc * f); c->Li[2][2] = 1.0 / f;
  c->vol = fabs(a * cc * f);
  const double* A = c->L[0]; const double* B = c->L[1]; const double* C = c->L[2];
  double bc[3] = {B[1]*C[2]-B[2]*C[1], B[2]*C[0]-B[0]*C[2], B[0]*C[1]-B[1]*C[0]};
  double ac[3] = {A[1]*C[2]-A[2]*C[1], A[2]*C[0]-A[0]*C[2], A[0]*C[1]-A[1]*C[0]};
  double ab[3] = {A[1]*B[2]-A[2]*B[1], A[2]*B[0]-A[0]*B[2], A[0]*B[1]-A[1]*B[0]};
  c->perp[0] = c->vol / sqrt(bc[0]*bc[0]+bc[1]*bc[1]+bc[2]*bc[2]);
  c->perp[1] = c->vol / sqrt(ac[0]*ac[0]+ac[1]*ac[1]+ac[2]*ac[2]);
  c->perp[2] = c->vol / sqrt(ab[0]*ab[0]+ab[1]*ab[1]+ab[2]*ab[2]);
}

/* x @ M for a lower-triangular 3x3 (core/utils/math.py:177-239, right side) */
static inline void rmul(const double M[3][3], const double* x, double* y) {
  y[0] = x[0]*M[0][0] + x[1]*M[1][0] + x[2]*M[2][0];
  y[1] = x[1]*M[1][1] + x[2]*M[2][1];
  y[2] = x[2]*M[2][2];
}

/* Image count per axis (core/neighborlist.py:455-459):
 * ceil(2 rc / perp), forced odd; 1 for an infinite cutoff. */
static void image_counts(const cell_t* c, double rc, int img[3]) {
  for (int a = 0; a < 3; ++a) {
    int n = isfinite(rc) ? (int)ceil(2.0 * rc / c->perp[a]) : 1;
    if (n < 1) n = 1;
    if (n % 2 == 0) n += 1;
    img[a] = n;
  }
}

/* ------------------------------------------------------- neighbor list -- */
/*
 * Full directed neighbor list of every system (rh = lh, no remap), the set
 * CellListNeighborList / DenseNearestNeighborList return
 * (core/neighborlist.py:743-786 basic_neighborlist, 692-740 filtering):
 *   - candidates: all ordered pairs (i, j) of the same system, i == j allowed;
 *   - if any axis of the system needs images (image_counts > 1 somewhere), each
 *     pair is replicated over the centred offset grid and the shift is the grid
 *     offset itself (595-624); otherwise shift = round_half_even(f_i - f_j)
 *     (542-564, jnp.round);
 *   - keep d^2 < rc^2 (strict, 589/621) where d = (f_i - f_j - s) @ L;
 *   - drop pairs with equal exclusion ids on the minimum image only (731-735);
 *     is_minimum_interaction = (s == round(f_i - f_j)) for replicated systems.
 * Positions are used as given (no wrapping), like the reference.
 * Writes at most `cap` edges (i, j, s0, s1, s2) in (i-major, j, offset) order
 * and returns the total count.  `sys_off` is the system-offset table.
 */
long orc_neighborlist(const double* pos, const int32_t* excl, int nsys,
                      const int32_t* sys_off, const double* tril,
                      const double* cutoff, long cap, int32_t* out_ij,
                      int32_t* out_shift) {
  long count = 0;
  for (int s = 0; s < nsys; ++s) {
    cell_t c; cell_from_tril(tril + 6 * s, &c);
    double rc = cutoff[s];
    int img[3]; image_counts(&c, rc, img);
    int replicated = img[0] * img[1] * img[2] > 1;
    int h[3] = {(img[0]-1)/2, (img[1]-1)/2, (img[2]-1)/2};
    int i0 = sys_off[s], i1 = sys_off[s+1];
    for (int i = i0; i < i1; ++i) {
      double fi[3]; rmul(c.Li, pos + 3*i, fi);
      for (int j = i0; j < i1; ++j) {
        double fj[3]; rmul(c.Li, pos + 3*j, fj);
        double df[3] = {fi[0]-fj[0], fi[1]-fj[1], fi[2]-fj[2]};
        double mins[3] = {rint(df[0]), rint(df[1]), rint(df[2])};
        int same_excl = excl[i] == excl[j];
        if (!replicated) {
          double dd[3] = {df[0]-mins[0], df[1]-mins[1], df[2]-mins[2]}, d[3];
          rmul(c.L, dd, d);
          double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
          if (r2 < rc*rc && !same_excl) {
            if (count < cap) {
              out_ij[2*count] = i; out_ij[2*count+1] = j;
              for (int a = 0; a < 3; ++a) out_shift[3*count+a] = (int32_t)mins[a];
            }
            ++count;
          }
          continue;
        }
        for (int o0 = -h[0]; o0 <= h[0]; ++o0)
          for (int o1 = -h[1]; o1 <= h[1]; ++o1)
            for (int o2 = -h[2]; o2 <= h[2]; ++o2) {
              double dd[3] = {df[0]-o0, df[1]-o1, df[2]-o2}, d[3];
              rmul(c.L, dd, d);
              double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
              int is_min = (o0 == mins[0]) && (o1 == mins[1]) && (o2 == mins[2]);
              if (r2 < rc*rc && !(same_excl && is_min)) {
                if (count < cap) {
                  out_ij[2*count] = i; out_ij[2*count+1] = j;
                  out_shift[3*count] = o0; out_shift[3*count+1] = o1; out_shift[3*count+2] = o2;
                }
                ++count;
              }
            }
      }
    }
  }
  return count;
}

/*
 * rh-subset query (NearestNeighborList.__call__(lh, rh, systems, cutoffs,
 * rh_index_remap), core/neighborlist.py:233-255, through basic_neighborlist,
 * 743-786): candidates (i in lh, r in rh) of the same system
 * (_build_segmentation_mask, 505-539; an rh system outside [0, nsys) drops
 * r); with a remap, an lh particle that any remap entry names is paired only
 * with entries whose remapped index is <= i (536-537: lh_idx >= rh_idx_out);
 * shift = round(f_i - f_r) or the centred image grid of a replicated system
 * (542-564, 595-624); kept when |(f_i - f_r - s) @ L|^2 < rc^2 and not
 * (equal exclusion ids on the minimum image) (731-735); output (i, out_r, s)
 * with out_r = remap[r] (n_lh for a remap outside [0, n_lh)) or r, and with
 * a remap every edge again reversed, (out_r, i, -s), after all forward edges
 * (_compact_edges, 680-685).  Forward edges in (r, i, offset) order.  Returns
 * the total count (both directions); writes at most cap.
 */
long orc_neighborlist_query(const double* pos, const int32_t* excl, int nsys,
                            const int32_t* sys_off, int k, const double* rh_pos,
                            const int32_t* rh_sys, const int32_t* rh_excl,
                            const int32_t* remap, const double* tril, const double* cutoff,
                            long cap, int32_t* out_ij, int32_t* out_shift) {
  const int n = sys_off[nsys];
  char* named = (char*)calloc(n > 0 ? n : 1, 1);
  if (remap)
    for (int r = 0; r < k; ++r)
      if (remap[r] >= 0 && remap[r] < n) named[remap[r]] = 1;
  /* forward edges first (count), then written with their reverses */
  long fwd = 0;
  for (int pass = 0; pass < 2; ++pass) {
    long e = 0;
    for (int r = 0; r < k; ++r) {
      const int s = rh_sys[r];
      if (s < 0 || s >= nsys) continue;
      cell_t c; cell_from_tril(tril + 6 * s, &c);
      const double rc = cutoff[s];
      int img[3]; image_counts(&c, rc, img);
      const int replicated = img[0] * img[1] * img[2] > 1;
      const int h[3] = {(img[0]-1)/2, (img[1]-1)/2, (img[2]-1)/2};
      int out_r = r;
      if (remap) out_r = (remap[r] >= 0 && remap[r] < n) ? remap[r] : n;
      double fr[3]; rmul(c.Li, rh_pos + 3*r, fr);
      for (int i = sys_off[s]; i < sys_off[s+1]; ++i) {
        if (remap && named[i] && !(i >= out_r)) continue;
        double fi[3]; rmul(c.Li, pos + 3*i, fi);
        double df[3] = {fi[0]-fr[0], fi[1]-fr[1], fi[2]-fr[2]};
        double mins[3] = {rint(df[0]), rint(df[1]), rint(df[2])};
        const int same = excl[i] == rh_excl[r];
        for (int o0 = -h[0]; o0 <= h[0]; ++o0)
          for (int o1 = -h[1]; o1 <= h[1]; ++o1)
            for (int o2 = -h[2]; o2 <= h[2]; ++o2) {
              double sh[3] = {o0, o1, o2};
              if (!replicated) { sh[0] = mins[0]; sh[1] = mins[1]; sh[2] = mins[2]; }
              double dd[3] = {df[0]-sh[0], df[1]-sh[1], df[2]-sh[2]}, d[3];
              rmul(c.L, dd, d);
              const double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
              const int is_min = sh[0] == mins[0] && sh[1] == mins[1] && sh[2] == mins[2];
              if (r2 < rc*rc && !(same && is_min)) {
                if (pass == 1) {
                  if (e < cap) {
                    out_ij[2*e] = i; out_ij[2*e+1] = out_r;
                    for (int a = 0; a < 3; ++a) out_shift[3*e+a] = (int32_t)sh[a];
                  }
                  const long rv = fwd + e;
                  if (remap && rv < cap) {
                    out_ij[2*rv] = out_r; out_ij[2*rv+1] = i;
                    for (int a = 0; a < 3; ++a) out_shift[3*rv+a] = -(int32_t)sh[a];
                  }
                }
                ++e;
              }
            }
      }
    }
    fwd = e;
  }
  free(named);
  return remap ? 2 * fwd : fwd;
}

/* ------------------------------------------------------ pair potentials -- */
/* Edge difference vector d = r_j - r_i + s @ L (core/neighborlist.py:164-184). */
static inline void edge_vec(const double* pos, const cell_t* c, int i, int j,
                            const int32_t* s, double* d) {
  double sd[3] = {(double)s[0], (double)s[1], (double)s[2]}, sl[3];
  rmul(c->L, sd, sl);
  for (int a = 0; a < 3; ++a) d[a] = pos[3*j+a] - pos[3*i+a] + sl[a];
}

/*
 * Lennard-Jones over a directed edge list (potential/classical/lennard_jones.py:
 * 149-178): u = 4 eps (c6^2 - c6), c6 = (sigma^2/r^2)^3, hard cutoff r^2 < rc^2
 * without shift, per-system sum over directed edges / 2.  sigma/eps are
 * (nspec x nspec) mixed tables (Lorentz-Berthelot, lennard_jones.py:125-141).
 * Gradients: the reference differentiates with jax.vjp
 * (potential/common/energy.py:297-300); here analytically:
 *   F_i += 2 u'(r^2) d_ij per directed edge (i -> j), u' = du/d(r^2),
 *   virial W -= sum_directed u' d (x) d  (stress = W / V,
 *   observables/stress.py:69-82).
 * energy/forces/virial are accumulated (callers zero them).
 */
void orc_lj_edges(const double* pos, const int32_t* species, int nsys,
                  const int32_t* sys_off, const double* tril, const double* cutoff,
                  int nspec, const double* sigma, const double* eps, long nedges,
                  const int32_t* ij, const int32_t* shift, double* energy,
                  double* forces, double* virial) {
  int* sys_of = NULL;
  int natoms = sys_off[nsys];
  sys_of = (int*)malloc(sizeof(int) * (natoms > 0 ? natoms : 1));
  for (int s = 0; s < nsys; ++s)
    for (int i = sys_off[s]; i < sys_off[s+1]; ++i) sys_of[i] = s;
  cell_t* cells = (cell_t*)malloc(sizeof(cell_t) * (nsys > 0 ? nsys : 1));
  for (int s = 0; s < nsys; ++s) cell_from_tril(tril + 6*s, &cells[s]);
  for (long e = 0; e < nedges; ++e) {
    int i = ij[2*e], j = ij[2*e+1], s = sys_of[i];
    double d[3]; edge_vec(pos, &cells[s], i, j, shift + 3*e, d);
    double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
    if (!(r2 < cutoff[s]*cutoff[s])) continue;
    int a = species[i], b = species[j];
    double sg = sigma[a*nspec+b], ep = eps[a*nspec+b];
    double c6 = pow(sg*sg / r2, 3);
    double u = 4.0 * ep * (c6*c6 - c6);
    double du = 12.0 * ep * c6 * (1.0 - 2.0*c6) / r2;   /* du/d(r^2) */
    energy[s] += 0.5 * u;
    for (int k = 0; k < 3; ++k) forces[3*i+k] += 2.0 * du * d[k];
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) virial[9*s + 3*k + l] -= du * d[k] * d[l];
  }
  free(sys_of); free(cells);
}

/*
 * Morse over an explicit bond list (potential/classical/morse.py:142-168):
 * u = D (1 - exp(-a (r - r0)))^2 per listed edge, summed without /2.
 * Parameters indexed per species pair (nspec x nspec).
 */
void orc_morse_edges(const double* pos, const int32_t* species, int nsys,
                     const int32_t* sys_off, const double* tril, int nspec,
                     const double* D, const double* alpha, const double* r0,
                     long nedges, const int32_t* ij, const int32_t* shift,
                     double* energy, double* forces, double* virial) {
  int natoms = sys_off[nsys];
  int* sys_of = (int*)malloc(sizeof(int) * (natoms > 0 ? natoms : 1));
  for (int s = 0; s < nsys; ++s)
    for (int i = sys_off[s]; i < sys_off[s+1]; ++i) sys_of[i] = s;
  for (long e = 0; e < nedges; ++e) {
    int i = ij[2*e], j = ij[2*e+1], s = sys_of[i];
    cell_t c; cell_from_tril(tril + 6*s, &c);
    double d[3]; edge_vec(pos, &c, i, j, shift + 3*e, d);
    double r = sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2]);
    int p = species[i]*nspec + species[j];
    double ex = exp(-alpha[p] * (r - r0[p]));
    double u = D[p] * (1 - ex) * (1 - ex);
    double dudr = 2.0 * D[p] * (1 - ex) * alpha[p] * ex;
    energy[s] += u;
    /* E depends on r = |r_j - r_i + sL|: dE/dr_i = -dudr d/r */
    for (int k = 0; k < 3; ++k) {
      forces[3*i+k] += dudr * d[k] / r;
      forces[3*j+k] -= dudr * d[k] / r;
    }
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) virial[9*s+3*k+l] -= dudr * d[k] * d[l] / r;
  }
  free(sys_of);
}

/* ----------------------------------------------------------------- Ewald -- */
/*
 * Real-space Ewald sum over a directed edge list (potential/classical/ewald.py:
 * 305-325): E = K/2 sum_directed q_i q_j erfc(alpha r)/r [r < rc].
 * With vacuum != 0 the vacuum Coulomb energy K/2 sum q_i q_j / r is subtracted
 * instead of screened (the exclusion correction E_sr - E_vac,
 * ewald.py:328-346 + coulomb.py:64-83, over an intramolecular edge list;
 * `use_cutoff` = 0 there because the vacuum term has no cutoff mask but the
 * screened one does).
 */
void orc_ewald_real_edges(const double* pos, const double* q, int nsys,
                          const int32_t* sys_off, const double* tril,
                          const double* alpha, const double* cutoff, long nedges,
                          const int32_t* ij, const int32_t* shift, int mode,
                          double* energy, double* forces, double* virial) {
  /* mode 0: screened real-space sum; mode 1: exclusion correction
   * (screened - vacuum) */
  const double K = coulomb_k();
  int natoms = sys_off[nsys];
  int* sys_of = (int*)malloc(sizeof(int) * (natoms > 0 ? natoms : 1));
  for (int s = 0; s < nsys; ++s)
    for (int i = sys_off[s]; i < sys_off[s+1]; ++i) sys_of[i] = s;
  for (long e = 0; e < nedges; ++e) {
    int i = ij[2*e], j = ij[2*e+1], s = sys_of[i];
    cell_t c; cell_from_tril(tril + 6*s, &c);
    double d[3]; edge_vec(pos, &c, i, j, shift + 3*e, d);
    double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
    double r = sqrt(r2);
    double a = alpha[s];
    double qq = q[i] * q[j];
    double u = 0.0, dudr = 0.0;
    if (r < cutoff[s]) {
      u = qq * erfc(a*r) / r;
      dudr = qq * (-2.0*a/sqrt(M_PI) * exp(-a*a*r2) / r - erfc(a*r) / r2);
    }
    if (mode == 1) {
      u -= qq / r;
      dudr += qq / r2;
    }
    energy[s] += 0.5 * K * u;
    /* per directed edge: dE/dr_j = K/2 dudr d/r; the reverse edge gives the
     * same for i, so F_i = sum_directed K dudr d / r */
    for (int k = 0; k < 3; ++k) forces[3*i+k] += K * dudr * d[k] / r;
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) virial[9*s+3*k+l] -= 0.5 * K * dudr * d[k] * d[l] / r;
  }
  free(sys_of);
}

/*
 * Reciprocal-space Ewald sum (ewald.py:349-439, 585-601):
 *   k = 2 pi Linv n  (ewald.py:265-274),
 *   P(k) = (2 pi / V) exp(-k^2 / 4 alpha^2) / k^2 * (2 - [n_x == 0]), P(0) = 0
 *   (ewald.py:363-391),
 *   S(k) = sum_i q_i (cos k.r_i, sin k.r_i),  E = K sum_k P(k) |S(k)|^2.
 * Forces F_i = 2 K q_i sum_k P(k) (S_c sin(k.r_i) - S_s cos(k.r_i)) k.
 * Virial (reciprocal-space Ewald stress, standard form):
 *   W_ab = K sum_k P(k) |S|^2 (delta_ab - 2 k_a k_b (1/k^2 + 1/(4 alpha^2))).
 * kvec_n holds nk integer triples per system (row-major, nsys x nk x 3);
 * structure factors are written to sk (nsys x nk x 2) when non-NULL.
 */
void orc_ewald_recip(const double* pos, const double* q, int nsys,
                     const int32_t* sys_off, const double* tril,
                     const double* alpha, int nk, const int32_t* kvec_n,
                     double* energy, double* forces, double* virial, double* sk) {
  const double K = coulomb_k();
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < nsys; ++s) {
    cell_t c; cell_from_tril(tril + 6*s, &c);
    double a = alpha[s];
    for (int kk = 0; kk < nk; ++kk) {
      const int32_t* n = kvec_n + 3*((long)s*nk + kk);
      double kv[3];
      for (int x = 0; x < 3; ++x)
        kv[x] = 2*M_PI*(c.Li[x][0]*n[0] + c.Li[x][1]*n[1] + c.Li[x][2]*n[2]);
      double k2 = kv[0]*kv[0] + kv[1]*kv[1] + kv[2]*kv[2];
      double Sc = 0, Ss = 0;
      for (int i = sys_off[s]; i < sys_off[s+1]; ++i) {
        double th = kv[0]*pos[3*i] + kv[1]*pos[3*i+1] + kv[2]*pos[3*i+2];
        Sc += q[i] * cos(th); Ss += q[i] * sin(th);
      }
      if (sk) { sk[2*((long)s*nk+kk)] = Sc; sk[2*((long)s*nk+kk)+1] = Ss; }
      if (!(k2 > 0)) continue;
      double P = 2*M_PI / c.vol * exp(-k2 / (4*a*a)) / k2 * (n[0] == 0 ? 1.0 : 2.0);
      double S2 = Sc*Sc + Ss*Ss;
      energy[s] += K * P * S2;
      for (int i = sys_off[s]; i < sys_off[s+1]; ++i) {
        double th = kv[0]*pos[3*i] + kv[1]*pos[3*i+1] + kv[2]*pos[3*i+2];
        double g = 2*K*q[i]*P*(Sc*sin(th) - Ss*cos(th));
        for (int x = 0; x < 3; ++x) forces[3*i+x] += g * kv[x];
      }
      double fac = 2.0 * (1.0/k2 + 1.0/(4*a*a));
      for (int x = 0; x < 3; ++x)
        for (int y = 0; y < 3; ++y)
          virial[9*s+3*x+y] += K * P * S2 * ((x == y) - fac * kv[x] * kv[y]);
    }
  }
}

/* Ewald self term (ewald.py:277-302): E = -K alpha/sqrt(pi) sum q^2.
 * No neutralising-background term exists in the reference. */
void orc_ewald_self(const double* q, int nsys, const int32_t* sys_off,
                    const double* alpha, double* energy) {
  const double K = coulomb_k();
  for (int s = 0; s < nsys; ++s) {
    double q2 = 0;
    for (int i = sys_off[s]; i < sys_off[s+1]; ++i) q2 += q[i]*q[i];
    energy[s] += -K * alpha[s] / sqrt(M_PI) * q2;
  }
}

/* Half-space integer k grid n in [0..nx] x [-ny..ny] x [-nz..nz] filtered
 * to |k| <= kmax (kmax < 0 keeps the whole box), the recipe of
 * kvecs_from_kmax (ewald.py:1322-1339) with the box bound from its svd.
 * Returns the count; writes up to cap triples. */
long orc_kvecs(const double* tril, int nmax, double kmax, long cap, int32_t* out) {
  cell_t c; cell_from_tril(tril, &c);
  long cnt = 0;
  for (int n0 = 0; n0 <= nmax; ++n0)
    for (int n1 = -nmax; n1 <= nmax; ++n1)
      for (int n2 = -nmax; n2 <= nmax; ++n2) {
        double kv[3];
        int n[3] = {n0, n1, n2};
        for (int x = 0; x < 3; ++x)
          kv[x] = 2*M_PI*(c.Li[x][0]*n[0] + c.Li[x][1]*n[1] + c.Li[x][2]*n[2]);
        double kn = sqrt(kv[0]*kv[0] + kv[1]*kv[1] + kv[2]*kv[2]);
        if (kmax >= 0 && !(kn <= kmax)) continue;
        if (cnt < cap) { out[3*cnt] = n0; out[3*cnt+1] = n1; out[3*cnt+2] = n2; }
        ++cnt;
      }
  return cnt;
}

/* ------------------------------------------------- cell-list LJ forces -- */
/*
 * LJ energy/forces/virial of a batch of systems by a cell list, the same pair
 * set as orc_neighborlist + orc_lj_edges (used by the MD baseline where
 * building the full edge list would dominate).  Requires every system to be
 * non-replicated (all perpendicular widths >= 2 rc); callers check with
 * orc_needs_images and fall back to the edge-list path otherwise.  Exclusion:
 * i != j only (particle-id exclusion, the MD apps' setting).
 * Parallel over systems (OpenMP).
 */
int orc_needs_images(const double* tril, double rc) {
  cell_t c; cell_from_tril(tril, &c);
  int img[3]; image_counts(&c, rc, img);
  return img[0]*img[1]*img[2] > 1;
}

static void lj_system_dense(const double* pos, const int32_t* species, int i0, int i1,
                            const cell_t* c, double rc, int nspec, const double* sigma,
                            const double* eps, double* e_out, double* f, double* w) {
  /* replicated system: explicit offset grid, exactly the orc_neighborlist set */
  int img[3]; image_counts(c, rc, img);
  int h[3] = {(img[0]-1)/2, (img[1]-1)/2, (img[2]-1)/2};
  int rep = img[0]*img[1]*img[2] > 1;
  double e = 0;
  for (int i = i0; i < i1; ++i) {
    double fi[3]; rmul(c->Li, pos+3*i, fi);
    for (int j = i0; j < i1; ++j) {
      double fj[3]; rmul(c->Li, pos+3*j, fj);
      double df[3] = {fi[0]-fj[0], fi[1]-fj[1], fi[2]-fj[2]};
      double mins[3] = {rint(df[0]), rint(df[1]), rint(df[2])};
      int lo[3], hi[3];
      for (int a = 0; a < 3; ++a) {
        if (rep) { lo[a] = -h[a]; hi[a] = h[a]; } else { lo[a] = hi[a] = (int)mins[a]; }
      }
      for (int o0 = lo[0]; o0 <= hi[0]; ++o0)
        for (int o1 = lo[1]; o1 <= hi[1]; ++o1)
          for (int o2 = lo[2]; o2 <= hi[2]; ++o2) {
            int is_min = o0 == mins[0] && o1 == mins[1] && o2 == mins[2];
            if (i == j && is_min) continue;
            double dd[3] = {df[0]-o0, df[1]-o1, df[2]-o2}, d[3];
            rmul(c->L, dd, d);
            /* d here is r_i - r_j - sL = -(edge vector) */
            double r2 = d[0]*d[0]+d[1]*d[1]+d[2]*d[2];
            if (!(r2 < rc*rc)) continue;
            int p = species[i]*nspec + species[j];
            double c6 = pow(sigma[p]*sigma[p]/r2, 3);
            e += 0.5 * 4.0*eps[p]*(c6*c6 - c6);
            double du = 12.0*eps[p]*c6*(1.0-2.0*c6)/r2;
            for (int k = 0; k < 3; ++k) f[3*i+k] -= 2.0*du*d[k];
            for (int k = 0; k < 3; ++k)
              for (int l = 0; l < 3; ++l) w[3*k+l] -= du*d[k]*d[l];
          }
    }
  }
  *e_out += e;
}

static void lj_system_cells(const double* pos, const int32_t* species, int i0, int i1,
                            const cell_t* c, double rc, int nspec, const double* sigma,
                            const double* eps, double* e_out, double* f, double* w) {
  int n = i1 - i0;
  int nc[3];
  for (int a = 0; a < 3; ++a) { nc[a] = (int)(c->perp[a] / rc); if (nc[a] < 1) nc[a] = 1; }
  int ncell = nc[0]*nc[1]*nc[2];
  int* cid = (int*)malloc(sizeof(int)*(n>0?n:1));
  int* start = (int*)calloc(ncell+1, sizeof(int));
  int* order = (int*)malloc(sizeof(int)*(n>0?n:1));
  double* fr = (double*)malloc(sizeof(double)*3*(n>0?n:1));
  for (int k = 0; k < n; ++k) {
    double fk[3]; rmul(c->Li, pos+3*(i0+k), fk);
    int cc[3];
    for (int a = 0; a < 3; ++a) {
      double u = fk[a] - floor(fk[a]);   /* wrap to [0,1) */
      fr[3*k+a] = fk[a];
      int b = (int)(u * nc[a]); if (b >= nc[a]) b = nc[a]-1; if (b < 0) b = 0;
      cc[a] = b;
    }
    cid[k] = (cc[2]*nc[1] + cc[1])*nc[0] + cc[0];
    start[cid[k]+1]++;
  }
  for (int k = 0; k < ncell; ++k) start[k+1] += start[k];
  int* cur = (int*)malloc(sizeof(int)*(ncell>0?ncell:1));
  memcpy(cur, start, sizeof(int)*ncell);
  for (int k = 0; k < n; ++k) order[cur[cid[k]]++] = k;
  double e = 0;
  for (int k = 0; k < n; ++k) {
    int cx = cid[k] % nc[0], cy = (cid[k]/nc[0]) % nc[1], cz = cid[k]/(nc[0]*nc[1]);
    int visited[27], nv = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int x = (cx+dx+nc[0])%nc[0], y = (cy+dy+nc[1])%nc[1], z = (cz+dz+nc[2])%nc[2];
          int nb = (z*nc[1]+y)*nc[0]+x, dup = 0;
          for (int v = 0; v < nv; ++v) dup |= visited[v] == nb;
          if (dup) continue;
          visited[nv++] = nb;
          for (int t = start[nb]; t < start[nb+1]; ++t) {
            int m = order[t];
            if (m == k) continue;
            double df[3];
            for (int a = 0; a < 3; ++a) { df[a] = fr[3*k+a]-fr[3*m+a]; df[a] -= rint(df[a]); }
            double d[3]; rmul(c->L, df, d);
            double r2 = d[0]*d[0]+d[1]*d[1]+d[2]*d[2];
            if (!(r2 < rc*rc)) continue;
            int p = species[i0+k]*nspec + species[i0+m];
            double c6 = sigma[p]*sigma[p]/r2; c6 = c6*c6*c6;
            e += 2.0*eps[p]*(c6*c6 - c6);
            double du = 12.0*eps[p]*c6*(1.0-2.0*c6)/r2;
            for (int a = 0; a < 3; ++a) f[3*(i0+k)+a] -= 2.0*du*d[a];
            for (int a = 0; a < 3; ++a)
              for (int b = 0; b < 3; ++b) w[3*a+b] -= du*d[a]*d[b];
          }
        }
  }
  *e_out += e;
  free(cid); free(start); free(order); free(fr); free(cur);
}

void orc_lj_cells(const double* pos, const int32_t* species, int nsys,
                  const int32_t* sys_off, const double* tril, const double* cutoff,
                  int nspec, const double* sigma, const double* eps,
                  double* energy, double* forces, double* virial) {
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < nsys; ++s) {
    cell_t c; cell_from_tril(tril + 6*s, &c);
    int i0 = sys_off[s], i1 = sys_off[s+1];
    for (int i = i0; i < i1; ++i) forces[3*i] = forces[3*i+1] = forces[3*i+2] = 0;
    double e = 0, w[9] = {0};
    int img[3]; image_counts(&c, cutoff[s], img);
    if (img[0]*img[1]*img[2] > 1)
      lj_system_dense(pos, species, i0, i1, &c, cutoff[s], nspec, sigma, eps, &e, forces, w);
    else
      lj_system_cells(pos, species, i0, i1, &c, cutoff[s], nspec, sigma, eps, &e, forces, w);
    energy[s] = e;
    for (int k = 0; k < 9; ++k) virial[9*s+k] = w[k];
  }
}

/* ---------------------------------------------------------- integrators -- */
/* Wrap Cartesian r into the primary cell [-0.5, 0.5) in fractional space
 * (core/unitcell.py:139-171): f = r @ Linv; f = (f + 0.5) mod 1 - 0.5; r = f @ L. */
static void wrap_pos(const cell_t* c, double* r) {
  double f[3]; rmul(c->Li, r, f);
  for (int a = 0; a < 3; ++a) { double u = f[a] + 0.5; f[a] = (u - floor(u)) - 0.5; }
  rmul(c->L, f, r);
}

/*
 * One MD step of every system, FP64.  integrator 0 = velocity Verlet
 * (md/integrators.py:273-317: B(dt/2) A(dt) F B(dt/2)), 1 = BAOAB Langevin
 * (405-453: B(dt/2) A(dt/2) O(dt) A(dt/2) F B(dt/2)); A steps wrap positions
 * (104-141).  O: p <- c1 p + sqrt(kT (1 - c1^2)) sqrt(m) eta, c1 = exp(-gamma dt)
 * (330-402) with eta supplied by the caller (n x 3 normals) so the CPU and GPU
 * runs can share one noise stream.  `forces` holds F(t) on entry and F(t+dt)
 * on exit; energy/virial receive the values at t+dt.  dt, gamma in internal
 * units (md/data.py:195-205), kT in eV.
 */
void orc_md_step(int integrator, double* pos, double* mom, const double* mass,
                 const int32_t* species, int nsys, const int32_t* sys_off,
                 const double* tril, const double* cutoff, int nspec,
                 const double* sigma, const double* eps, const double* dt,
                 const double* kT, const double* gamma, const double* eta,
                 double* forces, double* energy, double* virial) {
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < nsys; ++s) {
    cell_t c; cell_from_tril(tril + 6*s, &c);
    double h = 0.5 * dt[s];
    double c1 = exp(-gamma[s] * dt[s]);
    double c2 = sqrt(kT[s] * (1.0 - c1*c1));
    for (int i = sys_off[s]; i < sys_off[s+1]; ++i) {
      for (int a = 0; a < 3; ++a) mom[3*i+a] += forces[3*i+a] * h;
      if (integrator == 0) {
        for (int a = 0; a < 3; ++a) pos[3*i+a] += mom[3*i+a] / mass[i] * dt[s];
        wrap_pos(&c, pos + 3*i);
      } else {
        for (int a = 0; a < 3; ++a) pos[3*i+a] += mom[3*i+a] / mass[i] * h;
        wrap_pos(&c, pos + 3*i);
        for (int a = 0; a < 3; ++a)
          mom[3*i+a] = c1 * mom[3*i+a] + c2 * sqrt(mass[i]) * eta[3*i+a];
        for (int a = 0; a < 3; ++a) pos[3*i+a] += mom[3*i+a] / mass[i] * h;
        wrap_pos(&c, pos + 3*i);
      }
    }
  }
  orc_lj_cells(pos, species, nsys, sys_off, tril, cutoff, nspec, sigma, eps, energy, forces, virial);
#pragma omp parallel for schedule(static)
  for (int s = 0; s < nsys; ++s) {
    double h = 0.5 * dt[s];
    for (int i = sys_off[s]; i < sys_off[s+1]; ++i)
      for (int a = 0; a < 3; ++a) mom[3*i+a] += forces[3*i+a] * h;
  }
}

/* FP32 variant of the velocity-Verlet LJ step (the precision the GPU runs
 * at), used for the FP32 CPU baseline: same cell list and step order. */
static void lj_system_cells_f32(float* pos, const int32_t* species, int i0, int i1,
                                const cell_t* c, float rc, int nspec, const float* sigma,
                                const float* eps, double* e_out, float* f, int* scratch_i,
                                float* scratch_f) {
  int n = i1 - i0;
  int nc[3];
  for (int a = 0; a < 3; ++a) { nc[a] = (int)(c->perp[a] / rc); if (nc[a] < 1) nc[a] = 1; }
  int ncell = nc[0]*nc[1]*nc[2];
  int* cid = scratch_i; int* order = scratch_i + n; int* start = scratch_i + 2*n;
  for (int k = 0; k <= ncell; ++k) start[k] = 0;
  float Li[3][3], L[3][3];
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) { Li[a][b] = (float)c->Li[a][b]; L[a][b] = (float)c->L[a][b]; }
  float* fr = scratch_f;
  for (int k = 0; k < n; ++k) {
    const float* r = pos + 3*(i0+k);
    float fk[3] = {r[0]*Li[0][0] + r[1]*Li[1][0] + r[2]*Li[2][0], r[1]*Li[1][1] + r[2]*Li[2][1], r[2]*Li[2][2]};
    int cc[3];
    for (int a = 0; a < 3; ++a) {
      fr[3*k+a] = fk[a];
      float u = fk[a] - floorf(fk[a]);
      int b = (int)(u * nc[a]); if (b >= nc[a]) b = nc[a]-1; if (b < 0) b = 0;
      cc[a] = b;
    }
    cid[k] = (cc[2]*nc[1] + cc[1])*nc[0] + cc[0];
    start[cid[k]+1]++;
  }
  for (int k = 0; k < ncell; ++k) start[k+1] += start[k];
  int* cur = start + ncell + 1;
  for (int k = 0; k < ncell; ++k) cur[k] = start[k];
  for (int k = 0; k < n; ++k) order[cur[cid[k]]++] = k;
  double e = 0;
  float rc2 = rc*rc;
  for (int k = 0; k < n; ++k) {
    int cx = cid[k] % nc[0], cy = (cid[k]/nc[0]) % nc[1], cz = cid[k]/(nc[0]*nc[1]);
    int visited[27], nv = 0;
    float fx = 0, fy = 0, fz = 0; float ek = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          int x = (cx+dx+nc[0])%nc[0], y = (cy+dy+nc[1])%nc[1], z = (cz+dz+nc[2])%nc[2];
          int nb = (z*nc[1]+y)*nc[0]+x, dup = 0;
          for (int v = 0; v < nv; ++v) dup |= visited[v] == nb;
          if (dup) continue;
          visited[nv++] = nb;
          for (int t = start[nb]; t < start[nb+1]; ++t) {
            int m = order[t];
            if (m == k) continue;
            float df[3];
            for (int a = 0; a < 3; ++a) { df[a] = fr[3*k+a]-fr[3*m+a]; df[a] -= rintf(df[a]); }
            float d0 = df[0]*L[0][0] + df[1]*L[1][0] + df[2]*L[2][0];
            float d1 = df[1]*L[1][1] + df[2]*L[2][1];
            float d2 = df[2]*L[2][2];
            float r2 = d0*d0+d1*d1+d2*d2;
            if (!(r2 < rc2)) continue;
            int p = species[i0+k]*nspec + species[i0+m];
            float c6 = sigma[p]*sigma[p]/r2; c6 = c6*c6*c6;
            ek += 2.0f*eps[p]*(c6*c6 - c6);
            float du = 12.0f*eps[p]*c6*(1.0f-2.0f*c6)/r2;
            fx -= 2.0f*du*d0; fy -= 2.0f*du*d1; fz -= 2.0f*du*d2;
          }
        }
    f[3*(i0+k)] = fx; f[3*(i0+k)+1] = fy; f[3*(i0+k)+2] = fz;
    e += ek;
  }
  *e_out = e;
}

/*
 * n_steps velocity-Verlet LJ steps of every system in FP32 (positions,
 * momenta, forces) with FP64 per-system energies, cell list rebuilt every
 * step, OpenMP over systems.  This is the FP32 CPU baseline of the batched MD
 * workload; only systems that need no periodic images are supported (returns
 * -1 otherwise).  `forces` carries F between calls.
 */
int orc_md_vv_f32(int n_steps, float* pos, float* mom, const float* mass,
                  const int32_t* species, int nsys, const int32_t* sys_off,
                  const double* tril, const double* cutoff, int nspec,
                  const float* sigma, const float* eps, const double* dt,
                  float* forces, double* energy) {
  for (int s = 0; s < nsys; ++s)
    if (orc_needs_images(tril + 6*s, cutoff[s])) return -1;
#pragma omp parallel for schedule(dynamic)
  for (int s = 0; s < nsys; ++s) {
    cell_t c; cell_from_tril(tril + 6*s, &c);
    int i0 = sys_off[s], i1 = sys_off[s+1], n = i1 - i0;
    int nc[3];
    for (int a = 0; a < 3; ++a) { nc[a] = (int)(c.perp[a] / cutoff[s]); if (nc[a] < 1) nc[a] = 1; }
    int ncell = nc[0]*nc[1]*nc[2];
    int* si = (int*)malloc(sizeof(int)*(2*n + 2*ncell + 2));
    float* sf = (float*)malloc(sizeof(float)*(3*n + 3));
    float Lf[3][3], Lif[3][3];
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) { Lf[a][b] = (float)c.L[a][b]; Lif[a][b] = (float)c.Li[a][b]; }
    float dtf = (float)dt[s], hf = 0.5f * (float)dt[s];
    for (int st = 0; st < n_steps; ++st) {
      for (int i = i0; i < i1; ++i) {
        float* r = pos + 3*i; float* p = mom + 3*i;
        for (int a = 0; a < 3; ++a) { p[a] += forces[3*i+a] * hf; r[a] += p[a] / mass[i] * dtf; }
        float f0 = r[0]*Lif[0][0] + r[1]*Lif[1][0] + r[2]*Lif[2][0];
        float f1 = r[1]*Lif[1][1] + r[2]*Lif[2][1];
        float f2 = r[2]*Lif[2][2];
        f0 = (f0 + 0.5f) - floorf(f0 + 0.5f) - 0.5f;
        f1 = (f1 + 0.5f) - floorf(f1 + 0.5f) - 0.5f;
        f2 = (f2 + 0.5f) - floorf(f2 + 0.5f) - 0.5f;
        r[0] = f0*Lf[0][0] + f1*Lf[1][0] + f2*Lf[2][0];
        r[1] = f1*Lf[1][1] + f2*Lf[2][1];
        r[2] = f2*Lf[2][2];
      }
      double e = 0;
      lj_system_cells_f32(pos, species, i0, i1, &c, (float)cutoff[s], nspec, sigma, eps, &e, forces, si, sf);
      energy[s] = e;
      for (int i = i0; i < i1; ++i)
        for (int a = 0; a < 3; ++a) mom[3*i+a] += forces[3*i+a] * hf;
    }
    free(si); free(sf);
  }
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ================================================================ GCMC == */
/*
 * Rigid-molecule GCMC of a batch of chains (application/simulations/
 * mcmc_rigid.py:281-307 propagator; mcmc/moves.py:122-295, 573-835,
 * 1078-1230 moves; mcmc/probability.py:60-225 acceptance;
 * core/propagator.py:244-344 step; core/parameter_scheduler.py:161-208
 * step-width adaptation).  FP64 restatement of the GPU kernel's algorithm,
 * drawing the same counter-based random numbers, so accept sequences can be
 * compared trial by trial.
 *
 * Random numbers (shared with the GPU kernel; the reference draws from
 * jax.random, which cannot be reproduced bit for bit):
 *   batch draw  philox({0xFFFFFFFF, t, t>>32, TAG_BATCH}, seed): u32[0] picks
 *               the move type by the cumulative weights, u32[1] >> 31 the
 *               insert (0) / delete (1) coin (one move type and one coin for
 *               every chain, core/propagator.py:263-268, moves.py:1217);
 *   chain draw  philox({chain, t, t>>32, TAG_A}): [group bits, accept u,
 *               spare, spare]; TAG_B: quaternion uniforms; TAG_C: offsets or
 *               the translation normals (Box-Muller).
 * Float conversions of the uniforms are done in float, as on the device.
 */
#define GC_TAG_BATCH 0x42415443u
#define GC_TAG_A 0x47434d41u
#define GC_TAG_B 0x47434d42u
#define GC_TAG_C 0x47434d43u

static void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += W0; k1 += W1;
  }
}
static float gc_u01(uint32_t x) { return ((float)x + 0.5f) * 2.3283064365386963e-10f; }
static void gc_draw(uint64_t seed, uint32_t a, int64_t t, uint32_t tag, uint32_t out[4]) {
  out[0] = a; out[1] = (uint32_t)t; out[2] = (uint32_t)((uint64_t)t >> 32); out[3] = tag;
  philox10(out, (uint32_t)seed, (uint32_t)(seed >> 32));
}
void orc_philox(uint64_t seed, uint32_t a, int64_t t, uint32_t tag, uint32_t* out) {
  gc_draw(seed, a, t, tag, out);
}
/* The bare generator: out = Philox4x32-10(counter ctr[4], key key[2]). */
void orc_philox_raw(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  for (int k = 0; k < 4; ++k) out[k] = ctr[k];
  philox10(out, key[0], key[1]);
}

/* ------------------------------------------------------------- CSVR ---- */
/* Stochastic velocity rescaling (md/integrators.py:470-585, Bussi, Donadio
 * and Parrinello, J. Chem. Phys. 126, 014101 (2007)): per system
 *   alpha^2 = c1 + c2 (R1^2 + R2) + 2 R1 sqrt(c1 c2),
 *   c1 = exp(-dt / tau), c2 = (1 - c1) K_target / (K N_dof), K_target = N_dof kT / 2,
 *   N_dof = 3 N - 3, R1 ~ N(0, 1) ONE draw shared by every system of the batch
 *   (the reference draws a scalar, 552), R2 ~ chi^2(N_dof - 1) per system
 *   (0 when N_dof - 1 <= 0), p <- alpha p.
 * Random numbers (shared with the GPU kernel):
 *   R1: Box-Muller of words 0, 1 of philox({0xFFFFFFFF, t, t >> 32, CSVR_TAG_R1});
 *   R2: 2 Gamma((N_dof - 1) / 2) by Marsaglia and Tsang (ACM TOMS 26 (2000)
 *       363-372), attempt k drawing philox({system, t, t >> 32, CSVR_TAG_R2 + k}):
 *       x = Box-Muller(words 0, 1), u = word 2; shape < 1 boosts from shape + 1
 *       with U^(1/shape), U = word 0 of attempt 255.
 * Uniforms in FP64: u = (x + 0.5) 2^-32. */
#define CSVR_TAG_R1 0x43535631u
#define CSVR_TAG_R2 0x43535700u

static double csvr_u(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }
static double csvr_bm(uint32_t a, uint32_t b) {
  return sqrt(-2.0 * log(csvr_u(a))) * cos(2.0 * M_PI * csvr_u(b));
}

double orc_gamma_mt(double shape, uint64_t seed, uint32_t sys, int64_t t) {
  uint32_t c[4];
  double boost = 1.0;
  if (shape < 1.0) {
    gc_draw(seed, sys, t, CSVR_TAG_R2 + 255u, c);
    boost = pow(csvr_u(c[0]), 1.0 / shape);
    shape += 1.0;
  }
  const double d = shape - 1.0 / 3.0, cc = 1.0 / sqrt(9.0 * d);
  for (uint32_t k = 0; k < 255u; ++k) {
    gc_draw(seed, sys, t, CSVR_TAG_R2 + k, c);
    const double x = csvr_bm(c[0], c[1]);
    double v = 1.0 + cc * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    if (log(csvr_u(c[2])) < 0.5 * x * x + d - d * v + d * log(v)) return boost * d * v;
  }
  return boost * d;
}

double orc_csvr_r1(uint64_t seed, int64_t t) {
  uint32_t c[4];
  gc_draw(seed, 0xFFFFFFFFu, t, CSVR_TAG_R1, c);
  return csvr_bm(c[0], c[1]);
}

/* alpha of one system from its kinetic energy K (eV). */
double orc_csvr_alpha(double K, int n_atoms, double kT, double dt, double tau, double r1,
                      uint64_t seed, uint32_t sys, int64_t t) {
  const int ndof = 3 * n_atoms - 3;
  if (ndof <= 0 || !(K > 0.0)) return 1.0;
  const double r2 = ndof - 1 > 0 ? 2.0 * orc_gamma_mt(0.5 * (ndof - 1), seed, sys, t) : 0.0;
  const double c1 = exp(-dt / tau);
  const double c2 = (1.0 - c1) * (0.5 * ndof * kT) / (K * ndof);
  const double a2 = c1 + c2 * (r1 * r1 + r2) + 2.0 * r1 * sqrt(c1 * c2);
  return sqrt(a2 > 0.0 ? a2 : 0.0);
}

/* The CSVR step of every system at global step t: momenta rescaled in place,
 * K from the FP64 momenta; alpha[s] receives the factor. */
void orc_csvr_scale(uint64_t seed, int64_t t, int nsys, const int32_t* sys_off, const double* mass,
                    double* mom, const double* kT, const double* dt, const double* tau,
                    double* alpha) {
  const double r1 = orc_csvr_r1(seed, t);
  for (int s = 0; s < nsys; ++s) {
    double K = 0.0;
    for (int i = sys_off[s]; i < sys_off[s + 1]; ++i)
      for (int a = 0; a < 3; ++a) K += 0.5 * mom[3 * i + a] * mom[3 * i + a] / mass[i];
    const double al = orc_csvr_alpha(K, sys_off[s + 1] - sys_off[s], kT[s], dt[s], tau[s], r1,
                                     seed, (uint32_t)s, t);
    for (int i = sys_off[s]; i < sys_off[s + 1]; ++i)
      for (int a = 0; a < 3; ++a) mom[3 * i + a] *= al;
    if (alpha) alpha[s] = al;
  }
}

/* Stochastic cell rescaling (md/integrators.py:657-811; Bernetti and Bussi,
 * J. Chem. Phys. 153, 114107 (2020)) of every system at global step t:
 *   P = 2K / (3V) + tr(W) / (3V)   (W the virial, K from the momenta),
 *   d eps = (dt / tau_P) beta (P - P0) + sqrt(2 kT beta dt / (tau_P V)) R,
 *   mu = clip(exp(d eps / 3), mu_min, 1 / mu_min), L <- mu L, r <- mu r,
 * with ONE R ~ N(0, 1) for the batch: Box-Muller of words 0, 1 of
 * philox({0xFFFFFFFE, t, t >> 32, SCR_TAG}).  baro[4 s..] = P0, tau_P, beta,
 * mu_min; mu[s] receives the factor. */
#define SCR_TAG 0x53435231u

void orc_scr_scale(uint64_t seed, int64_t t, int nsys, const int32_t* sys_off, const double* mass,
                   const double* mom, double* pos, double* tril, const double* kT,
                   const double* dt, const double* baro, const double* virial, double* mu) {
  uint32_t c[4];
  gc_draw(seed, 0xFFFFFFFEu, t, SCR_TAG, c);
  const double R = csvr_bm(c[0], c[1]);
  for (int s = 0; s < nsys; ++s) {
    double K = 0.0;
    for (int i = sys_off[s]; i < sys_off[s + 1]; ++i)
      for (int a = 0; a < 3; ++a) K += 0.5 * mom[3 * i + a] * mom[3 * i + a] / mass[i];
    double* L = tril + 6 * s;
    const double* bp = baro + 4 * s;
    const double V = fabs(L[0] * L[2] * L[5]);
    const double* W = virial + 9 * s;
    const double P = 2.0 * K / (3.0 * V) + (W[0] + W[4] + W[8]) / (3.0 * V);
    const double de = dt[s] / bp[1] * bp[2] * (P - bp[0]) +
                      sqrt(2.0 * kT[s] * bp[2] * dt[s] / (bp[1] * V)) * R;
    double m = exp(de / 3.0);
    if (m < bp[3]) m = bp[3];
    if (m > 1.0 / bp[3]) m = 1.0 / bp[3];
    for (int k = 0; k < 6; ++k) L[k] *= m;
    for (int i = sys_off[s]; i < sys_off[s + 1]; ++i)
      for (int a = 0; a < 3; ++a) pos[3 * i + a] *= m;
    if (mu) mu[s] = m;
  }
}

/* Shoemake random quaternion (core/utils/quaternion.py:70-97), the power q^w
 * (_pow_quaternion: sign fix w >= 0, angle scaled) and the rotation matrix
 * (_quat_to_mat), applied as row-vector @ R (moves.py:190). */
static void gc_quat(const float u[3], double w, double R[3][3]) {
  double u1 = u[0], u2 = u[1], u3 = u[2];
  double q[4] = {sqrt(1 - u1) * sin(2 * M_PI * u2), sqrt(1 - u1) * cos(2 * M_PI * u2),
                 sqrt(u1) * sin(2 * M_PI * u3), sqrt(u1) * cos(2 * M_PI * u3)};
  if (q[0] < 0) for (int k = 0; k < 4; ++k) q[k] = -q[k];
  double cw = q[0] > 1 ? 1 : (q[0] < -1 ? -1 : q[0]);
  double ang = 2 * acos(cw);
  double sn = sqrt(q[1]*q[1] + q[2]*q[2] + q[3]*q[3]);
  double ax[3] = {1, 0, 0};
  if (sn > 1e-10) { ax[0] = q[1]/sn; ax[1] = q[2]/sn; ax[2] = q[3]/sn; }
  double na = ang * w;
  double W = cos(na / 2), X = ax[0]*sin(na/2), Y = ax[1]*sin(na/2), Z = ax[2]*sin(na/2);
  double x2 = X*X, y2 = Y*Y, z2 = Z*Z, w2 = W*W;
  R[0][0] = x2 - y2 - z2 + w2; R[0][1] = 2*(X*Y - Z*W); R[0][2] = 2*(X*Z + Y*W);
  R[1][0] = 2*(X*Y + Z*W); R[1][1] = -x2 + y2 - z2 + w2; R[1][2] = 2*(Y*Z - X*W);
  R[2][0] = 2*(X*Z - Y*W); R[2][1] = 2*(Y*Z + X*W); R[2][2] = -x2 - y2 + z2 + w2;
}

static void wrap_f(const cell_t* c, double* r) { wrap_pos(c, r); }

/* minimum-image difference vector b - a (round of fractional components) */
static void min_image(const cell_t* c, const double* a, const double* b, double* d) {
  double dd[3] = {b[0]-a[0], b[1]-a[1], b[2]-a[2]}, f[3];
  rmul(c->Li, dd, f);
  for (int k = 0; k < 3; ++k) f[k] -= rint(f[k]);
  rmul(c->L, f, d);
}

typedef struct {
  /* sizes */
  int n_chain, n_steps, gmax, m, nspec, mx, my, mz;
  int n_motif; const int32_t* motif_nsite; /* guest species and their site counts */
  int64_t step0;
  uint64_t seed;
  /* hosts */
  const int32_t* host_off; const double* host_pos; const double* host_q; const int32_t* host_spec;
  const double* tril;
  /* motifs [n_motif x m (x 3)] (templates centred at their geometric centres) */
  const double* motif_pos; const double* motif_q; const int32_t* motif_spec;
  /* potentials */
  const double* sig; const double* eps; double rc_lj; double rc_cl;
  const double* alpha; const double* kmax; const double* tail; /* n_chain x nspec x nspec */
  const double* mol_const; /* n_chain x n_motif: E_self + E_excl of one molecule */
  const double* kT; const double* log_activity; /* n_chain x n_motif */
  double w_move[4]; /* translation, rotation, reinsertion, exchange */
  double upper[4];  /* step-width upper bounds */
  /* state (in/out) */
  double* guest_pos; int32_t* n_guest; int32_t* guest_motif; /* n_chain x gmax */
  int32_t* counts; /* n_chain x nspec */
  double* sk; double* energy;
  float* width; uint32_t* hist; int32_t* hist_idx; /* n_chain x 4 (x 4 words) */
  int64_t* stats; /* n_chain x 8: trials[4], accepted[4] */
  int64_t* ins_acc; int64_t* del_acc;
  /* optional per-trial trace of chain 0..n_trace-1 */
  int n_trace; double* trace; /* n_trace x n_steps x 4: move, dE, logp, accepted */
} gc_args;

static double gc_pair(const gc_args* a, const cell_t* c, int ch, const double* r, double q, int sp,
                      const double* rj, double qj, int spj) {
  double d[3]; min_image(c, r, rj, d);
  double r2 = d[0]*d[0] + d[1]*d[1] + d[2]*d[2];
  double e = 0;
  int p = sp * a->nspec + spj;
  if (r2 < a->rc_lj * a->rc_lj && a->eps[p] != 0) {
    double c6 = pow(a->sig[p]*a->sig[p]/r2, 3);
    e += 4*a->eps[p]*(c6*c6 - c6);
  }
  double rr = sqrt(r2);
  if (rr < a->rc_cl) e += coulomb_k() * q * qj * erfc(a->alpha[ch]*rr) / rr;
  return e;
}

static void gc_grid_n(const gc_args* a, int k, int* n0, int* n1, int* n2) {
  int wy = 2*a->my + 1, wz = 2*a->mz + 1;
  *n0 = k / (wy*wz); *n1 = (k / wz) % wy - a->my; *n2 = k % wz - a->mz;
}

int orc_gcmc_run(gc_args* a) {
  if (a->n_motif < 1 || a->n_motif > 8 || a->m < 1 || a->m > 8) return -1;
  const double K = coulomb_k();
  const int nk = (a->mx + 1) * (2*a->my + 1) * (2*a->mz + 1);
  const int M = a->m;
#pragma omp parallel for schedule(dynamic)
  for (int ch = 0; ch < a->n_chain; ++ch) {
    cell_t c; cell_from_tril(a->tril + 6*ch, &c);
    double* P = (double*)malloc(sizeof(double) * nk);
    double (*kv)[3] = malloc(sizeof(double) * 3 * nk);
    double* dS = (double*)malloc(sizeof(double) * 2 * nk);
    for (int k = 0; k < nk; ++k) {
      int n0, n1, n2; gc_grid_n(a, k, &n0, &n1, &n2);
      int n[3] = {n0, n1, n2};
      for (int x = 0; x < 3; ++x)
        kv[k][x] = 2*M_PI*(c.Li[x][0]*n[0] + c.Li[x][1]*n[1] + c.Li[x][2]*n[2]);
      double k2 = kv[k][0]*kv[k][0] + kv[k][1]*kv[k][1] + kv[k][2]*kv[k][2];
      double al = a->alpha[ch];
      int in = a->kmax[ch] < 0 || sqrt(k2) <= a->kmax[ch];
      P[k] = (k2 > 0 && in) ? 2*M_PI/c.vol*exp(-k2/(4*al*al))/k2*(n0 == 0 ? 1.0 : 2.0) : 0.0;
    }
    const int h0 = a->host_off[ch], h1 = a->host_off[ch+1];
    double* gp = a->guest_pos + (size_t)ch * a->gmax * M * 3;
    double* S = a->sk + (size_t)ch * nk * 2;
    int32_t* cnt = a->counts + (size_t)ch * a->nspec;
    const double* T = a->tail + (size_t)ch * a->nspec * a->nspec;
    for (int st = 0; st < a->n_steps; ++st) {
      int64_t t = a->step0 + st;
      uint32_t bb[4]; gc_draw(a->seed, 0xFFFFFFFFu, t, GC_TAG_BATCH, bb);
      double wsum = a->w_move[0] + a->w_move[1] + a->w_move[2] + a->w_move[3];
      double um = gc_u01(bb[0]) * wsum, acc_w = 0;
      int move = 3;
      for (int k = 0; k < 4; ++k) { acc_w += a->w_move[k]; if (um < acc_w) { move = k; break; } }
      int deleting = (bb[1] >> 31) & 1;
      uint32_t ra[4], rb[4], rcn[4];
      gc_draw(a->seed, (uint32_t)ch, t, GC_TAG_A, ra);
      gc_draw(a->seed, (uint32_t)ch, t, GC_TAG_B, rb);
      gc_draw(a->seed, (uint32_t)ch, t, GC_TAG_C, rcn);
      int ng = a->n_guest[ch];
      int32_t* gm = a->guest_motif + (size_t)ch * a->gmax;
      /* molecules of each species */
      int nmol[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int g = 0; g < ng; ++g) nmol[gm[g]] += 1;
      /* old / new sites; an exchange picks the species uniformly (moves.py:
       * 600-602, 711-712), a deletion then one guest of it; displacement
       * moves pick a guest of any species */
      double oldp[8][3], newp[8][3];
      int n_old = 0, n_new = 0, slot = -1;
      int tm = move == 3 ? (int)(ra[2] % (uint32_t)a->n_motif) : 0;
      float* wid = a->width + (size_t)ch * 4;
      if (move < 3 && ng > 0) {
        slot = (int)(ra[0] % (uint32_t)ng);
        tm = gm[slot];
      } else if (move == 3 && deleting && nmol[tm] > 0) {
        int k = (int)(ra[0] % (uint32_t)nmol[tm]);
        for (int g = 0; g < ng; ++g)
          if (gm[g] == tm && k-- == 0) { slot = g; break; }
      }
      const int Mt = a->motif_nsite[tm];
      const double* mpos = a->motif_pos + (size_t)tm * M * 3;
      const double* mq = a->motif_q + (size_t)tm * M;
      const int32_t* msp = a->motif_spec + (size_t)tm * M;
      if (slot >= 0) {
        n_old = Mt;
        for (int s = 0; s < Mt; ++s) for (int x = 0; x < 3; ++x) oldp[s][x] = gp[(slot*M + s)*3 + x];
      }
      if (move == 0 && n_old) {
        /* translation N(0,1) * width per axis (moves.py:222-244) */
        float u0 = gc_u01(rcn[0]), u1 = gc_u01(rcn[1]), u2 = gc_u01(rcn[2]), u3 = gc_u01(rcn[3]);
        double r0 = sqrt(-2.0*log((double)u0)), r1 = sqrt(-2.0*log((double)u2));
        double g3[3] = {r0*cos(2*M_PI*u1), r0*sin(2*M_PI*u1), r1*cos(2*M_PI*u3)};
        for (int s = 0; s < Mt; ++s) {
          for (int x = 0; x < 3; ++x) newp[s][x] = oldp[s][x] + g3[x] * wid[0];
          wrap_f(&c, newp[s]);
        }
        n_new = Mt;
      } else if ((move == 1 || move == 2) && n_old) {
        /* rotation about the geometric centre (moves.py:158-197), q^w; reinsertion
         * = full rotation + uniform fractional offset (268-295) */
        float uq[3] = {gc_u01(rb[0]), gc_u01(rb[1]), gc_u01(rb[2])};
        double R[3][3]; gc_quat(uq, move == 1 ? wid[1] : 1.0, R);
        double rel[8][3], cen[3] = {0, 0, 0};
        for (int s = 0; s < Mt; ++s) {
          min_image(&c, oldp[Mt-1], oldp[s], rel[s]);
          for (int x = 0; x < 3; ++x) cen[x] += rel[s][x] / Mt;
        }
        double com[3] = {oldp[Mt-1][0] + cen[0], oldp[Mt-1][1] + cen[1], oldp[Mt-1][2] + cen[2]};
        wrap_f(&c, com);
        double off[3] = {0, 0, 0};
        if (move == 2) {
          double fu[3] = {gc_u01(rcn[0]), gc_u01(rcn[1]), gc_u01(rcn[2])};
          rmul(c.L, fu, off);
        }
        for (int s = 0; s < Mt; ++s) {
          double rr[3]; min_image(&c, com, oldp[s], rr);
          for (int x = 0; x < 3; ++x)
            newp[s][x] = com[x] + rr[0]*R[0][x] + rr[1]*R[1][x] + rr[2]*R[2][x] + off[x];
          wrap_f(&c, newp[s]);
        }
        n_new = Mt;
      } else if (move == 3 && !deleting) {
        /* insertion: template @ R + U L (moves.py:573-675) */
        float uq[3] = {gc_u01(rb[0]), gc_u01(rb[1]), gc_u01(rb[2])};
        double R[3][3]; gc_quat(uq, 1.0, R);
        double fu[3] = {gc_u01(rcn[0]), gc_u01(rcn[1]), gc_u01(rcn[2])}, off[3];
        rmul(c.L, fu, off);
        if (ng < a->gmax) {
          for (int s = 0; s < Mt; ++s) {
            const double* tp = mpos + 3*s;
            for (int x = 0; x < 3; ++x)
              newp[s][x] = tp[0]*R[0][x] + tp[1]*R[1][x] + tp[2]*R[2][x] + off[x];
            wrap_f(&c, newp[s]);
          }
          n_new = Mt;
        }
      }
      /* energy difference */
      double dE = 0;
      for (int pass = 0; pass < 2; ++pass) {
        int ns = pass ? n_new : n_old;
        double sgn = pass ? 1.0 : -1.0;
        for (int s = 0; s < ns; ++s) {
          double* r = pass ? newp[s] : oldp[s];
          double q = mq[s]; int sp = msp[s];
          double e = 0;
          for (int j = h0; j < h1; ++j)
            e += gc_pair(a, &c, ch, r, q, sp, a->host_pos + 3*j, a->host_q[j], a->host_spec[j]);
          for (int g = 0; g < ng; ++g) {
            if (g == slot) continue;
            const int t2 = gm[g];
            for (int s2 = 0; s2 < a->motif_nsite[t2]; ++s2)
              e += gc_pair(a, &c, ch, r, q, sp, gp + (g*M + s2)*3, a->motif_q[t2*M + s2],
                           a->motif_spec[t2*M + s2]);
          }
          dE += sgn * e;
        }
      }
      double dElr = 0;
      for (int k = 0; k < nk; ++k) {
        double dr = 0, di = 0;
        for (int pass = 0; pass < 2; ++pass) {
          int ns = pass ? n_new : n_old;
          double sgn = pass ? 1.0 : -1.0;
          for (int s = 0; s < ns; ++s) {
            double* r = pass ? newp[s] : oldp[s];
            double th = kv[k][0]*r[0] + kv[k][1]*r[1] + kv[k][2]*r[2];
            dr += sgn * mq[s] * cos(th);
            di += sgn * mq[s] * sin(th);
          }
        }
        dS[2*k] = dr; dS[2*k+1] = di;
        double sr = S[2*k], si = S[2*k+1];
        dElr += K * P[k] * (2*(sr*dr + si*di) + dr*dr + di*di);
      }
      dE += dElr;
      double logp = 0;
      int dn = 0;
      if (move == 3) dn = n_new ? 1 : (n_old ? -1 : 0);
      if (dn) {
        dE += dn * a->mol_const[(size_t)ch * a->n_motif + tm];
        /* tail: sum_ab T_ab (N'_a N'_b - N_a N_b) */
        int np_[16];
        for (int s = 0; s < a->nspec; ++s) np_[s] = cnt[s];
        for (int s = 0; s < Mt; ++s) np_[msp[s]] += dn;
        double et = 0;
        for (int x = 0; x < a->nspec; ++x)
          for (int y = 0; y < a->nspec; ++y)
            et += T[x*a->nspec + y] * ((double)np_[x]*np_[y] - (double)cnt[x]*cnt[y]);
        dE += et;
        /* fugacity term of the exchanged species (probability.py:117-173) */
        double lnfV = a->log_activity[(size_t)ch * a->n_motif + tm] + log(c.vol);
        if (dn > 0) logp += lnfV - log((double)(nmol[tm] + 1));
        else logp += -lnfV + log((double)nmol[tm]);
      }
      logp += -dE / a->kT[ch];
      double lu = log((double)gc_u01(ra[1]));
      int accept = logp > lu;
      if (ch < a->n_trace) {
        double* tr = a->trace + ((size_t)ch * a->n_steps + st) * 4;
        tr[0] = move * 2 + (move == 3 ? deleting : 0); tr[1] = dE; tr[2] = logp; tr[3] = accept;
      }
      int64_t* stt = a->stats + (size_t)ch * 8;
      stt[move] += 1;
      if (accept) {
        stt[4 + move] += 1;
        if (n_new && n_old) {
          for (int s = 0; s < Mt; ++s) for (int x = 0; x < 3; ++x) gp[(slot*M + s)*3 + x] = newp[s][x];
        } else if (n_new) {
          for (int s = 0; s < Mt; ++s) for (int x = 0; x < 3; ++x) gp[(ng*M + s)*3 + x] = newp[s][x];
          gm[ng] = tm;
          a->n_guest[ch] = ng + 1;
          a->ins_acc[ch] += 1;
        } else if (n_old) {
          for (int s = 0; s < M; ++s) for (int x = 0; x < 3; ++x) gp[(slot*M + s)*3 + x] = gp[((ng-1)*M + s)*3 + x];
          gm[slot] = gm[ng - 1];
          a->n_guest[ch] = ng - 1;
          a->del_acc[ch] += 1;
        }
        if (dn) for (int s = 0; s < Mt; ++s) cnt[msp[s]] += dn;
        for (int k = 0; k < nk; ++k) { S[2*k] += dS[2*k]; S[2*k+1] += dS[2*k+1]; }
        a->energy[ch] += dE;
      }
      /* acceptance-target schedule of the chosen move (parameter_scheduler.py:161-208) */
      uint32_t* hb = a->hist + ((size_t)ch * 4 + move) * 4;
      int32_t* hi = a->hist_idx + (size_t)ch * 4 + move;
      int idx = *hi;
      if (accept) hb[idx >> 5] |= (1u << (idx & 31)); else hb[idx >> 5] &= ~(1u << (idx & 31));
      idx = (idx + 1) % 100;
      *hi = idx;
      if (idx == 0) {
        int ones = 0;
        for (int w = 0; w < 4; ++w) ones += __builtin_popcount(hb[w] & (w == 3 ? 0xFu : 0xFFFFFFFFu));
        double avg = ones / 100.0;
        if (fabs(avg - 0.5) > 0.05) {
          double v = wid[move];
          v = avg < 0.5 ? v / 1.1 : v * 1.1;
          if (v < 0) v = 0;
          if (v > a->upper[move]) v = a->upper[move];
          wid[move] = (float)v;
        }
      }
    }
    free(P); free(kv); free(dS);
  }
  return 0;
}
