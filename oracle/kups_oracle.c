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
