/*
 * kupsmi.h -- C ABI of libkupsmi.so, the MI355X (gfx950) implementation of
 * kUPS's batched particle-interaction hot path.
 *
 * Conventions (all entry points):
 *   - Every pointer argument except `ctx` and host scalars is a DEVICE pointer
 *     owned by the caller.  Nothing here allocates on the hot path: scratch
 *     lives in the context and grows on first use by a larger call.  Growing
 *     a scratch slot frees the old buffer, so work recorded with a context
 *     (a captured HIP graph) must keep that context to itself and freeze it
 *     (kupsmi_ctx_freeze), which makes any further growth an error.
 *   - `stream` is a hipStream_t passed as void*; all work is enqueued on it
 *     and no entry point synchronises the host.  Capacity overflows are
 *     reported through device counters that the caller reads at its own sync
 *     point (the reference's Capacity/fix_or_raise convention,
 *     core/capacity.py:175-234, core/propagator.py:137-172).
 *   - Return value: 0 ok, <0 error (see KUPSMI_ERR_*; kupsmi_last_error()
 *     gives the message).  No C++ exception crosses the ABI.
 *   - Systems are contiguous atom ranges given by a system-offset table
 *     sys_off[n_sys + 1] (the kUPS particle Table argsorted by system,
 *     cf. potential/common/graph.py:122-164).
 *   - Unit cells: lower-triangular lattice, tril[n_sys][6] =
 *     [L00, L10, L11, L20, L21, L22] in FP64 (core/unitcell.py:251-300).
 *   - Positions, momenta, forces: FP32, row-major (n, 3) like kUPS arrays.
 *     Energies and virials: FP64 per system.
 *   - Units: eV, Angstrom, amu; time in kUPS internal units
 *     (1 fs = core/constants.py FEMTO_SECOND).
 */
#ifndef KUPSMI_H
#define KUPSMI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KUPSMI_OK 0
#define KUPSMI_ERR_ARG (-1)
#define KUPSMI_ERR_HIP (-2)
#define KUPSMI_ERR_UNSUPPORTED (-3)

/* largest system the LDS-resident kernels hold (atoms per system) */
#define KUPSMI_LDS_MAX_ATOMS 4096

typedef struct kupsmi_ctx kupsmi_ctx;

/* Library version (major*10000 + minor*100 + patch). */
int kupsmi_version(void);
/* Message of the last error returned on this thread ("" if none). */
const char* kupsmi_last_error(void);

/* Context: owns scratch on the current HIP device.  No counterpart in the
 * reference (JAX/XLA owns buffers there); it is the handle the jax.ffi
 * adapter would keep per device. */
int kupsmi_ctx_create(kupsmi_ctx** out);
int kupsmi_ctx_destroy(kupsmi_ctx* ctx);
/* frozen != 0: a call that needs more scratch than the context holds fails
 * (KUPSMI_ERR_ARG) instead of reallocating, so addresses held by recorded
 * work stay valid; frozen == 0 lifts it.  No counterpart in the reference. */
int kupsmi_ctx_freeze(kupsmi_ctx* ctx, int frozen);

/*
 * Full directed neighbor list of every system (lh = rh, no remap).
 * Replaces NearestNeighborList.__call__(lh, None, systems, cutoffs)
 * (core/neighborlist.py:233-255) as implemented by CellListNeighborList /
 * DenseNearestNeighborList (985-1077, 902-982) via basic_neighborlist
 * (743-786): same system, strict |d| < rc decided in FP64, shifts
 * round-half-even(f_i - f_j) or the centred image grid for systems whose
 * perpendicular width is below 2 rc (453-487, 595-624), pairs with equal
 * exclusion id dropped on the minimum image only (731-735).
 * Edges are written i-major: edges_ij[e] = (i, j), edges_shift[e] = (s0,s1,s2);
 * at most `capacity` edges are written and *count (device int64) receives the
 * required number (> capacity means: regrow and call again).
 */
int kupsmi_neighborlist(kupsmi_ctx* ctx, void* stream, int n_sys,
                        const int32_t* sys_off, int n_atoms, const float* pos,
                        const int32_t* exclusion, const double* tril,
                        const double* cutoff, int64_t capacity,
                        int32_t* edges_ij, int32_t* edges_shift, int64_t* count);

/*
 * The same directed pair set from a cell list.  Replaces CellListNeighborList
 * (core/neighborlist.py:985-1077: _cell_hash / _cell_stencil / _num_cells,
 * 258-292, and _cell_list_subselect, 295-364, feeding basic_neighborlist,
 * 743-786): the atoms of each system are counting-sorted into cells at
 * least rc wide (at most 16 per axis), and each atom tests the atoms of the
 * distinct cells among its 27 neighbours (the reference's jnp.unique over
 * the stencil hashes) with the minimum-image shift round(f_i - f_j) and the
 * FP64 decision of kupsmi_neighborlist; systems narrower than 2 rc use the
 * image grid over all their atoms.  The edge SET equals kupsmi_neighborlist's
 * (the reference's contract: its order comes from jnp.unique/argsort);
 * the order here is i-major, then stencil cell, then the cell's atoms in
 * index order.  Systems hold at most 16384 atoms (max_atoms_per_sys).
 * Capacity and *count as for kupsmi_neighborlist.
 */
int kupsmi_neighborlist_cells(kupsmi_ctx* ctx, void* stream, int n_sys,
                              const int32_t* sys_off, int n_atoms, int max_atoms_per_sys,
                              const float* pos, const int32_t* exclusion, const double* tril,
                              const double* cutoff, int64_t capacity, int32_t* edges_ij,
                              int32_t* edges_shift, int64_t* count);

/*
 * Edges between the full particle table (lh: pos, exclusion, systems by
 * sys_off) and k further particles (rh: rh_pos [k x 3], rh_system [k],
 * rh_exclusion [k]).  Replaces NearestNeighborList.__call__(lh, rh, systems,
 * cutoffs, rh_index_remap) (core/neighborlist.py:233-255), the form that
 * RadiusGraphConstructor's patch queries (potential/common/graph.py:252-267)
 * and neighborlist_changes (1467) make, with the rules of
 * _build_segmentation_mask (505-539) and _compact_edges (651-689):
 *   - a pair (i, r) needs i and r in the same system (rh_system outside
 *     [0, n_sys) drops r); shift round(f_i - f_r), or every offset of the
 *     image grid of a system narrower than 2 rc; kept when
 *     |(f_i - f_r - s) @ L|^2 < rc^2 (FP64, strict) and not (equal exclusion
 *     ids on the minimum image);
 *   - rh_index_remap (optional): rh particle r is lh particle remap[r]; an lh
 *     particle named by any remap entry pairs only with entries whose remap
 *     is <= its index (536-537), and every edge is output in both directions:
 *     all edges (i, remap[r], s) first, then (remap[r], i, -s) in the same
 *     order (680-685).  A remap outside [0, n_atoms) names no lh particle
 *     and is written as n_atoms (the reference's out-of-bounds index).
 *     Without a remap the edges are (i, r, s), r indexing the rh particles.
 * Forward edges are ordered by r, then i, then image offset.  Every rh
 * particle scans its system densely (O(k x system size)): the form suits
 * the small patches of MC moves.  Capacity and *count as for
 * kupsmi_neighborlist (the count includes both directions).
 */
int kupsmi_neighborlist_query(kupsmi_ctx* ctx, void* stream, int n_sys,
                              const int32_t* sys_off, int n_atoms, const float* pos,
                              const int32_t* exclusion, int k, const float* rh_pos,
                              const int32_t* rh_system, const int32_t* rh_exclusion,
                              const int32_t* rh_index_remap, const double* tril,
                              const double* cutoff, int64_t capacity, int32_t* edges_ij,
                              int32_t* edges_shift, int64_t* count);

/*
 * Parity output of the force kernels (no counterpart in the reference: it
 * exposes the pair set a force evaluation used, for comparison with the
 * neighbor list above).  Arms the next pair-evaluating call made with the
 * context (kupsmi_md_lj, kupsmi_lj_forces, kupsmi_pair_forces) or with the
 * Verlet handle (kupsmi_pair_forces_verlet) to write every directed pair its
 * force loops kept: for kupsmi_md_lj the pairs of the last step's final
 * force pass, at the positions the call writes back; for the others the
 * pairs at the caller's positions.  ij[e] = (i, j) global particle indices,
 * shift[e] = s with d = r_j - r_i + s @ L in the caller's frame, terms[e]
 * (optional) = 1 LJ | 2 real-space Coulomb; entries in no particular order;
 * *count (device uint64) receives the number of pairs (entries past capacity
 * are not written).  The armed call runs the recording instantiation of its
 * kernels: the same source and pair decisions plus the writes.  One call
 * consumes the arming.
 */
int kupsmi_ctx_record_pairs(kupsmi_ctx* ctx, int64_t capacity, int32_t* ij, int32_t* shift,
                            uint8_t* terms, unsigned long long* count);

/*
 * Edges that moving k particles adds and removes.  Replaces
 * neighborlist_changes (core/neighborlist.py:1424-1515) for the full list
 * above: moved[k] are distinct particle indices, moved_pos[k x 3] their new
 * positions.  removed = every directed edge of the current configuration with
 * an end among the moved particles; added = every directed edge of the
 * configuration with the moved particles at moved_pos that has an end among
 * them.  Both use the pair decision of kupsmi_neighborlist and report the
 * particles' own indices (the reference's remap of appended entries).  Each
 * output is written up to its capacity, with the required edge count in
 * *count_added / *count_removed (device int64; > capacity means regrow and
 * call again, -1 in both means an index in moved is out of range or
 * repeated).  Order: moved particle, partner, image, edge before its reverse.
 */
int kupsmi_neighborlist_changes(kupsmi_ctx* ctx, void* stream, int n_sys,
                                const int32_t* sys_off, int n_atoms, const float* pos,
                                const int32_t* exclusion, const double* tril,
                                const double* cutoff, int k, const int32_t* moved,
                                const float* moved_pos, int64_t capacity_added,
                                int32_t* added_ij, int32_t* added_shift, int64_t* count_added,
                                int64_t capacity_removed, int32_t* removed_ij,
                                int32_t* removed_shift, int64_t* count_removed);

/*
 * Morse bond energy, forces and virial over an explicit bond list.  Replaces
 * morse_bond_energy (potential/classical/morse.py:142-168) and its gradient:
 * for bond b = (i, j) with image shift s_b, d = r_j - r_i + s_b @ L and
 * E = D [1 - exp(-alpha (r - r0))]^2 with D/alpha/r0 taken from the
 * (n_species x n_species) tables at (species[i], species[j]); every bond is
 * counted once (no 1/2).  bonds: n_bonds x 2 int32 global particle indices
 * (particles sorted by system, sys_off as everywhere); shifts: n_bonds x 3
 * int32 or NULL.  Bond vector and energy in FP64; forces (f32) and per-system
 * energy/virial (f64) are added with atomics (summation order not fixed).
 * accumulate = 0 clears forces, energy and virial first.
 */
int kupsmi_morse_bonds(void* stream, int n_sys, const int32_t* sys_off, int n_atoms,
                       const float* pos, const int32_t* species, int n_species,
                       const double* D, const double* alpha, const double* r0,
                       long long n_bonds, const int32_t* bonds, const int32_t* shifts,
                       const double* tril, int accumulate, float* forces, double* energy,
                       double* virial);

/*
 * Lennard-Jones energy, forces and virial of every system at the given
 * positions.  Replaces the LJ Potential (potential/classical/lennard_jones.py:
 * 149-178 energy, differentiated by potential/common/energy.py:263-355):
 * per directed edge u = 4 eps (c6^2 - c6), c6 = (sigma^2/r^2)^3, hard cutoff
 * r^2 < rc^2 without shift, E = sum / 2; F = -dE/dr; virial
 * W = -sum_directed u'(r^2) d (x) d (stress = W / V, observables/stress.py).
 * Pair arithmetic FP32, cutoff band re-decided in FP64, energies/virials
 * accumulated in FP64.  sigma/eps: (n_species x n_species) mixed tables in
 * FP64 (Lorentz-Berthelot, lennard_jones.py:125-141), n_species <= 16 (also
 * for kupsmi_pair_forces_verlet).
 * Exclusion is by particle (i != j on the minimum image), as in md_lj.
 * Positions are wrapped into the cell by whole lattice vectors (in FP64,
 * rounded once to FP32) before the pairs are decided; a position already
 * inside the cell is used bit for bit.
 * Systems must hold <= KUPSMI_LDS_MAX_ATOMS atoms.
 * Outputs: forces (n_atoms x 3, f32), energy[n_sys], virial[n_sys x 9] (f64).
 */
int kupsmi_lj_forces(kupsmi_ctx* ctx, void* stream, int n_sys,
                     const int32_t* sys_off, int n_atoms, const float* pos,
                     const int32_t* species, int n_species, const double* sigma,
                     const double* epsilon, const double* tril,
                     const double* cutoff, float* forces, double* energy,
                     double* virial);

/*
 * Batched LJ molecular dynamics: n_steps steps of every system in ONE
 * persistent launch, each system resident in LDS (one workgroup per system,
 * work queue largest-first).  Replaces the MD step propagators built by
 * make_md_step_from_state (md/integrators.py:904-962):
 *   integrator 0: velocity Verlet  B(dt/2) A(dt) F B(dt/2)      (273-317)
 *   integrator 1: BAOAB Langevin   B(dt/2) A(dt/2) O A(dt/2) F B(dt/2) (405-453)
 *   integrator 2: CSVR + VV        T B(dt/2) A(dt) F B(dt/2)          (470-635)
 *                 T: momenta scaled by the stochastic velocity rescaling
 *                 factor (Bussi, Donadio, Parrinello 2007) of each system,
 *                 one R1 ~ N(0,1) for the batch and R2 ~ chi^2(3N - 4) per
 *                 system from Philox4x32-10 (seed, step), as the oracle draws.
 *   integrator 3: CSVR + VV + stochastic cell rescaling (829-888):
 *                 T B(dt/2) A(dt) F B(dt/2) S F, S scaling the cell and the
 *                 positions by mu = exp(d eps / 3) (Bernetti and Bussi 2020),
 *                 d eps = (dt/tau_P) beta (P - P0) + sqrt(2 kT beta dt /
 *                 (tau_P V)) R with one R ~ N(0,1) for the batch, P from the
 *                 kinetic energy and the virial of the first F; the cell list
 *                 is rebuilt every step (skin ignored).
 * A steps wrap positions to [-0.5, 0.5) fractional (104-141); O uses
 * p <- c1 p + sqrt(kT (1 - c1^2)) sqrt(m) eta with eta from Philox4x32-10
 * keyed by (seed, global atom index, step), see kupsmi_normals.
 * The neighbor set of every step is exactly the set of kupsmi_neighborlist
 * with particle exclusion (the reference rebuilds it every step).  With
 * skin == 0 it is rebuilt every step from a cell list in LDS.  With skin > 0
 * (and n_steps >= 2) each system keeps Verlet lists of radius cutoff + skin
 * (max_neighbors entries per atom, in HBM) and rebuilds them whenever an
 * atom may have moved by skin/2 since the last build; every step still
 * tests every listed pair against the cutoff, so the pair set is unchanged.
 * Systems thinner than 2 cutoffs, and any system whose lists overflow,
 * use the per-step cell list instead.
 * In/out: pos, mom, forces (F at the current positions on entry: the
 * reference initialises it to zero, md/data.py:188).
 * Per system: mass per atom (f32), dt, kT (eV), gamma (1/internal time),
 * tau (CSVR coupling time, internal units; NULL unless integrator >= 2),
 * baro[n_sys x 4] (target pressure eV/A^3, coupling time, compressibility
 * A^3/eV, minimum scale factor; integrator 3) and tril_out (the rescaled
 * cell, may alias tril; integrator 3, else NULL).
 * Outputs: pot_energy[n_sys x n_steps] and kin_energy[n_sys x n_steps] (f64,
 * after each step; either may be NULL), virial[n_sys x 9] of the last step
 * (may be NULL).  step0 is the global index of the first step (RNG stream
 * position).  n_steps == 0 only evaluates forces/energy at the input state.
 */
int kupsmi_md_lj(kupsmi_ctx* ctx, void* stream, int integrator, int n_steps,
                 int64_t step0, uint64_t seed, int n_sys, const int32_t* sys_off,
                 int n_atoms, float* pos, float* mom, float* forces,
                 const float* mass, const int32_t* species, int n_species,
                 const double* sigma, const double* epsilon, const double* tril,
                 const double* cutoff, double skin, int max_neighbors,
                 const double* dt, const double* kT, const double* gamma,
                 const double* tau, const double* baro, double* tril_out,
                 double* pot_energy, double* kin_energy, double* virial);

/*
 * The standard normals eta[n_atoms x 3] that kupsmi_md_lj's O step uses at
 * global step `step` (Philox4x32-10, Salmon et al. SC'11, counter =
 * (atom, step), key = seed; Box-Muller).  For parity tests of the stochastic
 * step.  No counterpart in the reference (jax.random.normal there).
 */
int kupsmi_normals(void* stream, uint64_t seed, int64_t step, int n_atoms,
                   float* eta);

/*
 * The bare Philox4x32-10 generator that every random stream of the library
 * draws from: out[i] = Philox4x32-10(counters[4 i..], keys[2 i..]) for
 * n independent (counter, key) pairs.  For known-answer checks of the
 * shared stream (Salmon et al. SC'11; Random123's published vectors).  No
 * counterpart in the reference (it draws from jax.random).
 */
int kupsmi_philox(void* stream, int n, const uint32_t* counters, const uint32_t* keys,
                  uint32_t* out);

/*
 * LJ + Ewald real-space energies, forces and virials of systems of any size
 * (<= 16384 atoms each), state in HBM.  Replaces, for one evaluation, the LJ
 * energy (potential/classical/lennard_jones.py:149-178) and the Ewald
 * short-range energy (potential/classical/ewald.py:305-325:
 * K/2 sum_directed q_i q_j erfc(alpha r)/r [r < rc]) with their jax.vjp
 * gradients (potential/common/energy.py:297-300).
 * LJ is skipped when sigma/epsilon/lj_cutoff are NULL; Coulomb when
 * alpha/coul_cutoff/charges are NULL.  lj_exclusion / coul_exclusion: ids
 * whose equal pairs are dropped on the minimum image (NULL: particle id),
 * so GCMC hosts can exclude host-host LJ while keeping host-host Coulomb
 * (application/simulations/mcmc_rigid.py:156-165).
 * energy: n_sys x 2 (LJ, real-space) FP64; virial n_sys x 9.  accumulate != 0
 * adds into forces/energy/virial instead of overwriting.
 * lj_truncation (n_sys FP64 or NULL): the smooth pairwise switch of
 * pair_tail_corrected_lennard_jones_energy (lennard_jones.py:200-222): pair
 * energies between r_tr and the LJ cutoff are multiplied by
 * S = (rc^2 - r^2)^2 (2 r^2 + rc^2 - 3 r_tr^2) / (rc^2 - r_tr^2)^3 and the
 * forces and virial are those of the switched energy; a system with
 * r_tr < 0 or r_tr >= rc is not switched.
 */
int kupsmi_pair_forces(kupsmi_ctx* ctx, void* stream, int n_sys, const int32_t* sys_off,
                       int n_atoms, int max_atoms_per_sys, const float* pos,
                       const float* charges, const int32_t* species, int n_species,
                       const double* sigma, const double* epsilon, const double* lj_cutoff,
                       const int32_t* lj_exclusion, const double* alpha,
                       const double* coul_cutoff, const int32_t* coul_exclusion,
                       const double* tril, int accumulate, float* forces, double* energy,
                       double* virial, const double* lj_truncation);

/*
 * The same evaluation over a Verlet list of radius rc + skin kept in a
 * kupsmi_verlet handle and reused across calls.  Replaces the per-step
 * rebuild of CellListNeighborList (core/neighborlist.py:985-1077) with the
 * neighbor-list reuse of a skin (no counterpart in the reference): each call
 * moves the stored positions by the lattice vectors the caller's wrap added,
 * and rebuilds the list of a system on the device once one of its atoms has
 * moved more than skin / 2 since the last build, so every pair within the
 * cutoff is in the list and the results equal kupsmi_pair_forces'.
 * max_neighbors: list entries per atom; an atom with more candidates within
 * rc + skin is evaluated over the cell stencil instead (exact, slower).
 * Between calls only positions may change (cells, charges, species,
 * exclusion ids and sizes fixed); kupsmi_verlet_invalidate forces a rebuild.
 * The handle allocates its buffers on first use (and when n_sys, n_atoms,
 * max_neighbors or the largest system grow) and is bound to one device and
 * to one stream's sequence of calls.
 */
typedef struct kupsmi_verlet kupsmi_verlet;
int kupsmi_verlet_create(kupsmi_verlet** out);
int kupsmi_verlet_destroy(kupsmi_verlet* v);
int kupsmi_verlet_invalidate(kupsmi_verlet* v, void* stream);
/* out[0]: system rebuilds, out[1]: atom lists that overflowed (host array,
 * synchronises the stream). */
int kupsmi_verlet_counters(kupsmi_verlet* v, void* stream, unsigned long long* out);
/* frozen != 0: a call that would resize the handle's buffers (more systems,
 * atoms, list entries or a larger system) fails instead of reallocating
 * them; for handles used by recorded work. */
int kupsmi_verlet_freeze(kupsmi_verlet* v, int frozen);
/* Kernel timing of the main pair pass (the kernel that evaluates the list):
 * returns in *total_ms / *calls the summed HIP-event time and the number of
 * calls timed since the previous kupsmi_verlet_profile (synchronising on
 * their events), then enables (enable != 0) or disables the timing of later
 * calls.  No counterpart in the reference; for measurement. */
int kupsmi_verlet_profile(kupsmi_verlet* v, int enable, double* total_ms, long long* calls);
/* Arms the handle's next kupsmi_pair_forces_verlet to make its stream wait
 * (hipStreamWaitEvent) for `event` after its list upkeep (binning and any
 * list rebuild) and before its pair passes; that call takes the gate.  With
 * the event of kupsmi_ewald_phase_event, the list upkeep runs beside the
 * other stream's structure-factor kernels and the pair passes beside its
 * force pass.  Results are unchanged (ordering only).  No counterpart in the
 * reference (a scheduling hook of the two-stream MD step); recordable in a
 * HIP graph like the call itself. */
int kupsmi_verlet_gate(kupsmi_verlet* v, void* event);
/* Arms the handle's next kupsmi_pair_forces_verlet to record its pair set,
 * as kupsmi_ctx_record_pairs does for the context's calls. */
int kupsmi_verlet_record_pairs(kupsmi_verlet* v, int64_t capacity, int32_t* ij, int32_t* shift,
                               uint8_t* terms, unsigned long long* count);
int kupsmi_pair_forces_verlet(kupsmi_verlet* v, void* stream, int n_sys, const int32_t* sys_off,
                              int n_atoms, int max_atoms_per_sys, const float* pos,
                              const float* charges, const int32_t* species, int n_species,
                              const double* sigma, const double* epsilon,
                              const double* lj_cutoff, const int32_t* lj_exclusion,
                              const double* alpha, const double* coul_cutoff,
                              const int32_t* coul_exclusion, const double* tril, double skin,
                              int max_neighbors, int accumulate, float* forces, double* energy,
                              double* virial);

/*
 * Reciprocal-space Ewald energy, forces, virial and structure factors.
 * Replaces ewald_long_range_energy (potential/classical/ewald.py:585-601,
 * structure factor 394-439, prefactor 363-391) and its gradient.
 * k grid: n in [0..mx] x [-my..my] x [-mz..mz] (row-major, n0 slowest), kept
 * where |k| <= kmax[s] (kmax NULL or < 0: all), k = 0 excluded
 * (kvecs_from_kmax, ewald.py:1322-1339).  mx, my, mz <= 12.
 * structure_factor (optional): n_sys x nk x 2 FP64 output S(k).
 */
int kupsmi_ewald_recip(kupsmi_ctx* ctx, void* stream, int n_sys, const int32_t* sys_off,
                       int n_atoms, int max_atoms_per_sys, const float* pos,
                       const float* charges, const double* tril, const double* alpha,
                       const double* kmax, int mx, int my, int mz, int accumulate,
                       float* forces, double* energy, double* virial,
                       double* structure_factor);
/* The context's phase event (a hipEvent_t, owned by the context): every
 * kupsmi_ewald_recip on ctx records it on its stream once the structure
 * factors, energies and force coefficients are done, right before the force
 * pass.  No counterpart in the reference (a scheduling hook, see
 * kupsmi_verlet_gate). */
int kupsmi_ewald_phase_event(kupsmi_ctx* ctx, void** event);

/*
 * kupsmi_ewald_recip without a context: the same kernels and results, with
 * the scratch in a file-scope device array of 64 MiB (one per GPU, held
 * until the process ends) instead of context memory, so the call allocates
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  Limit: the scratch a call needs,
 * n_sys x nk x (24 + 16 ceil(max_atoms_per_sys / 512)) bytes with
 * nk = (mx + 1)(2 my + 1)(2 mz + 1) (less 16 n_sys nk when structure_factor
 * is given), must fit in 64 MiB, else KUPSMI_ERR_ARG with a message.  Every
 * call on one GPU uses that GPU's single array: calls that overlap on a GPU
 * (other streams, other threads) race; serialise them.
 */
int kupsmi_ewald_recip_noalloc(void* stream, int n_sys, const int32_t* sys_off, int n_atoms,
                               int max_atoms_per_sys, const float* pos, const float* charges,
                               const double* tril, const double* alpha, const double* kmax,
                               int mx, int my, int mz, int accumulate, float* forces,
                               double* energy, double* virial, double* structure_factor);

/*
 * Elementwise integrator phases of the HBM-resident MD path
 * (md/integrators.py:162-402): phase 0 = B(dt/2) A(dt) (integrator 0) or
 * B(dt/2) A(dt/2) O A(dt/2) (integrator 1) and the fractional wrap; phase 1 =
 * B(dt/2) plus the per-system kinetic energy; phase 2 = kinetic energy only.
 * The O step draws the same Philox4x32-10 normals as kupsmi_md_lj.
 * step_dev (optional): a device int64 read by the kernels instead of `step`,
 * so steps recorded in a HIP graph draw fresh noise on every replay.
 */
int kupsmi_integrate(void* stream, int phase, int integrator, int n_sys,
                     const int32_t* sys_off, int max_atoms_per_sys, float* pos, float* mom,
                     const float* forces, const float* mass, const double* tril,
                     const double* dt, const double* kT, const double* gamma, uint64_t seed,
                     int64_t step, double* kin_energy, const int64_t* step_dev);

/*
 * Thermostat and barostat steps of the HBM-resident MD path, one workgroup
 * per system.  Phase 0: CSVR rescaling p <- alpha p (CSVRStep,
 * md/integrators.py:470-635) from each system's kinetic energy, with one R1
 * for the batch as the reference draws it and R2 = 2 Gamma((Ndof - 1) / 2)
 * per system; tau: coupling time per system.  Phase 1: stochastic cell
 * rescaling (StochasticCellRescalingStep, md/integrators.py:657-811) from
 * P = 2K / 3V + tr(virial) / 3V with one R for the batch; baro: n_sys x 4
 * (target pressure, coupling time, compressibility, minimum scale factor);
 * tril is scaled in place and the positions with it (momenta kept).  The
 * random streams and formulas are those of kupsmi_md_lj's integrators 2
 * and 3, so both paths run the same chains.
 */
int kupsmi_thermostat_barostat(void* stream, int phase, int n_sys, const int32_t* sys_off,
                               float* pos, float* mom, const float* mass, double* tril,
                               const double* dt, const double* kT, const double* tau,
                               const double* baro, const double* virial,
                               const double* kin_energy, uint64_t seed, int64_t step,
                               const int64_t* step_dev);

/*
 * *step_dev += n on the stream: the device-side step counter that
 * kupsmi_integrate / kupsmi_thermostat_barostat read when they are recorded
 * in a HIP graph (the host-side step of a graph is fixed at capture).  No
 * counterpart in the reference (its step counter is a JAX array).
 */
int kupsmi_step_advance(void* stream, int64_t* step_dev, int64_t n);

/*
 * Rigid-molecule GCMC: n_steps trial moves of every chain, one workgroup per
 * chain with its host framework and guests resident in LDS.  Replaces the
 * MCMC cycle of kups_mcmc_rigid (application/simulations/mcmc_rigid.py:
 * 281-307): per step ONE move type for the whole batch (translation,
 * rotation, reinsertion, or exchange with one insert/delete coin;
 * core/propagator.py:244-273, mcmc/moves.py:1078-1230), proposals of
 * moves.py:122-295 and 573-835, Delta E from LJ + Ewald real-space pairs of
 * the changed sites, the structure-factor update S' = S + dS
 * (ewald.py:442-486), the self/exclusion energy of an inserted or deleted
 * molecule and the LJ tail change, acceptance log p = -dE/kT + fugacity term
 * (mcmc/probability.py:60-225) against log U in FP64, and the
 * acceptance-target step-width schedule (core/parameter_scheduler.py:161-208).
 * Mixtures (mcmc/moves.py:573-773, mcmc/probability.py:117-173, 228-256):
 * n_motif <= 4 guest species; an insertion or deletion picks the species
 * uniformly per chain, a deletion then one guest of that species (a no-op when
 * there is none), and the fugacity term is ln(a_t V) - ln(N_t + 1) /
 * -ln(a_t V) + ln N_t with the chosen species' activity and count;
 * translation, rotation and reinsertion move a guest of any species.
 *   host_xyzq      [n_host_total x 4] f32 (x, y, z, charge), chains by host_off
 *   motif_*        the guest templates [n_motif x m (x 3)], centred; a species
 *                  with fewer than m sites leaves its last rows unused
 *   motif_nsite    HOST array[n_motif]: sites of each species (1..m <= 4)
 *   mol_const, log_activity  [n_chain x n_motif] f64 (self + exclusion energy
 *                  of one molecule, ln(fugacity / kT))
 *   guest_motif    [n_chain x gmax] int32: species of each guest slot, in/out
 *   kn / pk        per chain k list: nk packed (n0 | (n1+64) << 8 | (n2+64) << 16)
 *                  and FP64 prefactors P(k) (0 = padding); nm = max |n|
 *   guest          [n_chain x gmax x m x 3] f64 (rigid-body moves in FP64, so molecules
 *                  stay rigid), n_guest, counts[n_chain x n_species],
 *                  structure_factor [n_chain x nk x 2] f64, energy, width[n_chain x 4],
 *                  hist[n_chain x 4 x 4], hist_idx[n_chain x 4]: state, in/out
 *   move_weights   HOST array[4]: unnormalised weights of translation, rotation,
 *                  reinsertion, exchange (default 1, 1, 1, 3 in kUPS's mcmc app)
 *   width_upper    HOST array[4]: upper bounds of the adapted step widths
 *   stats          [n_chain x 12] int64: trials[4], accepted[4] (moves that changed the
 *                  configuration; a move of an empty chain is accepted but not counted),
 *                  insertions, deletions, trials that evaluated any site (a displacement or
 *                  deletion of an empty chain evaluates none), site evaluations (sites of
 *                  the old and the new configuration summed over trials)
 *   trace          optional [n_trace x n_steps x 4] f64 (move, dE, log p, accepted)
 * Random numbers: Philox4x32-10 keyed by seed, counters (chain, step, tag); the
 * first trial's step counter is step0, or *step_dev when step_dev is not NULL
 * (a device int64 that a caller advances between replays of a captured HIP
 * graph, e.g. with kupsmi_step_advance).
 */
int kupsmi_gcmc_run(void* stream, int n_chain, int n_steps, int64_t step0, uint64_t seed,
                    int max_host, const int32_t* host_off, const float* host_xyzq,
                    const int32_t* host_species, const double* tril, int gmax, int m,
                    int n_motif, const int32_t* motif_nsite, const float* motif_pos,
                    const float* motif_q, const int32_t* motif_species,
                    int n_species, const double* sigma, const double* epsilon, double rc_lj,
                    double rc_coul, const double* alpha, const double* tail,
                    const double* mol_const, const double* kT, const double* log_activity,
                    int nk, int nm, const int32_t* kn, const double* pk,
                    const double* move_weights, const double* width_upper, double* guest,
                    int32_t* n_guest, int32_t* guest_motif, int32_t* counts,
                    double* structure_factor,
                    double* energy, float* width, uint32_t* hist, int32_t* hist_idx,
                    long long* stats, int n_trace, double* trace, const float* prop_cdf,
                    const int32_t* prop_guide, int prop_nv, int* sample_count, int sample_cap,
                    int32_t* sample_row, float* sample_y, const int64_t* step_dev);

/* ---- learned insertion proposal (no counterpart in the reference; it enters
 * kUPS's MCMC through the ChangesFn proposal log ratio,
 * core/propagator.py:230-235, added to the acceptance at 332-334) ----
 * kupsmi_gcmc_run above takes the proposal table prop_cdf [n_chain x nv^3]
 * (cumulative q over the fractional voxel grid, last entry 1) and its guide
 * table prop_guide [n_chain x nv^3] (both from kupsmi_proposal_table): an insertion
 * draws voxel v ~ q and a uniform offset inside it and adds
 * -ln(q(v) nv^3) to log p; a deletion adds +ln(q(v) nv^3) for the voxel of the
 * molecule's centre (detailed balance).  NULL: uniform insertion.  With
 * sample_count != NULL every insertion trial appends (chain * nv^3 + voxel,
 * -dE/kT) to sample_row/sample_y (up to sample_cap): the training set. */

/* number of MLP parameters (32 -> 128 -> 128 -> 128 -> 1, SiLU): 37377 */
int kupsmi_mlp_param_count(void);

/* 32-bin radial host-density histogram (r < rmax, normalised by shell volume
 * x rho0) around every voxel centre of every chain, BF16 [n_chain x nv^3 x 32]. */
int kupsmi_voxel_descriptors(void* stream, int n_chain, int nv, int max_host,
                             const int32_t* host_off, const float* host_xyzq,
                             const double* tril, float rmax, float rho0, uint16_t* desc);

/* MLP logits of n_rows descriptors on BF16 MFMA (FP32 accumulate).
 * weights_bf16 / weights_f32: the flat parameter vector, BF16 and the same
 * values as FP32. */
int kupsmi_mlp_infer(void* stream, int n_rows, const uint16_t* desc,
                     const uint16_t* weights_bf16, const float* weights_f32, float* logits);

/* One training step on a minibatch of descriptor rows and targets:
 * loss = 1/2 mean (f(x) - y)^2, forward/backward with the BF16-rounded
 * weights, FP32 gradients (written to grad_out if non-NULL), Adam on the FP32
 * master weights (apply != 0) and refresh of both rounded copies.
 * loss_out (optional): per-workgroup loss sums. */
int kupsmi_mlp_train_step(kupsmi_ctx* ctx, void* stream, int n_samples, const int32_t* rows,
                          const float* targets, const uint16_t* desc, float* master,
                          float* m1, float* m2, float* weights_f32, uint16_t* weights_bf16,
                          float lr, int adam_step, int apply, float* grad_out,
                          float* loss_out);

/* Rounded copies (FP32 and BF16) of the FP32 master weights. */
int kupsmi_mlp_round_weights(void* stream, const float* master, float* weights_f32,
                             uint16_t* weights_bf16);

/* Proposal table: cdf of q = (1 - mix) softmax(logits / temperature) + mix / nvox
 * per chain (FP64 sums, fixed order), and the guide table of the sampler:
 * guide[j] = first voxel v with j / nvox < cdf[v] (an insertion starts its
 * search for u there, so it reads ~2 entries instead of log2(nvox)). */
int kupsmi_proposal_table(void* stream, int n_chain, int nvox, const float* logits,
                          float temperature, float mix, float* cdf, int32_t* guide);

#ifdef __cplusplus
}
#endif
#endif /* KUPSMI_H */
