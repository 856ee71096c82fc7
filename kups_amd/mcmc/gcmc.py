"""Rigid-molecule GCMC driver (application/simulations/mcmc_rigid.py).

``RigidGCMC`` builds the state of kups_mcmc_rigid for a batch of chains (hosts
with the reference's exclusion semantics, one or several guest species,
TraPPE LJ with Lorentz-Berthelot mixing and the global tail correction, Ewald
with the reference's estimator, Peng-Robinson mixture fugacities) and
advances all chains with the LDS-resident HIP kernel ``kupsmi_gcmc_run``.  Set-up that the reference
does on the host (estimator, k vectors, fugacity, tail matrix) is done on the
host here too; the initial structure factors and energies come from the GPU
Ewald and pair kernels.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .. import _lib
from ..core.batch import ParticleBatch, SystemBatch, offsets_from_sizes, segment_sum
from ..core.constants import BOLTZMANN_CONSTANT, PASCAL, TO_STANDARD_UNITS
from ..potential.ewald import EwaldParameters, EwaldPotential, estimate_ewald_parameters, kvec_grid
from ..potential.lennard_jones import LennardJonesParameters, global_tail_correction_pressure
from .fugacity import peng_robinson_log_fugacity

MOVES = ("translation", "rotation", "reinsertion", "exchange")


@dataclass
class Motif:
    """Rigid guest template (positions centred on the geometric centre)."""

    species: tuple[str, ...]
    positions: np.ndarray
    charges: np.ndarray
    critical_temperature: float
    critical_pressure: float
    acentric_factor: float

    @classmethod
    def from_dict(cls, d) -> "Motif":
        p = np.asarray(d["positions"], float)
        return cls(tuple(d["symbols"]), p - p.mean(axis=0), np.asarray(d["charges"], float),
                   float(d["critical_temperature"]), float(d["critical_pressure"]),
                   float(d["acentric_factor"]))


def molecule_constant(motif: Motif, alpha: float) -> float:
    """E_self + E_excl of one rigid molecule: -K alpha/sqrt(pi) sum q^2
    (ewald.py:277-302) + sum_{i<j} K q_i q_j (erfc(alpha r) - 1)/r (328-346)."""
    from scipy.special import erfc

    q = motif.charges
    e = -TO_STANDARD_UNITS * alpha / math.sqrt(math.pi) * float(q @ q)
    for i in range(len(q)):
        for j in range(i + 1, len(q)):
            r = float(np.linalg.norm(motif.positions[i] - motif.positions[j]))
            e += TO_STANDARD_UNITS * q[i] * q[j] * (erfc(alpha * r) - 1.0) / r
    return e


def tail_matrix(lj: LennardJonesParameters, volume: float) -> np.ndarray:
    """T_ab with E_tail = sum_ab T_ab N_a N_b (lennard_jones.py:276-313, pairs with
    eps > 0)."""
    t1 = (lj.sigma / lj.cutoff) ** 3
    T = (8 / 3) * math.pi / volume * lj.epsilon * lj.sigma**3 * (t1**3 / 3 - t1)
    return T * (lj.epsilon > 0)


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform rotation matrix from a uniform unit quaternion (Shoemake's
    subgroup algorithm, Graphics Gems III), as Quaternion.random."""
    u1, u2, u3 = rng.random(3)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    w, x, y, z = (a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3))
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


class RigidGCMC:
    """GCMC of rigid guest species in a batch of rigid hosts.

    ``motif`` is one Motif or a sequence of them (a mixture, at most 4
    species of at most 4 sites), ``guest_species`` the LJ species of each
    site (a tuple, or one tuple per motif); ``composition`` the gas-phase
    mole fractions and ``interaction`` the Peng-Robinson binary interaction
    parameters of a mixture (application/mcmc/data.py:122-125, 380-391)."""

    def __init__(self, host_positions, host_charges, host_species, host_sizes, lattices,
                 pressure, temperature, motif, lj: LennardJonesParameters,
                 guest_species, *, ewald_precision: float = 1e-6,
                 ewald_cutoff: float = 12.0, weights=(1.0, 1.0, 1.0, 3.0), gmax: int = 64,
                 seed: int = 0, tail_correction: bool = True, composition=None,
                 interaction=None, device=None):
        dev = device or _lib.require_gpu()
        self.dev = dev
        motifs = [motif] if isinstance(motif, Motif) else list(motif)
        if isinstance(motif, Motif):
            guest_species = [tuple(guest_species)]
        guest_species = [tuple(int(x) for x in g) for g in guest_species]
        if not 1 <= len(motifs) <= 4 or len(guest_species) != len(motifs):
            raise ValueError("1..4 motifs, with the LJ species of each one's sites")
        for m_, g_ in zip(motifs, guest_species):
            if len(g_) != len(m_.charges) or not 1 <= len(m_.charges) <= 4:
                raise ValueError("a motif has 1..4 sites, each with an LJ species")
        self.motifs = motifs
        self.motif = motifs[0]
        self.n_motif = len(motifs)
        self.motif_nsite = np.array([len(m_.charges) for m_ in motifs], np.int32)
        self.m = int(self.motif_nsite.max())
        self.lj = lj
        self.seed = int(seed)
        self.step = 0
        self.gmax = int(gmax)
        self.weights = np.asarray(weights, float)
        sizes = np.asarray(host_sizes, np.int64)
        self.sizes = sizes
        n = len(sizes)
        self.n_chain = n
        lat = np.asarray(lattices, float).reshape(n, 3, 3)
        tril = lat[:, [0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]]
        vol = np.abs(np.linalg.det(lat))
        self.volume = vol
        self.temperature = np.broadcast_to(np.asarray(temperature, float), (n,)).copy()
        kT = self.temperature * BOLTZMANN_CONSTANT
        # fugacity (mcmc/data.py:286-290, 380-391): ln a_t = ln f_t - ln kT, f in
        # eV / A^3, f_t from the Peng-Robinson mixture at the gas composition
        # (composition and interaction may also be given per chain)
        mt = self.n_motif
        frac = np.broadcast_to(np.full(mt, 1.0 / mt) if composition is None else
                               np.asarray(composition, float), (n, mt))
        if np.any(np.abs(frac.sum(1) - 1) > 1e-9) or np.any(frac < 0):
            raise ValueError("composition: one mole fraction per motif, summing to 1")
        kij = np.broadcast_to(np.zeros((mt, mt)) if interaction is None else
                              np.asarray(interaction, float), (n, mt, mt))
        crit = ([m_.critical_pressure for m_ in motifs], [m_.critical_temperature for m_ in motifs],
                [m_.acentric_factor for m_ in motifs])
        lnf = np.array([peng_robinson_log_fugacity(p, t, *crit, frac[s], kij[s])[0]
                        for s, (p, t) in enumerate(zip(np.broadcast_to(pressure, (n,)),
                                                       self.temperature))]).reshape(n, -1)
        self.log_activity = lnf + math.log(PASCAL) - np.log(kT)[:, None]
        # Ewald: fixed real-space cutoff -> alpha and k_max (ewald.py:1254-1277)
        hq = np.asarray(host_charges, float)
        off = np.concatenate([[0], np.cumsum(sizes)])
        est = [estimate_ewald_parameters(hq[off[s]:off[s + 1]] if sizes[s] else motifs[0].charges,
                                         tril[s], real_cutoff=ewald_cutoff,
                                         epsilon_total=ewald_precision) for s in range(n)]
        self.alpha = np.array([e["alpha"] for e in est])
        self.kmax = np.array([e["k_max"] for e in est])
        self.ewald_cutoff = float(ewald_cutoff)
        kl = [kvec_grid(tril[s], self.kmax[s]) for s in range(n)]
        nk = max(len(k) for k in kl)
        self.nm = int(max(np.abs(k).max() for k in kl))
        if self.nm > 12:
            raise ValueError("k vectors beyond |n| = 12 are not supported")
        kn = np.zeros((n, nk), np.int32)
        pk = np.zeros((n, nk))
        gidx = np.zeros((n, nk), np.int64)  # index into the (nm, nm, nm) box grid
        w = 2 * self.nm + 1
        for s in range(n):
            k = kl[s]
            kn[s, :len(k)] = k[:, 0] | ((k[:, 1] + 64) << 8) | ((k[:, 2] + 64) << 16)
            kv = 2 * np.pi * k @ np.linalg.inv(lat[s]).T
            k2 = (kv**2).sum(-1)
            P = np.where(k2 > 0, 2 * np.pi / vol[s] * np.exp(-k2 / (4 * self.alpha[s]**2))
                         / np.where(k2 > 0, k2, 1) * np.where(k[:, 0] == 0, 1.0, 2.0), 0.0)
            pk[s, :len(k)] = P
            gidx[s, :len(k)] = (k[:, 0] * w + (k[:, 1] + self.nm)) * w + (k[:, 2] + self.nm)
        self.nk = nk
        self.k_counts = np.array([len(k) for k in kl])  # k vectors of each chain
        self._kgrid_index = gidx
        # species: host types then guest types
        self.n_species = lj.n_species
        self.guest_species_per_motif = [np.asarray(g_, np.int32) for g_ in guest_species]
        self.guest_species = self.guest_species_per_motif[0]
        self.mol_const = np.array([[molecule_constant(m_, a) for m_ in motifs] for a in self.alpha])
        self.tail = np.stack([tail_matrix(lj, v) for v in vol])
        self.tail_correction = bool(tail_correction)
        if not tail_correction:
            self.tail[:] = 0.0
        counts = np.zeros((n, self.n_species), np.int32)
        hs = np.asarray(host_species, np.int32)
        for s in range(n):
            counts[s] = np.bincount(hs[off[s]:off[s + 1]], minlength=self.n_species)
        # translation width <= half the smallest perpendicular width of the batch
        # (application/simulations/mcmc_rigid.py:357-371); others unbounded
        self.upper = np.array([self._min_half_box(lat), 1e30, 1e30, 1e30])
        t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)
        f32, f64, i32 = torch.float32, torch.float64, torch.int32
        M = self.m
        mpos = np.zeros((self.n_motif, M, 3))
        mq = np.zeros((self.n_motif, M))
        msp = np.zeros((self.n_motif, M), np.int32)
        for k, (m_, g_) in enumerate(zip(motifs, self.guest_species_per_motif)):
            mpos[k, :len(g_)] = m_.positions
            mq[k, :len(g_)] = m_.charges
            msp[k, :len(g_)] = g_
        self.d = dict(
            host_off=t(off.astype(np.int32), i32),
            host=t(np.concatenate([np.asarray(host_positions, float), hq[:, None]], 1), f32),
            host_species=t(hs, i32), tril=t(tril, f64),
            motif_pos=t(mpos, f32), motif_q=t(mq, f32),
            motif_species=t(msp, i32), sigma=t(lj.sigma, f64),
            epsilon=t(lj.epsilon, f64), alpha=t(self.alpha, f64), tail=t(self.tail, f64),
            mol_const=t(self.mol_const, f64), kT=t(kT, f64), log_activity=t(self.log_activity, f64),
            kn=t(kn, i32), pk=t(pk, f64),
            guest=torch.zeros((n, self.gmax, M, 3), dtype=f64, device=dev),
            n_guest=torch.zeros(n, dtype=i32, device=dev),
            guest_motif=torch.zeros((n, self.gmax), dtype=i32, device=dev), counts=t(counts, i32),
            sk=torch.zeros((n, nk, 2), dtype=f64, device=dev),
            energy=torch.zeros(n, dtype=f64, device=dev),
            width=torch.full((n, 4), 0.1, dtype=f32, device=dev),
            hist=torch.zeros((n, 4, 4), dtype=torch.int32, device=dev),
            hist_idx=torch.zeros((n, 4), dtype=i32, device=dev),
            stats=torch.zeros((n, 12), dtype=torch.int64, device=dev),
        )
        # move weights and width bounds are host arrays of the entry point
        self._w_host = np.ascontiguousarray(self.weights, dtype=np.float64)
        self._u_host = np.ascontiguousarray(self.upper, dtype=np.float64)
        self._n_host = np.ascontiguousarray(self.motif_nsite, dtype=np.int32)
        self.max_host = int(sizes.max()) if n else 0
        self.recompute_state()

    @staticmethod
    def _min_half_box(lat) -> float:
        lat = np.asarray(lat).reshape(-1, 3, 3)
        return min(float((abs(np.linalg.det(L)) /
                          np.linalg.norm(np.cross(L[[1, 0, 0]], L[[2, 2, 1]]), axis=-1)).min())
                   for L in lat) / 2.0

    def sr_intramolecular(self) -> np.ndarray:
        """E_sr over the intramolecular pairs of one guest of each motif,
        (n_chain, n_motif).  The reference's full evaluation contains it for
        every molecule present at that time, its incremental Delta E does not."""
        from scipy.special import erfc

        out = np.zeros((self.n_chain, self.n_motif))
        for t, m_ in enumerate(self.motifs):
            q, p = m_.charges, m_.positions
            for i in range(len(q)):
                for j in range(i + 1, len(q)):
                    r = float(np.linalg.norm(p[i] - p[j]))
                    out[:, t] += TO_STANDARD_UNITS * q[i] * q[j] * erfc(self.alpha * r) / r
        return out

    def motif_counts(self) -> np.ndarray:
        """Guests of each motif in every chain, (n_chain, n_motif)."""
        ng = self.d["n_guest"].cpu().numpy()
        gm = self.d["guest_motif"].cpu().numpy()
        out = np.zeros((self.n_chain, self.n_motif), np.int64)
        for s in range(self.n_chain):
            out[s] = np.bincount(gm[s, :ng[s]], minlength=self.n_motif)
        return out

    # ---- full evaluation (initial energies, structure factors, checks)
    def configuration(self):
        """Host + guest atoms of every chain as a ParticleBatch (chain-sorted)."""
        d = self.d
        n = self.n_chain
        ng = d["n_guest"].cpu().numpy()
        gm = d["guest_motif"].cpu().numpy()
        host = d["host"].cpu().numpy()
        hoff = d["host_off"].cpu().numpy()
        hspec = d["host_species"].cpu().numpy()
        gpos = d["guest"].cpu().numpy()
        mq = [m_.charges for m_ in self.motifs]
        pos, q, sp, grp, sizes = [], [], [], [], []
        for s in range(n):
            h = host[hoff[s]:hoff[s + 1]]
            ms = gm[s, :ng[s]]
            ns = self.motif_nsite[ms]
            g = np.concatenate([gpos[s, k, :ns[k]] for k in range(ng[s])]) if ng[s] else \
                np.zeros((0, 3))
            pos += [h[:, :3], g]
            q += [h[:, 3]] + [mq[t] for t in ms]
            sp += [hspec[hoff[s]:hoff[s + 1]]] + [self.guest_species_per_motif[t] for t in ms]
            grp += [np.full(len(h), -1), np.repeat(np.arange(ng[s]), ns)]
            sizes.append(len(h) + len(g))
        return (np.concatenate(pos), np.concatenate(q), np.concatenate(sp).astype(np.int32),
                np.concatenate(grp).astype(np.int32), np.asarray(sizes))

    def full_energy(self) -> np.ndarray:
        """Absolute energy of every chain with the reference's composition
        (application/simulations/mcmc_rigid.py:281-307): Ewald over the atomic
        view (all pairs, intramolecular included, ewald.py:946-967) + self +
        exclusion correction of every guest + LJ with host-host and
        intramolecular pairs excluded + global LJ tail."""
        pos, q, sp, grp, sizes = self.configuration()
        dev = self.dev
        n = self.n_chain
        if len(pos) == 0:
            return np.zeros(n)
        system = np.repeat(np.arange(n), sizes).astype(np.int32)
        off = offsets_from_sizes(sizes, dev)
        pb = ParticleBatch(torch.as_tensor(pos, dtype=torch.float32, device=dev),
                           torch.as_tensor(system, device=dev),
                           species=torch.as_tensor(sp, device=dev),
                           charges=torch.as_tensor(q, dtype=torch.float32, device=dev))
        sb = SystemBatch(self.d["tril"])
        ep = EwaldParameters(self.alpha, np.full(n, self.ewald_cutoff), self.kmax,
                             (self.nm, self.nm, self.nm))
        out = EwaldPotential(ep)(pb, sb)
        # LJ: exclusion id = group (hosts share one id per chain)
        big = np.int32(2**30)
        excl = np.where(grp < 0, big, grp + np.repeat(np.arange(n), sizes) * self.gmax).astype(np.int32)
        excl_t = torch.as_tensor(excl, device=dev)
        sig = torch.as_tensor(self.lj.sigma, device=dev)
        eps = torch.as_tensor(self.lj.epsilon, device=dev)
        rc = torch.full((n,), self.lj.cutoff, dtype=torch.float64, device=dev)
        f = torch.zeros((len(pos), 3), dtype=torch.float32, device=dev)
        e2 = torch.zeros((n, 2), dtype=torch.float64, device=dev)
        _lib.check(_lib.load().kupsmi_pair_forces(
            _lib.ctx(), _lib.stream_ptr(), n, _lib.ptr(off), len(pos), int(sizes.max()),
            _lib.ptr(pb.positions), None, _lib.ptr(pb.species), self.n_species, _lib.ptr(sig),
            _lib.ptr(eps), _lib.ptr(rc), _lib.ptr(excl_t), None, None, None, _lib.ptr(sb.tril), 0,
            _lib.ptr(f), _lib.ptr(e2), None, None))
        nt = self.motif_counts()
        counts = self.d["counts"].cpu().numpy().astype(float)
        e_tail = np.einsum("sab,sa,sb->s", self.tail, counts, counts)
        # intramolecular exclusion correction of every guest molecule; the
        # atomic view above already holds E_sr(intra) of each guest, as the
        # reference's full evaluation does (ewald.py:946-967 over all pairs)
        e_excl = (nt * (self.mol_const - self._self_mol())).sum(1)
        return (out.short_range + out.long_range + out.self_energy).cpu().numpy() + e_excl + \
            e2[:, 0].cpu().numpy() + e_tail

    def _self_mol(self):
        """Ewald self energy of one guest of each motif, (n_chain, n_motif)."""
        qq = np.array([float(m_.charges @ m_.charges) for m_ in self.motifs])
        return -TO_STANDARD_UNITS * self.alpha[:, None] / math.sqrt(math.pi) * qq[None, :]

    def recompute_state(self):
        """Structure factors of the current configuration (GPU) and the tracked
        energy set to the full evaluation (the composition origin moves to the
        current loading)."""
        self._ng_ref = self.motif_counts()
        if self._structure_factors():
            self.d["energy"].copy_(torch.as_tensor(self.full_energy(), device=self.dev))

    def resync(self):
        """Re-evaluate S(k) and the tracked energies in full while keeping the
        composition of the incremental bookkeeping: a guest inserted since the
        composition origin carries no intramolecular real-space term, one
        removed keeps it (the reference's tracked-energy convention).  Clears
        the rounding that FP32 pair terms leave in the FP64 running energy,
        which matters after a start from overlapping molecules, where the
        early energies are astronomically large."""
        nt = self.motif_counts()
        if self._structure_factors():
            e = self.full_energy() - ((nt - self._ng_ref) * self.sr_intramolecular()).sum(1)
            self.d["energy"].copy_(torch.as_tensor(e, device=self.dev))

    def _structure_factors(self) -> bool:
        """S(k) of the current configuration into the chain state; False (and
        S = 0, E = 0) when the batch holds no atom."""
        pos, q, sp, grp, sizes = self.configuration()
        dev = self.dev
        n = self.n_chain
        if len(pos) == 0:
            self.d["sk"].zero_()
            self.d["energy"].zero_()
            return False
        off = offsets_from_sizes(sizes, dev)
        p = torch.as_tensor(pos, dtype=torch.float32, device=dev)
        qq = torch.as_tensor(q, dtype=torch.float32, device=dev)
        alpha = torch.as_tensor(self.alpha, dtype=torch.float64, device=dev)
        nm = self.nm
        sk = torch.empty((n, (nm + 1) * (2 * nm + 1) ** 2, 2), dtype=torch.float64, device=dev)
        _lib.check(_lib.load().kupsmi_ewald_recip(
            _lib.ctx(), _lib.stream_ptr(), n, _lib.ptr(off), len(pos), int(sizes.max()),
            _lib.ptr(p), _lib.ptr(qq), _lib.ptr(self.d["tril"]), _lib.ptr(alpha), None, nm, nm, nm,
            0, None, None, None, _lib.ptr(sk)))
        gi = torch.as_tensor(self._kgrid_index, device=dev)
        self.d["sk"].copy_(torch.gather(sk, 1, gi[:, :, None].expand(-1, -1, 2)))
        return True

    def energy_error(self) -> np.ndarray:
        """Tracked energy minus a full re-evaluation, per chain (eV).  The full
        evaluation holds E_sr(intra) of every guest, the incremental Delta E of
        the kernel does not, so guests added since the last full evaluation are
        accounted for."""
        nt = self.motif_counts()
        full = self.full_energy() - ((nt - self._ng_ref) * self.sr_intramolecular()).sum(1)
        return self.d["energy"].cpu().numpy() - full

    def insert_random(self, n_init, seed: int = 0):
        """Place n_init[s] molecules in chain s (n_init[s, t] of motif t when
        given per motif), each at a uniform fractional centre with a uniformly
        random orientation and no overlap test (application/mcmc/data.py:
        293-340), then rebuild S(k) and the tracked energies from a full
        evaluation."""
        n_init = np.asarray(n_init, np.int64)
        if n_init.ndim < 2:
            n_init = np.stack([np.broadcast_to(n_init, (self.n_chain,))] +
                              [np.zeros(self.n_chain, np.int64)] * (self.n_motif - 1), 1)
        n_init = np.broadcast_to(n_init, (self.n_chain, self.n_motif))
        ng = self.d["n_guest"].cpu().numpy()
        if np.any(ng + n_init.sum(1) > self.gmax):
            raise ValueError(f"initial loading exceeds the {self.gmax}-molecule buffer")
        rng = np.random.default_rng(seed)
        guest = self.d["guest"].cpu().numpy()
        gm = self.d["guest_motif"].cpu().numpy()
        counts = self.d["counts"].cpu().numpy()
        lat = np.zeros((self.n_chain, 3, 3))
        tr = self.d["tril"].cpu().numpy()
        lat[:, [0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]] = tr
        for s in range(self.n_chain):
            for t, m_ in enumerate(self.motifs):
                k = len(m_.charges)
                for _ in range(int(n_init[s, t])):
                    com = rng.uniform(-0.5, 0.5, 3) @ lat[s]
                    guest[s, ng[s]] = 0.0
                    guest[s, ng[s], :k] = com + m_.positions @ random_rotation(rng).T
                    gm[s, ng[s]] = t
                    ng[s] += 1
                    np.add.at(counts[s], self.guest_species_per_motif[t], 1)
        self.d["guest"].copy_(torch.as_tensor(guest))
        self.d["guest_motif"].copy_(torch.as_tensor(gm))
        self.d["n_guest"].copy_(torch.as_tensor(ng))
        self.d["counts"].copy_(torch.as_tensor(counts))
        self.recompute_state()

    def guest_pressure(self) -> dict:
        """Pressure terms of every chain as the reference's MCMC logging
        computes them (application/simulations/mcmc_rigid.py:385-431):

        * ``potential``: molecular virial pressure of the guest-only system
          (observables/stress.py:84-120, molecules at their geometric
          centres): P = (tr W - sum_i d_i . F_i) / 3V with W and F of LJ,
          real-space and reciprocal Ewald over the guests alone, d_i the
          minimum-image offset of site i from its molecule's centre.  Pair
          terms inside one molecule drop out of this trace exactly
          (r_ij = d_i - d_j), so those pairs are excluded here;
        * ``tail``: the global LJ tail pressure of all species counts
          (lennard_jones.py:316-340);
        * ``ideal``: N kT / V over the molecules (observables/pressure.py).

        Returns FP64 numpy arrays (n_chain,) under those keys (eV/A^3)."""
        d = self.d
        dev = self.dev
        n = self.n_chain
        ng = d["n_guest"].long()
        vol = torch.as_tensor(self.volume, dtype=torch.float64, device=dev)
        kT = torch.as_tensor(self.temperature * BOLTZMANN_CONSTANT, dtype=torch.float64, device=dev)
        ideal = (ng.double() * kT / vol).cpu().numpy()
        counts = d["counts"].cpu().numpy()
        tail = global_tail_correction_pressure(self.lj, counts, self.volume) * \
            float(self.tail_correction)
        pot = np.zeros(n)
        total = int(ng.sum().item())
        if total == 0:
            return {"potential": pot, "tail": tail, "ideal": ideal}
        # guest-only configuration, chain-sorted, molecules contiguous
        keep = torch.arange(self.gmax, device=dev)[None, :] < ng[:, None]          # (n, gmax)
        g = d["guest"][keep]                                                      # (mols, M, 3)
        mot = d["guest_motif"][keep].long()                                       # (mols,)
        chain_of_mol = torch.arange(n, device=dev)[:, None].expand(n, self.gmax)[keep]
        nsite = torch.as_tensor(self.motif_nsite, device=dev).long()[mot]         # (mols,)
        real = torch.arange(self.m, device=dev)[None, :] < nsite[:, None]         # (mols, M)
        tril = d["tril"]
        L = torch.zeros((n, 3, 3), dtype=torch.float64, device=dev)
        L[:, [0, 1, 1, 2, 2, 2], [0, 0, 1, 0, 1, 2]] = tril
        Linv = torch.linalg.inv(L)
        Lm, Lim = L[chain_of_mol], Linv[chain_of_mol]
        # offsets from the first site on the minimum image, then from the centre
        rel = g - g[:, :1]
        f = torch.einsum("mkc,mcd->mkd", rel, Lim)
        rel = torch.einsum("mkc,mcd->mkd", f - torch.round(f), Lm)
        rel = torch.where(real[:, :, None], rel, 0.0)
        centre = rel.sum(1, keepdim=True) / nsite[:, None, None].double()
        dvec = (rel - centre)[real]                                               # (atoms, 3)
        pos = g[real]
        atoms = pos.shape[0]
        sizes = segment_sum(nsite.double(), chain_of_mol, n).round().long().cpu().numpy()
        off = offsets_from_sizes(sizes, dev)
        q = d["motif_q"][mot][real].contiguous()
        sp = d["motif_species"][mot][real].contiguous()
        mol_id = torch.arange(g.shape[0], device=dev, dtype=torch.int32)[:, None] \
            .expand(-1, self.m)[real].contiguous()
        chain_of_atom = chain_of_mol[:, None].expand(-1, self.m)[real]
        p32 = pos.to(torch.float32).contiguous()
        forces = torch.zeros((atoms, 3), dtype=torch.float32, device=dev)
        e2 = torch.zeros((n, 2), dtype=torch.float64, device=dev)
        vir = torch.zeros((n, 3, 3), dtype=torch.float64, device=dev)
        alpha = torch.as_tensor(self.alpha, dtype=torch.float64, device=dev)
        rc_c = torch.full((n,), self.ewald_cutoff, dtype=torch.float64, device=dev)
        rc_lj = torch.full((n,), self.lj.cutoff, dtype=torch.float64, device=dev)
        sig = torch.as_tensor(self.lj.sigma, dtype=torch.float64, device=dev)
        eps = torch.as_tensor(self.lj.epsilon, dtype=torch.float64, device=dev)
        lib = _lib.load()
        maxn = int(sizes.max())
        _lib.check(lib.kupsmi_pair_forces(
            _lib.ctx(), _lib.stream_ptr(), n, _lib.ptr(off), atoms, maxn, _lib.ptr(p32),
            _lib.ptr(q), _lib.ptr(sp), self.n_species, _lib.ptr(sig), _lib.ptr(eps), _lib.ptr(rc_lj),
            _lib.ptr(mol_id), _lib.ptr(alpha), _lib.ptr(rc_c), _lib.ptr(mol_id), _lib.ptr(tril), 0,
            _lib.ptr(forces), _lib.ptr(e2), _lib.ptr(vir), None))
        nm = self.nm
        elr = torch.zeros(n, dtype=torch.float64, device=dev)
        kmax = torch.as_tensor(self.kmax, dtype=torch.float64, device=dev)
        _lib.check(lib.kupsmi_ewald_recip(
            _lib.ctx(), _lib.stream_ptr(), n, _lib.ptr(off), atoms, maxn, _lib.ptr(p32), _lib.ptr(q),
            _lib.ptr(tril), _lib.ptr(alpha), _lib.ptr(kmax), nm, nm, nm, 1, _lib.ptr(forces),
            _lib.ptr(elr), _lib.ptr(vir), None))
        dF = (dvec * forces.double()).sum(-1)                                     # (atoms,)
        corr = segment_sum(dF, chain_of_atom, n)
        tr = torch.diagonal(vir, dim1=-2, dim2=-1).sum(-1)
        pot = ((tr - corr) / (3.0 * vol)).cpu().numpy()
        return {"potential": pot, "tail": tail, "ideal": ideal}

    _STATE = ("guest", "n_guest", "guest_motif", "counts", "sk", "energy", "width", "hist", "hist_idx", "stats")

    def snapshot(self) -> dict:
        """Device copy of the mutable chain state (guests, S(k), energies,
        schedulers, counters) and the step counter."""
        snap = {k: self.d[k].clone() for k in self._STATE}
        snap["_step"], snap["_ng_ref"] = self.step, self._ng_ref.copy()
        return snap

    def restore(self, snap: dict):
        for k in self._STATE:
            self.d[k].copy_(snap[k])
        self.step, self._ng_ref = snap["_step"], snap["_ng_ref"].copy()

    # ---- the MC cycle
    def run(self, n_steps: int, n_trace: int = 0, proposal=None, samples=None):
        """Advance every chain by n_steps trial moves.  ``proposal``: a
        LearnedProposal (its table biases insertions, with the exact density in
        the acceptance ratio) or None (uniform).  ``samples``: an InsertionSamples
        buffer that records (voxel row, -dE/kT) of every insertion trial."""
        trace = torch.zeros((n_trace, max(n_steps, 1), 4), dtype=torch.float64,
                            device=self.dev) if n_trace else None
        self._launch(n_steps, n_trace, trace, proposal, samples)
        self.step += int(n_steps)
        return trace

    def capture(self, n_steps: int, launches: int = 1, proposal=None) -> "GraphedTrials":
        """``launches`` x ``n_steps`` trial moves recorded in one HIP graph."""
        return GraphedTrials(self, n_steps, launches, proposal)

    def _launch(self, n_steps, n_trace=0, trace=None, proposal=None, samples=None, step_dev=None):
        d = self.d
        cdf = proposal.cdf if proposal is not None else None
        guide = proposal.guide if proposal is not None else None
        nv = proposal.nv if proposal is not None else (samples.nv if samples is not None else 0)
        if samples is not None and proposal is not None and samples.nv != proposal.nv:
            raise ValueError("samples and proposal use different voxel grids")
        _lib.check(_lib.load().kupsmi_gcmc_run(
            _lib.stream_ptr(), self.n_chain, int(n_steps), int(self.step), self.seed & (2**64 - 1),
            self.max_host, _lib.ptr(d["host_off"]), _lib.ptr(d["host"]), _lib.ptr(d["host_species"]),
            _lib.ptr(d["tril"]), self.gmax, self.m, self.n_motif, self._n_host.ctypes.data,
            _lib.ptr(d["motif_pos"]), _lib.ptr(d["motif_q"]),
            _lib.ptr(d["motif_species"]), self.n_species, _lib.ptr(d["sigma"]), _lib.ptr(d["epsilon"]),
            float(self.lj.cutoff), float(self.ewald_cutoff), _lib.ptr(d["alpha"]), _lib.ptr(d["tail"]),
            _lib.ptr(d["mol_const"]), _lib.ptr(d["kT"]), _lib.ptr(d["log_activity"]), self.nk,
            self.nm, _lib.ptr(d["kn"]), _lib.ptr(d["pk"]), self._w_host.ctypes.data,
            self._u_host.ctypes.data, _lib.ptr(d["guest"]), _lib.ptr(d["n_guest"]),
            _lib.ptr(d["guest_motif"]), _lib.ptr(d["counts"]), _lib.ptr(d["sk"]), _lib.ptr(d["energy"]), _lib.ptr(d["width"]), _lib.ptr(d["hist"]),
            _lib.ptr(d["hist_idx"]), _lib.ptr(d["stats"]), int(n_trace), _lib.ptr(trace),
            _lib.ptr(cdf), _lib.ptr(guide), int(nv), _lib.ptr(samples.count) if samples is not None else None,
            int(samples.capacity) if samples is not None else 0,
            _lib.ptr(samples.rows) if samples is not None else None,
            _lib.ptr(samples.y) if samples is not None else None, _lib.ptr(step_dev)))

    @property
    def loading(self) -> torch.Tensor:
        return self.d["n_guest"]

    @property
    def loading_per_motif(self) -> torch.Tensor:
        """Guests of each motif, (n_chain, n_motif) int32 on the device."""
        d = self.d
        live = torch.arange(self.gmax, device=self.dev)[None, :] < d["n_guest"][:, None]
        oh = torch.nn.functional.one_hot(d["guest_motif"].long(), self.n_motif)
        return (oh * live[:, :, None]).sum(1).to(torch.int32)

    @property
    def stats(self) -> dict:
        s = self.d["stats"].cpu().numpy()
        return {"trials": s[:, :4], "accepted": s[:, 4:8], "insertions": s[:, 8],
                "deletions": s[:, 9], "work_trials": s[:, 10], "site_evaluations": s[:, 11]}


class GraphedTrials:
    """``launches`` x ``n_steps`` trial moves of a RigidGCMC batch recorded in
    one HIP graph (a whole MC cycle replayed with one call).  The trial counter
    of the random streams is a device int64 that the recorded GCMC launches
    read and that a recorded kernel advances between them; it is set from
    ``gcmc.step`` before each replay, so a replay draws the same numbers as the
    same trials run eagerly.  The learned proposal's table, if given, is read
    from its buffers at replay time (rebuilding the table in place is seen by
    later replays)."""

    def __init__(self, gcmc: RigidGCMC, n_steps: int, launches: int = 1, proposal=None):
        if n_steps < 1 or launches < 1:
            raise ValueError("a captured cycle needs at least one launch of one trial")
        self.gcmc, self.n, self.launches = gcmc, int(n_steps), int(launches)
        self.step_dev = torch.zeros(1, dtype=torch.int64, device=gcmc.dev)
        lib = _lib.load()
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            for _ in range(self.launches):
                gcmc._launch(self.n, proposal=proposal, step_dev=self.step_dev)
                _lib.check(lib.kupsmi_step_advance(_lib.stream_ptr(), _lib.ptr(self.step_dev),
                                                   self.n))
        torch.cuda.synchronize()

    def replay(self) -> None:
        self.step_dev.fill_(int(self.gcmc.step))
        self.graph.replay()
        self.gcmc.step += self.n * self.launches


def make_co2_gcmc(hosts, *, gmax: int = 64, seed: int = 0, weights=(1.0, 1.0, 1.0, 3.0),
                  device=None) -> RigidGCMC:
    """GCMC of CO2 in the synthetic rigid hosts of kups_amd.workloads.gcmc_hosts."""
    return make_mixture_gcmc(hosts, ("CO2",), gmax=gmax, seed=seed, weights=weights,
                             device=device)


def mixture_species(adsorbates):
    """Motifs, LJ table (host types of kups_amd.workloads first, then each new
    guest label once, in order of appearance) and the LJ species of every
    motif's sites for the named adsorbates."""
    from .. import workloads as w

    motifs = [Motif.from_dict(w.ADSORBATES[a]) for a in adsorbates]
    labels = list(w.HOST_TYPES)
    for m_ in motifs:
        for sym in m_.species:
            if sym not in labels:
                labels.append(sym)
    lj = LennardJonesParameters.from_dict(12.0, {k: w.GCMC_LJ[k] for k in labels})
    guest_species = [tuple(labels.index(sym) for sym in m_.species) for m_ in motifs]
    return motifs, lj, guest_species


def make_mixture_gcmc(hosts, adsorbates, composition=None, *, gmax: int = 64, seed: int = 0,
                      weights=(1.0, 1.0, 1.0, 3.0), device=None) -> RigidGCMC:
    """GCMC of a gas mixture (names from kups_amd.workloads.ADSORBATES, e.g.
    ("CO2", "N2")) at the gas-phase mole fractions ``composition`` in the
    synthetic rigid hosts of kups_amd.workloads.gcmc_hosts."""
    motifs, lj, guest_species = mixture_species(adsorbates)
    if len(motifs) == 1:
        motifs, guest_species = motifs[0], guest_species[0]
    return RigidGCMC(hosts.positions, hosts.charges, hosts.species, hosts.sizes, hosts.lattices,
                     hosts.pressure, hosts.temperature, motifs, lj, guest_species, gmax=gmax,
                     seed=seed, weights=weights, composition=composition, device=device)
