"""ctypes bindings of liboracle.so plus numpy restatements (test infrastructure).

Every array argument is converted to a C-contiguous numpy array of the
required dtype; positions and the like are FP64 here.  See kups_oracle.c for
the reference citations of the compiled functions.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _lib.orc_neighborlist.restype = C.c_long
        _lib.orc_neighborlist_query.restype = C.c_long
        _lib.orc_kvecs.restype = C.c_long
        _lib.orc_coulomb_k.restype = C.c_double
        _lib.orc_md_vv_f32.restype = C.c_int
        _lib.orc_num_threads.restype = C.c_int
        _lib.orc_gamma_mt.restype = C.c_double
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def coulomb_k() -> float:
    """HARTREE*BOHR in eV*Angstrom (potential/classical/ewald.py:82)."""
    return lib().orc_coulomb_k()


def num_threads() -> int:
    return lib().orc_num_threads()


def neighborlist(pos, excl, sys_off, tril, cutoff):
    """Directed pair set (i, j, shift) of every system; see orc_neighborlist."""
    pos, excl, sys_off, tril, cutoff = _d(pos), _i(excl), _i(sys_off), _d(tril), _d(cutoff)
    nsys = len(sys_off) - 1
    L = lib()
    dummy_ij = np.zeros(2, np.int32)
    dummy_s = np.zeros(3, np.int32)
    n = L.orc_neighborlist(_p(pos), _p(excl), nsys, _p(sys_off), _p(tril), _p(cutoff),
                           C.c_long(0), _p(dummy_ij), _p(dummy_s))
    ij = np.zeros((max(n, 1), 2), np.int32)
    sh = np.zeros((max(n, 1), 3), np.int32)
    n2 = L.orc_neighborlist(_p(pos), _p(excl), nsys, _p(sys_off), _p(tril), _p(cutoff),
                            C.c_long(n), _p(ij), _p(sh))
    assert n2 == n
    return ij[:n], sh[:n]


def neighborlist_query(pos, excl, sys_off, tril, cutoff, rh_pos, rh_sys, rh_excl, remap=None):
    """rh-subset query (i, out_r, shift) of orc_neighborlist_query: forward
    edges in (r, i, image) order, then (with a remap) their reverses."""
    pos, excl, sys_off, tril, cutoff = _d(pos), _i(excl), _i(sys_off), _d(tril), _d(cutoff)
    rh_pos, rh_sys, rh_excl = _d(rh_pos).reshape(-1, 3), _i(rh_sys), _i(rh_excl)
    rm = None if remap is None else _i(remap)
    nsys, k = len(sys_off) - 1, len(rh_sys)
    L = lib()
    args = (_p(pos), _p(excl), nsys, _p(sys_off), k, _p(rh_pos), _p(rh_sys), _p(rh_excl),
            None if rm is None else _p(rm), _p(tril), _p(cutoff))
    dummy_ij, dummy_s = np.zeros(2, np.int32), np.zeros(3, np.int32)
    n = L.orc_neighborlist_query(*args, C.c_long(0), _p(dummy_ij), _p(dummy_s))
    ij = np.zeros((max(n, 1), 2), np.int32)
    sh = np.zeros((max(n, 1), 3), np.int32)
    n2 = L.orc_neighborlist_query(*args, C.c_long(n), _p(ij), _p(sh))
    assert n2 == n
    return ij[:n], sh[:n]


def lj_edges(pos, species, sys_off, tril, cutoff, sigma, eps, ij, shift):
    """LJ energy/forces/virial over a directed edge list (lennard_jones.py:149-178)."""
    pos, species, sys_off, tril, cutoff = _d(pos), _i(species), _i(sys_off), _d(tril), _d(cutoff)
    sigma, eps, ij, shift = _d(sigma), _d(eps), _i(ij), _i(shift)
    nsys = len(sys_off) - 1
    e = np.zeros(nsys)
    f = np.zeros_like(pos)
    w = np.zeros((nsys, 9))
    lib().orc_lj_edges(_p(pos), _p(species), nsys, _p(sys_off), _p(tril), _p(cutoff),
                       sigma.shape[0], _p(sigma), _p(eps), C.c_long(len(ij)), _p(ij), _p(shift),
                       _p(e), _p(f), _p(w))
    return e, f, w.reshape(nsys, 3, 3)


def lj_switch_factor(r2, rc, rtr):
    """S(r^2) and dS/d(r^2) of the pairwise switch (lennard_jones.py:210-220):
    S = 1 for r <= r_tr, (rc^2 - r^2)^2 (2 r^2 + rc^2 - 3 r_tr^2) / (rc^2 - r_tr^2)^3
    for r_tr < r < rc, 0 for r >= rc."""
    r2 = np.asarray(r2, np.float64)
    rc2, t2 = rc * rc, rtr * rtr
    div = (rc2 - t2) ** 3
    x = rc2 - r2
    band = (r2 > t2) & (r2 < rc2)
    with np.errstate(divide="ignore", invalid="ignore"):
        S = np.where(band, x * x * (2 * r2 + rc2 - 3 * t2) / div, np.where(r2 >= rc2, 0.0, 1.0))
        dS = np.where(band, 6 * x * (t2 - r2) / div, 0.0)
    return S, dS


def lj_switched_edges(pos, species, sys_off, tril, cutoff, rtr, sigma, eps, ij, shift):
    """LJ with the smooth pairwise switch of pair_tail_corrected_lennard_jones_energy
    (lennard_jones.py:200-222) over a directed edge list, numpy FP64: energy
    1/2 u S per directed edge, forces and virial of the switched energy, with
    lj_edges' conventions (d = r_j - r_i + shift @ L, virial -= d(uS)/d(r^2) d d)."""
    pos, species, sys_off = _d(pos), _i(species), _i(sys_off)
    tril, cutoff, rtr = _d(tril).reshape(-1, 6), _d(cutoff), _d(rtr)
    sigma, eps, ij, shift = _d(sigma), _d(eps), _i(ij).reshape(-1, 2), _d(shift).reshape(-1, 3)
    nsys = len(sys_off) - 1
    sys_of = np.repeat(np.arange(nsys), np.diff(sys_off))
    lat = np.zeros((nsys, 3, 3))
    lat[:, 0, 0] = tril[:, 0]
    lat[:, 1, :2] = tril[:, 1:3]
    lat[:, 2, :] = tril[:, 3:6]
    i, j = ij[:, 0], ij[:, 1]
    s = sys_of[i]
    d = pos[j] - pos[i] + np.einsum("ek,ekl->el", shift, lat[s])
    r2 = (d * d).sum(1)
    keep = r2 < cutoff[s] ** 2
    sg, ep = sigma[species[i], species[j]], eps[species[i], species[j]]
    c6 = (sg * sg / r2) ** 3
    u = 4 * ep * (c6 * c6 - c6)
    du = 12 * ep * c6 * (1 - 2 * c6) / r2
    S, dS = lj_switch_factor(r2, cutoff[s], rtr[s])
    us = np.where(keep, u * S, 0.0)
    dus = np.where(keep, du * S + u * dS, 0.0)
    e = np.zeros(nsys)
    np.add.at(e, s, 0.5 * us)
    f = np.zeros_like(pos)
    np.add.at(f, i, 2 * dus[:, None] * d)
    w = np.zeros((nsys, 3, 3))
    np.add.at(w, s, -dus[:, None, None] * d[:, :, None] * d[:, None, :])
    return e, f, w


def lj_cells(pos, species, sys_off, tril, cutoff, sigma, eps):
    """LJ energy/forces/virial by cell list, same pair set (particle exclusion)."""
    pos, species, sys_off, tril, cutoff = _d(pos), _i(species), _i(sys_off), _d(tril), _d(cutoff)
    sigma, eps = _d(sigma), _d(eps)
    nsys = len(sys_off) - 1
    e = np.zeros(nsys)
    f = np.zeros_like(pos)
    w = np.zeros((nsys, 9))
    lib().orc_lj_cells(_p(pos), _p(species), nsys, _p(sys_off), _p(tril), _p(cutoff),
                       sigma.shape[0], _p(sigma), _p(eps), _p(e), _p(f), _p(w))
    return e, f, w.reshape(nsys, 3, 3)


def morse_edges(pos, species, sys_off, tril, D, alpha, r0, ij, shift):
    """Morse bond energy/forces/virial over explicit edges (morse.py:142-168)."""
    pos, species, sys_off, tril = _d(pos), _i(species), _i(sys_off), _d(tril)
    D, alpha, r0, ij, shift = _d(D), _d(alpha), _d(r0), _i(ij), _i(shift)
    nsys = len(sys_off) - 1
    e = np.zeros(nsys)
    f = np.zeros_like(pos)
    w = np.zeros((nsys, 9))
    lib().orc_morse_edges(_p(pos), _p(species), nsys, _p(sys_off), _p(tril), D.shape[0],
                          _p(D), _p(alpha), _p(r0), C.c_long(len(ij)), _p(ij), _p(shift),
                          _p(e), _p(f), _p(w))
    return e, f, w.reshape(nsys, 3, 3)


def ewald_real_edges(pos, q, sys_off, tril, alpha, cutoff, ij, shift, exclusion_correction=False):
    """Real-space Ewald (ewald.py:305-325) or the exclusion correction (328-346)."""
    pos, q, sys_off, tril = _d(pos), _d(q), _i(sys_off), _d(tril)
    alpha, cutoff, ij, shift = _d(alpha), _d(cutoff), _i(ij), _i(shift)
    nsys = len(sys_off) - 1
    e = np.zeros(nsys)
    f = np.zeros_like(pos)
    w = np.zeros((nsys, 9))
    lib().orc_ewald_real_edges(_p(pos), _p(q), nsys, _p(sys_off), _p(tril), _p(alpha),
                               _p(cutoff), C.c_long(len(ij)), _p(ij), _p(shift),
                               int(bool(exclusion_correction)), _p(e), _p(f), _p(w))
    return e, f, w.reshape(nsys, 3, 3)


def ewald_recip(pos, q, sys_off, tril, alpha, kvec_n):
    """Reciprocal-space Ewald (ewald.py:349-439, 585-601); kvec_n (nsys, nk, 3)."""
    pos, q, sys_off, tril, alpha = _d(pos), _d(q), _i(sys_off), _d(tril), _d(alpha)
    kvec_n = _i(kvec_n)
    nsys = len(sys_off) - 1
    nk = kvec_n.shape[1]
    e = np.zeros(nsys)
    f = np.zeros_like(pos)
    w = np.zeros((nsys, 9))
    sk = np.zeros((nsys, nk, 2))
    lib().orc_ewald_recip(_p(pos), _p(q), nsys, _p(sys_off), _p(tril), _p(alpha), nk,
                          _p(kvec_n), _p(e), _p(f), _p(w), _p(sk))
    return e, f, w.reshape(nsys, 3, 3), sk


def ewald_self(q, sys_off, alpha):
    """Ewald self term -K alpha/sqrt(pi) sum q^2 (ewald.py:277-302)."""
    q, sys_off, alpha = _d(q), _i(sys_off), _d(alpha)
    nsys = len(sys_off) - 1
    e = np.zeros(nsys)
    lib().orc_ewald_self(_p(q), nsys, _p(sys_off), _p(alpha), _p(e))
    return e


def kvecs(tril, nmax, kmax=-1.0):
    """Integer half-space k grid (ewald.py:1322-1339 recipe); kmax < 0 keeps the box."""
    tril = _d(tril)
    L = lib()
    n = L.orc_kvecs(_p(tril), int(nmax), C.c_double(kmax), C.c_long(0), _p(np.zeros(3, np.int32)))
    out = np.zeros((max(n, 1), 3), np.int32)
    L.orc_kvecs(_p(tril), int(nmax), C.c_double(kmax), C.c_long(n), _p(out))
    return out[:n]


def md_step(integrator, pos, mom, mass, species, sys_off, tril, cutoff, sigma, eps, dt, kT,
            gamma, eta, forces):
    """One FP64 VV (0) / BAOAB (1) step in place; returns (energy, virial)."""
    for a in (pos, mom, forces):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    mass, species, sys_off, tril, cutoff = _d(mass), _i(species), _i(sys_off), _d(tril), _d(cutoff)
    sigma, eps, dt, kT, gamma, eta = _d(sigma), _d(eps), _d(dt), _d(kT), _d(gamma), _d(eta)
    nsys = len(sys_off) - 1
    e = np.zeros(nsys)
    w = np.zeros((nsys, 9))
    lib().orc_md_step(int(integrator), _p(pos), _p(mom), _p(mass), _p(species), nsys,
                      _p(sys_off), _p(tril), _p(cutoff), sigma.shape[0], _p(sigma), _p(eps),
                      _p(dt), _p(kT), _p(gamma), _p(eta), _p(forces), _p(e), _p(w))
    return e, w.reshape(nsys, 3, 3)


def gamma_mt(shape: float, seed: int, sys: int, t: int) -> float:
    """Gamma(shape, 1) draw of the CSVR chi-square (Marsaglia-Tsang, shared stream)."""
    return float(lib().orc_gamma_mt(C.c_double(shape), C.c_uint64(int(seed) & (2**64 - 1)),
                                    C.c_uint32(sys), C.c_int64(t)))


def csvr_scale(seed, t, mom, mass, sys_off, kT, dt, tau):
    """CSVR rescaling of FP64 momenta in place at global step t; returns alpha per system."""
    assert mom.dtype == np.float64 and mom.flags.c_contiguous
    mass, sys_off, kT, dt, tau = _d(mass), _i(sys_off), _d(kT), _d(dt), _d(tau)
    nsys = len(sys_off) - 1
    alpha = np.zeros(nsys)
    lib().orc_csvr_scale(C.c_uint64(int(seed) & (2**64 - 1)), C.c_int64(t), nsys, _p(sys_off),
                         _p(mass), _p(mom), _p(kT), _p(dt), _p(tau), _p(alpha))
    return alpha


def scr_scale(seed, t, mom, pos, tril, mass, sys_off, kT, dt, baro, virial):
    """Stochastic cell rescaling of FP64 positions and cells in place at
    global step t (virial W per system, (n_sys, 3, 3)); returns mu per system."""
    for a in (mom, pos, tril):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    mass, sys_off, kT, dt, baro = _d(mass), _i(sys_off), _d(kT), _d(dt), _d(baro)
    virial = _d(np.asarray(virial).reshape(-1, 9))
    nsys = len(sys_off) - 1
    mu = np.zeros(nsys)
    lib().orc_scr_scale(C.c_uint64(int(seed) & (2**64 - 1)), C.c_int64(t), nsys, _p(sys_off),
                        _p(mass), _p(mom), _p(pos), _p(tril), _p(kT), _p(dt), _p(baro),
                        _p(virial), _p(mu))
    return mu


def md_vv_f32(n_steps, pos, mom, mass, species, sys_off, tril, cutoff, sigma, eps, dt, forces):
    """n_steps FP32 velocity-Verlet LJ steps in place (the FP32 CPU baseline)."""
    for a in (pos, mom, forces):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    mass, species, sys_off, tril, cutoff = _f(mass), _i(species), _i(sys_off), _d(tril), _d(cutoff)
    sigma, eps, dt = _f(sigma), _f(eps), _d(dt)
    nsys = len(sys_off) - 1
    e = np.zeros(nsys)
    rc = lib().orc_md_vv_f32(int(n_steps), _p(pos), _p(mom), _p(mass), _p(species), nsys,
                             _p(sys_off), _p(tril), _p(cutoff), sigma.shape[0], _p(sigma),
                             _p(eps), _p(dt), _p(forces), _p(e))
    if rc != 0:
        raise ValueError("orc_md_vv_f32 supports only systems that need no periodic images")
    return e


class _GcArgs(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("n_chain", "n_steps", "gmax", "m", "nspec", "mx", "my", "mz")] + [
        ("n_motif", C.c_int), ("motif_nsite", C.c_void_p),
        ("step0", C.c_int64), ("seed", C.c_uint64),
        ("host_off", C.c_void_p), ("host_pos", C.c_void_p), ("host_q", C.c_void_p),
        ("host_spec", C.c_void_p), ("tril", C.c_void_p), ("motif_pos", C.c_void_p),
        ("motif_q", C.c_void_p), ("motif_spec", C.c_void_p), ("sig", C.c_void_p),
        ("eps", C.c_void_p), ("rc_lj", C.c_double), ("rc_cl", C.c_double),
        ("alpha", C.c_void_p), ("kmax", C.c_void_p), ("tail", C.c_void_p),
        ("mol_const", C.c_void_p), ("kT", C.c_void_p), ("log_activity", C.c_void_p),
        ("w_move", C.c_double * 4), ("upper", C.c_double * 4),
        ("guest_pos", C.c_void_p), ("n_guest", C.c_void_p), ("guest_motif", C.c_void_p),
        ("counts", C.c_void_p),
        ("sk", C.c_void_p), ("energy", C.c_void_p), ("width", C.c_void_p), ("hist", C.c_void_p),
        ("hist_idx", C.c_void_p), ("stats", C.c_void_p), ("ins_acc", C.c_void_p),
        ("del_acc", C.c_void_p), ("n_trace", C.c_int), ("trace", C.c_void_p)]


def gcmc_run(state: dict, n_steps: int, step0: int, seed: int, n_trace: int = 0):
    """FP64 GCMC of every chain in place (see orc_gcmc_run).  ``state`` holds
    numpy arrays: host_off, host_pos, host_q, host_spec, tril, motif_pos, motif_q,
    motif_spec, motif_nsite (n_motif), sigma, epsilon, alpha, kmax, tail, mol_const
    (n, n_motif), kT, log_activity (n, n_motif), guest_pos (n, gmax, m, 3), n_guest,
    guest_motif (n, gmax), counts, sk (n, nk_grid, 2), energy, width (n, 4) f32,
    hist (n, 4, 4) u32, hist_idx (n, 4), stats (n, 8) i64, ins_acc, del_acc;
    scalars rc_lj, rc_cl, grid (mx, my, mz), w_move, upper.  Motif arrays are
    (n_motif, m[, 3]) with m the largest site count."""
    a = _GcArgs()
    n = len(state["n_guest"])
    gmax, m = state["guest_pos"].shape[1:3]
    a.n_chain, a.n_steps, a.gmax, a.m = n, int(n_steps), gmax, m
    nsite = np.ascontiguousarray(state["motif_nsite"], dtype=np.int32)
    a.n_motif, a.motif_nsite = len(nsite), _p(nsite)
    a.nspec = state["sigma"].shape[0]
    a.mx, a.my, a.mz = state["grid"]
    a.step0, a.seed = int(step0), int(seed) & (2**64 - 1)
    keep = {}
    for key, dt in (("host_off", np.int32), ("host_pos", np.float64), ("host_q", np.float64),
                    ("host_spec", np.int32), ("tril", np.float64), ("motif_pos", np.float64),
                    ("motif_q", np.float64), ("motif_spec", np.int32), ("sigma", np.float64),
                    ("epsilon", np.float64), ("alpha", np.float64), ("kmax", np.float64),
                    ("tail", np.float64), ("mol_const", np.float64), ("kT", np.float64),
                    ("log_activity", np.float64)):
        keep[key] = np.ascontiguousarray(state[key], dtype=dt)
    for key, dt in (("guest_pos", np.float64), ("n_guest", np.int32), ("guest_motif", np.int32),
                    ("counts", np.int32),
                    ("sk", np.float64), ("energy", np.float64), ("width", np.float32),
                    ("hist", np.uint32), ("hist_idx", np.int32), ("stats", np.int64),
                    ("ins_acc", np.int64), ("del_acc", np.int64)):
        arr = state[key]
        assert arr.dtype == dt and arr.flags.c_contiguous, key
        keep[key] = arr
    trace = np.zeros((max(n_trace, 1), max(n_steps, 1), 4))
    a.host_off, a.host_pos, a.host_q = _p(keep["host_off"]), _p(keep["host_pos"]), _p(keep["host_q"])
    a.host_spec, a.tril = _p(keep["host_spec"]), _p(keep["tril"])
    a.motif_pos, a.motif_q, a.motif_spec = _p(keep["motif_pos"]), _p(keep["motif_q"]), _p(keep["motif_spec"])
    a.sig, a.eps = _p(keep["sigma"]), _p(keep["epsilon"])
    a.rc_lj, a.rc_cl = float(state["rc_lj"]), float(state["rc_cl"])
    a.alpha, a.kmax, a.tail = _p(keep["alpha"]), _p(keep["kmax"]), _p(keep["tail"])
    a.mol_const, a.kT, a.log_activity = _p(keep["mol_const"]), _p(keep["kT"]), _p(keep["log_activity"])
    for i in range(4):
        a.w_move[i] = float(state["w_move"][i])
        a.upper[i] = float(state["upper"][i])
    a.guest_pos, a.n_guest, a.counts = _p(keep["guest_pos"]), _p(keep["n_guest"]), _p(keep["counts"])
    a.guest_motif = _p(keep["guest_motif"])
    a.sk, a.energy, a.width = _p(keep["sk"]), _p(keep["energy"]), _p(keep["width"])
    a.hist, a.hist_idx, a.stats = _p(keep["hist"]), _p(keep["hist_idx"]), _p(keep["stats"])
    a.ins_acc, a.del_acc = _p(keep["ins_acc"]), _p(keep["del_acc"])
    a.n_trace, a.trace = int(n_trace), _p(trace)
    if lib().orc_gcmc_run(C.byref(a)) != 0:
        raise ValueError("orc_gcmc_run: unsupported sizes (1..8 motifs of 1..8 sites)")
    return trace[:n_trace]


def gcmc_state(g, chains=None) -> dict:
    """The oracle's state dict for chains of a kups_amd RigidGCMC (all chains by
    default), copied from its device arrays: S(k) scattered back onto the
    oracle's full k grid, host arrays re-based to the selected chains."""
    d = g.d
    idx = np.arange(g.n_chain) if chains is None else np.asarray(chains)
    cpu = lambda k: d[k].cpu().numpy()
    nm = g.nm
    w = 2 * nm + 1
    skc = cpu("sk")[idx]
    sk = np.zeros((len(idx), (nm + 1) * w * w, 2))
    for r, s in enumerate(idx):
        sk[r, g._kgrid_index[s]] = skc[r]
    off = cpu("host_off")
    host = cpu("host").astype(np.float64)
    spec = cpu("host_species")
    rows = np.concatenate([np.arange(off[s], off[s + 1]) for s in idx]).astype(np.int64)
    sizes = off[idx + 1] - off[idx]
    return dict(
        host_off=np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32),
        host_pos=np.ascontiguousarray(host[rows, :3]), host_q=np.ascontiguousarray(host[rows, 3]),
        host_spec=spec[rows], tril=cpu("tril")[idx],
        motif_pos=cpu("motif_pos").astype(np.float64), motif_q=cpu("motif_q").astype(np.float64),
        motif_spec=cpu("motif_species"), motif_nsite=np.asarray(g.motif_nsite, np.int32),
        sigma=g.lj.sigma, epsilon=g.lj.epsilon,
        alpha=g.alpha[idx], kmax=g.kmax[idx], tail=g.tail[idx],
        mol_const=np.ascontiguousarray(g.mol_const[idx]),
        kT=cpu("kT")[idx], log_activity=np.ascontiguousarray(g.log_activity[idx]),
        guest_pos=np.ascontiguousarray(cpu("guest")[idx].astype(np.float64)),
        n_guest=np.ascontiguousarray(cpu("n_guest")[idx]),
        guest_motif=np.ascontiguousarray(cpu("guest_motif")[idx]),
        counts=np.ascontiguousarray(cpu("counts")[idx]), sk=sk,
        energy=np.ascontiguousarray(cpu("energy")[idx]), width=np.ascontiguousarray(cpu("width")[idx]),
        hist=np.ascontiguousarray(cpu("hist")[idx].view(np.uint32)),
        hist_idx=np.ascontiguousarray(cpu("hist_idx")[idx]),
        stats=np.zeros((len(idx), 8), np.int64), ins_acc=np.zeros(len(idx), np.int64),
        del_acc=np.zeros(len(idx), np.int64), rc_lj=g.lj.cutoff, rc_cl=g.ewald_cutoff,
        grid=(nm, nm, nm), w_move=g.weights, upper=g.upper)


def philox_raw(ctr, key) -> np.ndarray:
    """Philox4x32-10 of counter ctr[4] under key key[2] (the bare generator)."""
    c, k = np.ascontiguousarray(ctr, np.uint32), np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox_raw(_p(c), _p(k), _p(out))
    return out


def philox(seed: int, a: int, t: int, tag: int) -> np.ndarray:
    """The four Philox4x32-10 words of counter (a, t, t >> 32, tag)."""
    out = np.zeros(4, np.uint32)
    lib().orc_philox(C.c_uint64(int(seed) & (2**64 - 1)), C.c_uint32(a), C.c_int64(t),
                     C.c_uint32(tag), _p(out))
    return out


# ---------------------------------------------------------------- MLP --
# FP64 restatement of the proposal network's forward pass and gradient
# (textbook backpropagation; the reference has no such network).  Layout of the
# flat parameter vector: W1[32][128] b1 W2[128][128] b2 W3[128][128] b3 w4[128] b4.

def _unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    out = []
    for shape in ((32, 128), (128,), (128, 128), (128,), (128, 128), (128,), (128,), (1,)):
        n = int(np.prod(shape))
        out.append(p[o:o + n].reshape(shape))
        o += n
    assert o == p.size
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def mlp_forward(params, X):
    W1, b1, W2, b2, W3, b3, w4, b4 = _unpack(params)
    h1 = _silu(X @ W1 + b1)
    h2 = _silu(h1 @ W2 + b2)
    h3 = _silu(h2 @ W3 + b3)
    return h3 @ w4 + b4[0]


def mlp_loss_and_grad(params, X, y):
    """loss = 1/2 mean (f(X) - y)^2 and its gradient w.r.t. the flat parameters."""
    W1, b1, W2, b2, W3, b3, w4, b4 = _unpack(params)
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    n = len(y)
    z1 = X @ W1 + b1; h1 = _silu(z1)
    z2 = h1 @ W2 + b2; h2 = _silu(z2)
    z3 = h2 @ W3 + b3; h3 = _silu(z3)
    out = h3 @ w4 + b4[0]
    diff = out - y
    loss = 0.5 * np.mean(diff**2)
    dout = diff / n
    gw4 = h3.T @ dout
    gb4 = np.array([dout.sum()])
    d3 = np.outer(dout, w4) * _dsilu(z3)
    gW3 = h2.T @ d3; gb3 = d3.sum(0)
    d2 = (d3 @ W3.T) * _dsilu(z2)
    gW2 = h1.T @ d2; gb2 = d2.sum(0)
    d1 = (d2 @ W2.T) * _dsilu(z1)
    gW1 = X.T @ d1; gb1 = d1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gw4, gb4)])
    return loss, g
