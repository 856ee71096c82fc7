"""ctypes binding of libkupsmi.so (the HIP/gfx950 library).

The library is the only compute path: if it is missing, or no GPU is
visible, every operation raises.  There is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libkupsmi.so"

_lib = None
_ctx: dict[int, C.c_void_p] = {}

vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
u64 = C.c_uint64

_SIGNATURES = {
    "kupsmi_version": ([], C.c_int),
    "kupsmi_last_error": ([], C.c_char_p),
    "kupsmi_ctx_create": ([C.POINTER(vp)], C.c_int),
    "kupsmi_ctx_destroy": ([vp], C.c_int),
    "kupsmi_ctx_freeze": ([vp, i32], C.c_int),
    "kupsmi_verlet_freeze": ([vp, i32], C.c_int),
    "kupsmi_verlet_profile": ([vp, i32, vp, vp], C.c_int),
    "kupsmi_verlet_gate": ([vp, vp], C.c_int),
    "kupsmi_ewald_phase_event": ([vp, C.POINTER(vp)], C.c_int),
    "kupsmi_neighborlist": ([vp, vp, i32, vp, i32, vp, vp, vp, vp, i64, vp, vp, vp], C.c_int),
    "kupsmi_neighborlist_cells": (
        [vp, vp, i32, vp, i32, i32, vp, vp, vp, vp, i64, vp, vp, vp], C.c_int),
    "kupsmi_neighborlist_query": (
        [vp, vp, i32, vp, i32, vp, vp, i32, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp], C.c_int),
    "kupsmi_ctx_record_pairs": ([vp, i64, vp, vp, vp, vp], C.c_int),
    "kupsmi_verlet_record_pairs": ([vp, i64, vp, vp, vp, vp], C.c_int),
    "kupsmi_neighborlist_changes": (
        [vp, vp, i32, vp, i32, vp, vp, vp, vp, i32, vp, vp, i64, vp, vp, vp, i64, vp, vp, vp],
        C.c_int,
    ),
    "kupsmi_lj_forces": ([vp, vp, i32, vp, i32, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp], C.c_int),
    "kupsmi_md_lj": (
        [vp, vp, i32, i32, i64, u64, i32, vp, i32, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp,
         C.c_double, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp],
        C.c_int,
    ),
    "kupsmi_normals": ([vp, u64, i64, i32, vp], C.c_int),
    "kupsmi_philox": ([vp, i32, vp, vp, vp], C.c_int),
    "kupsmi_pair_forces": (
        [vp, vp, i32, vp, i32, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp,
         vp, vp, vp],
        C.c_int,
    ),
    "kupsmi_ewald_recip": (
        [vp, vp, i32, vp, i32, i32, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp],
        C.c_int,
    ),
    "kupsmi_ewald_recip_noalloc": (
        [vp, i32, vp, i32, i32, vp, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp, vp, vp],
        C.c_int,
    ),
    "kupsmi_gcmc_run": (
        [vp, i32, i32, i64, u64, i32, vp, vp, vp, vp, i32, i32, i32, vp, vp, vp, vp, i32, vp,
         vp, C.c_double, C.c_double, vp, vp, vp, vp, vp, i32, i32, vp, vp, vp, vp, vp, vp,
         vp, vp, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, i32, vp, i32, vp, vp, vp],
        C.c_int,
    ),
    "kupsmi_verlet_create": ([vp], C.c_int),
    "kupsmi_verlet_destroy": ([vp], C.c_int),
    "kupsmi_verlet_invalidate": ([vp, vp], C.c_int),
    "kupsmi_verlet_counters": ([vp, vp, vp], C.c_int),
    "kupsmi_pair_forces_verlet": (
        [vp, vp, i32, vp, i32, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, C.c_double,
         i32, i32, vp, vp, vp],
        C.c_int),
    "kupsmi_morse_bonds": (
        [vp, i32, vp, i32, vp, vp, i32, vp, vp, vp, C.c_longlong, vp, vp, vp, i32, vp, vp, vp],
        C.c_int),
    "kupsmi_mlp_param_count": ([], C.c_int),
    "kupsmi_voxel_descriptors": ([vp, i32, i32, i32, vp, vp, vp, C.c_float, C.c_float, vp], C.c_int),
    "kupsmi_mlp_infer": ([vp, i32, vp, vp, vp, vp], C.c_int),
    "kupsmi_mlp_train_step": (
        [vp, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, C.c_float, i32, i32, vp, vp], C.c_int),
    "kupsmi_mlp_round_weights": ([vp, vp, vp, vp], C.c_int),
    "kupsmi_proposal_table": ([vp, i32, i32, vp, C.c_float, C.c_float, vp, vp], C.c_int),
    "kupsmi_integrate": (
        [vp, i32, i32, i32, vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, u64, i64, vp, vp],
        C.c_int,
    ),
    "kupsmi_thermostat_barostat": (
        [vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, u64, i64, vp], C.c_int),
    "kupsmi_step_advance": ([vp, vp, i64], C.c_int),
}


class KupsmiError(RuntimeError):
    """An entry point of libkupsmi returned an error."""


def load():
    """Load libkupsmi.so and declare its entry points (no GPU needed)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise KupsmiError(
                f"{LIB_PATH} is missing: build it with `python -m kups_amd.build` "
                "(there is no CPU fallback)"
            )
        lib = C.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def check(rc: int) -> None:
    if rc != 0:
        msg = load().kupsmi_last_error().decode(errors="replace")
        raise KupsmiError(f"libkupsmi error {rc}: {msg}")


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise KupsmiError("no GPU visible: the kups_amd kernels run only on the MI355X")
    return torch.device("cuda", torch.cuda.current_device())


def ctx() -> C.c_void_p:
    """The library context of the current device (created on first use)."""
    dev = require_gpu().index
    if dev not in _ctx:
        h = vp()
        check(load().kupsmi_ctx_create(C.byref(h)))
        _ctx[dev] = h
    return _ctx[dev]


class Context:
    """A library context of its own (scratch that no other caller can grow or
    free), e.g. for work recorded in a HIP graph; ``freeze()`` then makes any
    call that would have to grow its scratch fail instead of reallocating."""

    def __init__(self):
        require_gpu()
        h = vp()
        check(load().kupsmi_ctx_create(C.byref(h)))
        self._h = h

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def freeze(self, frozen: bool = True) -> None:
        check(load().kupsmi_ctx_freeze(self._h, int(frozen)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                load().kupsmi_ctx_destroy(h)
            except Exception:
                pass
            self._h = None


def stream_ptr(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise KupsmiError("kups_amd kernels take device tensors")
    if not t.is_contiguous():
        raise KupsmiError("kups_amd kernels take contiguous tensors")
    return t.data_ptr()
