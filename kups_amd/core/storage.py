"""HDF5 trajectory storage (core/storage.py:101-381) on the HDF5 C library.

The file layout is the reference's ``HDF5StorageWriter``'s:

* file attributes ``config_class_name`` (str), ``group_names`` (JSON list of
  the group names) and, after the run, ``actual_steps`` (int64);
* one HDF5 group per leaf of the logging config, named ``"group"`` + the
  leaf's key path (``group.init``, ``group.step``, ...), holding one dataset
  per array leaf of what the leaf's view returns, named ``"array"`` + that
  leaf's key path (``array.potential_energy``, ``array.atoms.positions``),
  of shape leading + leaf shape, where the leading shape is ``()`` for
  ``Once`` and ``(ceil(total_steps / n),)`` for ``EveryNStep(n)``; group
  attributes ``data_class_name`` and ``paths`` (JSON list of the dataset
  names, in leaf order);
* key paths follow JAX's key-path strings: ``.name`` for a dataclass field,
  ``['key']`` for a mapping key, ``[i]`` for a sequence index.

Dtypes are the little-endian standard HDF5 types of the numpy dtype (h5py's
mapping; bool as h5py's FALSE / TRUE enum over int8), strings are
variable-length UTF-8.  The reference additionally stores pickled JAX tree
definitions (the ``config_pytree`` and ``tree_def`` datasets); those are
JAX-internal objects, so here the structures are stored as JSON instead (file
and group attribute ``tree_def_json``), which ``HDF5StorageReader`` uses to
rebuild nested dicts.  Writes are synchronous (the reference queues them to a
background thread; the data written is the same).

The library is located through ``KUPS_AMD_HDF5_LIB``, then the dynamic
loader's search path, then ``/opt/conda/lib``; without it the writer and
reader raise ``HDF5Unavailable``.
"""

from __future__ import annotations

import ctypes as C
import ctypes.util
import dataclasses
import glob
import json
import os
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Callable

import numpy as np

__all__ = ["HDF5Unavailable", "LoggingFrequency", "Once", "EveryNStep", "WriterGroupConfig",
           "HDF5StorageWriter", "HDF5StorageReader", "GroupReader", "flatten_with_paths"]


class HDF5Unavailable(RuntimeError):
    """The HDF5 C library could not be loaded."""


# ------------------------------------------------------------ C library
_hid = C.c_int64
_herr = C.c_int
_hsize = C.c_ulonglong
_P_DEFAULT = 0
_S_ALL = 0
_F_ACC_RDONLY, _F_ACC_TRUNC = 0x0000, 0x0002
_S_SCALAR = 0
_S_SELECT_SET = 0
_T_VARIABLE = C.c_size_t(-1).value
_T_CSET_UTF8 = 1
_T_INTEGER, _T_FLOAT, _T_STRING, _T_ENUM = 0, 1, 3, 8
_T_SGN_NONE = 0


class _H5:
    def __init__(self, path: str):
        self.lib = L = C.CDLL(path)
        sig = {
            "H5open": (_herr, []),
            "H5Fcreate": (_hid, [C.c_char_p, C.c_uint, _hid, _hid]),
            "H5Fopen": (_hid, [C.c_char_p, C.c_uint, _hid]),
            "H5Fclose": (_herr, [_hid]),
            "H5Gcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid]),
            "H5Gopen2": (_hid, [_hid, C.c_char_p, _hid]),
            "H5Gclose": (_herr, [_hid]),
            "H5Screate": (_hid, [C.c_int]),
            "H5Screate_simple": (_hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Sselect_hyperslab": (_herr, [_hid, C.c_int, C.POINTER(_hsize), C.POINTER(_hsize),
                                            C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Sget_simple_extent_ndims": (C.c_int, [_hid]),
            "H5Sget_simple_extent_dims": (C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
            "H5Sclose": (_herr, [_hid]),
            "H5Dcreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
            "H5Dopen2": (_hid, [_hid, C.c_char_p, _hid]),
            "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
            "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
            "H5Dget_space": (_hid, [_hid]),
            "H5Dget_type": (_hid, [_hid]),
            "H5Dclose": (_herr, [_hid]),
            "H5Acreate2": (_hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid]),
            "H5Aopen": (_hid, [_hid, C.c_char_p, _hid]),
            "H5Aexists": (C.c_int, [_hid, C.c_char_p]),
            "H5Awrite": (_herr, [_hid, _hid, C.c_void_p]),
            "H5Aread": (_herr, [_hid, _hid, C.c_void_p]),
            "H5Aget_type": (_hid, [_hid]),
            "H5Aclose": (_herr, [_hid]),
            "H5Lexists": (C.c_int, [_hid, C.c_char_p, _hid]),
            "H5Tcopy": (_hid, [_hid]),
            "H5Tset_size": (_herr, [_hid, C.c_size_t]),
            "H5Tset_cset": (_herr, [_hid, C.c_int]),
            "H5Tget_class": (C.c_int, [_hid]),
            "H5Tget_size": (C.c_size_t, [_hid]),
            "H5Tget_sign": (C.c_int, [_hid]),
            "H5Tis_variable_str": (C.c_int, [_hid]),
            "H5Tenum_create": (_hid, [_hid]),
            "H5Tenum_insert": (_herr, [_hid, C.c_char_p, C.c_void_p]),
            "H5Tclose": (_herr, [_hid]),
            "H5free_memory": (_herr, [C.c_void_p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        if L.H5open() < 0:
            raise HDF5Unavailable(f"H5open failed in {path}")
        g = lambda n: _hid.in_dll(L, n).value  # noqa: E731
        # keyed by kind + size ("f8", "i4", ...); stored little-endian
        self.types = {
            "f8": g("H5T_IEEE_F64LE_g"), "f4": g("H5T_IEEE_F32LE_g"),
            "i8": g("H5T_STD_I64LE_g"), "i4": g("H5T_STD_I32LE_g"),
            "i2": g("H5T_STD_I16LE_g"), "i1": g("H5T_STD_I8LE_g"),
            "u8": g("H5T_STD_U64LE_g"), "u4": g("H5T_STD_U32LE_g"),
            "u2": g("H5T_STD_U16LE_g"), "u1": g("H5T_STD_U8LE_g"),
        }
        self.c_s1 = g("H5T_C_S1_g")
        # h5py's boolean: enum FALSE = 0, TRUE = 1 over int8
        self.bool_t = L.H5Tenum_create(g("H5T_NATIVE_INT8_g"))
        for name, v in ((b"FALSE", 0), (b"TRUE", 1)):
            val = C.c_int8(v)
            self._ok(L.H5Tenum_insert(self.bool_t, name, C.byref(val)), "H5Tenum_insert")
        self.str_t = L.H5Tcopy(self.c_s1)
        self._ok(L.H5Tset_size(self.str_t, _T_VARIABLE), "H5Tset_size")
        self._ok(L.H5Tset_cset(self.str_t, _T_CSET_UTF8), "H5Tset_cset")

    @staticmethod
    def _ok(rc, what):
        if rc < 0:
            raise OSError(f"HDF5 call {what} failed ({rc})")
        return rc

    def type_of(self, dt: np.dtype) -> int:
        """Storage (and memory) type of a little-endian numpy dtype."""
        dt = np.dtype(dt)
        if dt == np.bool_:
            return self.bool_t
        if dt.byteorder == ">":
            raise TypeError(f"big-endian dtype {dt}: convert to little-endian first")
        t = self.types.get(f"{dt.kind}{dt.itemsize}")
        if t is None:
            raise TypeError(f"no HDF5 storage type for dtype {dt}")
        return t

    def dtype_of(self, t: int) -> np.dtype:
        L = self.lib
        cls, size = L.H5Tget_class(t), L.H5Tget_size(t)
        if cls == _T_FLOAT:
            return np.dtype(f"<f{size}")
        if cls == _T_INTEGER:
            return np.dtype(f"<{'u' if L.H5Tget_sign(t) == _T_SGN_NONE else 'i'}{size}")
        if cls == _T_ENUM and size == 1:
            return np.dtype(np.bool_)
        raise TypeError(f"HDF5 type class {cls} is not read by this reader")


_LIB: _H5 | None = None


def _h5() -> _H5:
    global _LIB
    if _LIB is not None:
        return _LIB
    cands = []
    if os.environ.get("KUPS_AMD_HDF5_LIB"):
        cands.append(os.environ["KUPS_AMD_HDF5_LIB"])
    found = ctypes.util.find_library("hdf5")
    if found:
        cands.append(found)
    cands += sorted(glob.glob("/opt/conda/lib/libhdf5.so*"), key=len)
    errs = []
    for c in cands:
        try:
            _LIB = _H5(c)
            return _LIB
        except (OSError, AttributeError, HDF5Unavailable) as e:
            errs.append(f"{c}: {e}")
    raise HDF5Unavailable("no usable HDF5 C library (set KUPS_AMD_HDF5_LIB): " +
                          ("; ".join(errs) or "none found"))


def _dims(shape) -> C.Array:
    return (_hsize * max(len(shape), 1))(*[int(d) for d in shape])


def _write_attr(loc: int, name: str, value) -> None:
    h = _h5()
    L = h.lib
    space = L.H5Screate(_S_SCALAR)
    try:
        if isinstance(value, str):
            t = h.str_t
            buf = (C.c_char_p * 1)(value.encode())
        else:
            arr = np.asarray(value, order="C")
            t = h.type_of(arr.dtype)
            buf = arr
        if L.H5Aexists(loc, name.encode()) > 0:
            raise ValueError(f"attribute {name} exists")
        a = h._ok(L.H5Acreate2(loc, name.encode(), t, space, _P_DEFAULT, _P_DEFAULT), "H5Acreate2")
        try:
            ptr = C.cast(buf, C.c_void_p) if isinstance(value, str) else buf.ctypes.data_as(C.c_void_p)
            h._ok(L.H5Awrite(a, t, ptr), "H5Awrite")
        finally:
            L.H5Aclose(a)
    finally:
        L.H5Sclose(space)


def _read_attr(loc: int, name: str):
    h = _h5()
    L = h.lib
    a = h._ok(L.H5Aopen(loc, name.encode(), _P_DEFAULT), f"H5Aopen({name})")
    try:
        t = L.H5Aget_type(a)
        try:
            if L.H5Tget_class(t) == _T_STRING:
                if L.H5Tis_variable_str(t) > 0:
                    buf = (C.c_void_p * 1)()
                    h._ok(L.H5Aread(a, t, buf), "H5Aread")
                    s = C.string_at(buf[0]).decode()
                    L.H5free_memory(buf[0])
                    return s
                sbuf = C.create_string_buffer(L.H5Tget_size(t))
                h._ok(L.H5Aread(a, t, sbuf), "H5Aread")
                return sbuf.value.decode()
            out = np.zeros((), h.dtype_of(t))
            h._ok(L.H5Aread(a, t, out.ctypes.data_as(C.c_void_p)), "H5Aread")
            return out[()]
        finally:
            L.H5Tclose(t)
    finally:
        L.H5Aclose(a)


# ------------------------------------------------------------ pytrees
def _as_numpy(x) -> np.ndarray:
    if hasattr(x, "detach"):  # torch tensor
        x = x.detach().cpu().numpy()
    return np.asarray(x)


def flatten_with_paths(tree, prefix: str = ""):
    """(key path, array) leaves in order, plus a JSON-able structure:
    dataclass fields ``.name``, mapping keys ``['key']``, sequence items
    ``[i]``; anything else is an array leaf."""
    if dataclasses.is_dataclass(tree) and not isinstance(tree, type):
        leaves, fields = [], {}
        for f in dataclasses.fields(tree):
            sub, st = flatten_with_paths(getattr(tree, f.name), f"{prefix}.{f.name}")
            leaves += sub
            fields[f.name] = st
        return leaves, {"dataclass": type(tree).__qualname__, "fields": fields}
    if isinstance(tree, dict):
        leaves, items = [], {}
        for k, v in tree.items():
            sub, st = flatten_with_paths(v, f"{prefix}['{k}']")
            leaves += sub
            items[str(k)] = st
        return leaves, {"dict": items}
    if isinstance(tree, (list, tuple)):
        leaves, items = [], []
        for i, v in enumerate(tree):
            sub, st = flatten_with_paths(v, f"{prefix}[{i}]")
            leaves += sub
            items.append(st)
        return leaves, {"seq": items}
    return [(prefix, _as_numpy(tree))], {"leaf": prefix}


def _unflatten(st, values: dict):
    if "leaf" in st:
        return values[st["leaf"]]
    if "dict" in st:
        return {k: _unflatten(v, values) for k, v in st["dict"].items()}
    if "seq" in st:
        return [_unflatten(v, values) for v in st["seq"]]
    return {k: _unflatten(v, values) for k, v in st["fields"].items()}


# ------------------------------------------------------------ logging frequency
class LoggingFrequency:
    """When a group is logged, its leading dataset shape and the dataset index
    of a step (core/storage.py:35-47)."""

    def should_log(self, step: int) -> bool: ...
    def leading_shape(self, total_steps: int) -> tuple[int, ...]: ...
    def dataset_index(self, step: int): ...


class Once(LoggingFrequency):
    """Step 0 only; datasets without a time axis (core/storage.py:50-60)."""

    def should_log(self, step: int) -> bool:
        return step == 0

    def leading_shape(self, total_steps: int) -> tuple[int, ...]:
        return ()

    def dataset_index(self, step: int):
        return None


@dataclass
class EveryNStep(LoggingFrequency):
    """Steps 0, n, 2n, ...; time axis of ceil(total / n) (core/storage.py:63-82)."""

    n: int

    def should_log(self, step: int) -> bool:
        return step % self.n == 0

    def leading_shape(self, total_steps: int) -> tuple[int, ...]:
        return ((total_steps + self.n - 1) // self.n,)

    def dataset_index(self, step: int):
        return step // self.n


@dataclass(frozen=True)
class WriterGroupConfig:
    """A view of the state and how often it is logged (core/storage.py:85-96)."""

    view: Callable[[Any], Any]
    logging_frequency: LoggingFrequency


# ------------------------------------------------------------ writer
@dataclass
class _GroupWriter:
    cfg: WriterGroupConfig
    gid: int
    dsets: list  # (dataset id, leaf shape, dtype)


@dataclass
class HDF5StorageWriter:
    """Writes the groups of ``config`` (a dataclass or dict whose leaves are
    ``WriterGroupConfig``) to ``out_path`` (core/storage.py:101-174).  Use as
    a context manager; ``log(state, step)`` writes the groups due at
    ``step``."""

    out_path: str | Path
    config: Any
    initial_state: Any
    total_steps: int
    _fid: int = field(init=False, default=-1, repr=False)
    _groups: list = field(init=False, default_factory=list, repr=False)
    _actual_steps: int = field(init=False, default=0, repr=False)

    def __enter__(self) -> "HDF5StorageWriter":
        h = _h5()
        L = h.lib
        self._fid = h._ok(L.H5Fcreate(str(self.out_path).encode(), _F_ACC_TRUNC, _P_DEFAULT,
                                      _P_DEFAULT), f"H5Fcreate({self.out_path})")
        confs = _config_leaves(self.config)
        _write_attr(self._fid, "config_class_name", type(self.config).__qualname__)
        names = []
        for path, cfg in confs:
            gname = "group" + path
            gid = h._ok(L.H5Gcreate2(self._fid, gname.encode(), _P_DEFAULT, _P_DEFAULT,
                                     _P_DEFAULT), f"H5Gcreate2({gname})")
            state = cfg.view(self.initial_state)
            leaves, st = flatten_with_paths(state)
            lead = cfg.logging_frequency.leading_shape(self.total_steps)
            dsets, paths = [], []
            for lp, arr in leaves:
                dname = "array" + lp
                shape = tuple(lead) + arr.shape
                t = h.type_of(arr.dtype)
                space = L.H5Screate_simple(len(shape), _dims(shape), None) if shape else \
                    L.H5Screate(_S_SCALAR)
                try:
                    d = h._ok(L.H5Dcreate2(gid, dname.encode(), t, space, _P_DEFAULT, _P_DEFAULT,
                                           _P_DEFAULT), f"H5Dcreate2({dname})")
                finally:
                    L.H5Sclose(space)
                dsets.append((d, arr.shape, arr.dtype))
                paths.append(dname)
            _write_attr(gid, "data_class_name", type(state).__qualname__)
            _write_attr(gid, "paths", json.dumps(paths))
            _write_attr(gid, "tree_def_json", json.dumps(st))
            self._groups.append(_GroupWriter(cfg, gid, dsets))
            names.append(gname)
        _write_attr(self._fid, "group_names", json.dumps(names))
        _write_attr(self._fid, "tree_def_json",
                    json.dumps(_config_structure(self.config, names)))
        return self

    def log(self, state, step: int) -> None:
        if self._fid < 0:
            raise RuntimeError("HDF5StorageWriter must be used inside a with-block")
        self._actual_steps = step + 1
        for g in self._groups:
            freq = g.cfg.logging_frequency
            if not freq.should_log(step):
                continue
            leaves, _ = flatten_with_paths(g.cfg.view(state))
            idx = freq.dataset_index(step)
            for (d, shape, dt), (_, arr) in zip(g.dsets, leaves):
                _write_rows(d, np.asarray(arr, dtype=dt, order="C"), shape, idx)

    def __exit__(self, *exc) -> None:
        if self._fid < 0:
            return
        h = _h5()
        L = h.lib
        try:
            _write_attr(self._fid, "actual_steps", np.int64(self._actual_steps))
        finally:
            for g in self._groups:
                for d, _, _ in g.dsets:
                    L.H5Dclose(d)
                L.H5Gclose(g.gid)
            self._groups = []
            L.H5Fclose(self._fid)
            self._fid = -1


def _config_leaves(config, prefix: str = ""):
    if isinstance(config, WriterGroupConfig):
        return [(prefix, config)]
    out = []
    if dataclasses.is_dataclass(config) and not isinstance(config, type):
        for f in dataclasses.fields(config):
            out += _config_leaves(getattr(config, f.name), f"{prefix}.{f.name}")
    elif isinstance(config, dict):
        for k, v in config.items():
            out += _config_leaves(v, f"{prefix}['{k}']")
    else:
        raise TypeError(f"logging config leaf at '{prefix}' is not a WriterGroupConfig")
    return out


def _config_structure(config, names):
    it = iter(names)

    def rec(c):
        if isinstance(c, WriterGroupConfig):
            return next(it)
        if dataclasses.is_dataclass(c) and not isinstance(c, type):
            return {f.name: rec(getattr(c, f.name)) for f in dataclasses.fields(c)}
        return {k: rec(v) for k, v in c.items()}

    return rec(config)


def _write_rows(d: int, arr: np.ndarray, shape, idx) -> None:
    h = _h5()
    L = h.lib
    t = h.type_of(arr.dtype)
    if arr.shape != tuple(shape):
        raise ValueError(f"leaf shape changed from {tuple(shape)} to {arr.shape}")
    if idx is None:
        h._ok(L.H5Dwrite(d, t, _S_ALL, _S_ALL, _P_DEFAULT, arr.ctypes.data_as(C.c_void_p)),
              "H5Dwrite")
        return
    fspace = L.H5Dget_space(d)
    rank = len(shape) + 1
    mshape = (1,) + tuple(shape)
    mspace = L.H5Screate_simple(rank, _dims(mshape), None)
    try:
        start = _dims((int(idx),) + (0,) * len(shape))
        h._ok(L.H5Sselect_hyperslab(fspace, _S_SELECT_SET, start, None, _dims(mshape), None),
              "H5Sselect_hyperslab")
        h._ok(L.H5Dwrite(d, t, mspace, fspace, _P_DEFAULT, arr.ctypes.data_as(C.c_void_p)),
              "H5Dwrite")
    finally:
        L.H5Sclose(mspace)
        L.H5Sclose(fspace)


# ------------------------------------------------------------ reader
@dataclass
class GroupReader:
    """One logging group: ``read(index)`` / ``[index]`` returns the group's
    structure (nested dicts) with every dataset indexed along its leading
    axis (core/storage.py:320-349)."""

    fid: int
    name: str

    def _gid(self) -> int:
        h = _h5()
        return h._ok(h.lib.H5Gopen2(self.fid, self.name.encode(), _P_DEFAULT),
                     f"H5Gopen2({self.name})")

    @property
    def paths(self) -> list[str]:
        gid = self._gid()
        try:
            return json.loads(_read_attr(gid, "paths"))
        finally:
            _h5().lib.H5Gclose(gid)

    def dataset(self, name: str) -> np.ndarray:
        h = _h5()
        L = h.lib
        gid = self._gid()
        try:
            d = h._ok(L.H5Dopen2(gid, name.encode(), _P_DEFAULT), f"H5Dopen2({name})")
            try:
                space, t = L.H5Dget_space(d), L.H5Dget_type(d)
                try:
                    rank = L.H5Sget_simple_extent_ndims(space)
                    dims = (_hsize * max(rank, 1))()
                    L.H5Sget_simple_extent_dims(space, dims, None)
                    out = np.empty(tuple(int(dims[i]) for i in range(rank)), h.dtype_of(t))
                    mem_t = h.type_of(out.dtype)
                    h._ok(L.H5Dread(d, mem_t, _S_ALL, _S_ALL, _P_DEFAULT,
                                    out.ctypes.data_as(C.c_void_p)), "H5Dread")
                    return out
                finally:
                    L.H5Sclose(space)
                    L.H5Tclose(t)
            finally:
                L.H5Dclose(d)
        finally:
            L.H5Gclose(gid)

    def read(self, index=None):
        gid = self._gid()
        try:
            st = json.loads(_read_attr(gid, "tree_def_json"))
        finally:
            _h5().lib.H5Gclose(gid)
        vals = {}
        for name in self.paths:
            arr = self.dataset(name)
            vals[name[len("array"):]] = arr if index is None else arr[index]
        return _unflatten(st, vals)

    def __getitem__(self, index):
        return self.read(index)


@dataclass
class HDF5StorageReader:
    """Reads files written by ``HDF5StorageWriter`` (core/storage.py:245-317)."""

    path: str | Path
    _fid: int = field(init=False, default=-1, repr=False)

    def __enter__(self) -> "HDF5StorageReader":
        h = _h5()
        self._fid = h._ok(h.lib.H5Fopen(str(self.path).encode(), _F_ACC_RDONLY, _P_DEFAULT),
                          f"H5Fopen({self.path})")
        return self

    def __exit__(self, *exc) -> None:
        if self._fid >= 0:
            _h5().lib.H5Fclose(self._fid)
            self._fid = -1

    def attr(self, name: str):
        return _read_attr(self._fid, name)

    def list_groups(self) -> list[str]:
        return json.loads(self.attr("group_names"))

    def focus_group(self, name_or_key: str) -> GroupReader:
        """A group by its HDF5 name (``group.step``) or by the config key
        (``step``)."""
        names = self.list_groups()
        if name_or_key not in names:
            cand = "group." + name_or_key
            if cand not in names:
                raise KeyError(f"no logging group {name_or_key!r} (groups: {names})")
            name_or_key = cand
        return GroupReader(self._fid, name_or_key)
