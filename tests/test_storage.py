"""HDF5 storage (core/storage.py of the reference) on the CPU.

The writer's files are read back by this package's reader and, independently,
by the HDF5 distribution's ``h5dump`` (values, storage types, group and
dataset names, attributes) when it is installed.
"""

from __future__ import annotations

import dataclasses
import json
import re
import shutil
import subprocess

import numpy as np
import pytest

from kups_amd.core import storage
from kups_amd.core.storage import (EveryNStep, HDF5StorageReader, HDF5StorageWriter, Once,
                                   WriterGroupConfig)

try:
    storage._h5()
except storage.HDF5Unavailable as e:  # pragma: no cover - the image ships libhdf5
    pytest.skip(f"HDF5 C library not available: {e}", allow_module_level=True)

H5DUMP = shutil.which("h5dump") or (
    "/opt/conda/bin/h5dump" if shutil.os.path.exists("/opt/conda/bin/h5dump") else None)


@dataclasses.dataclass
class Step:
    potential_energy: np.ndarray
    stress: np.ndarray
    flags: np.ndarray
    counts: np.ndarray


@dataclasses.dataclass
class Config:
    init: WriterGroupConfig
    step: WriterGroupConfig


def _state(k):
    return {"pos": (np.arange(12.0).reshape(4, 3) + k).astype(np.float32),
            "e": np.array([1.5 * k, -2.0 * k]),
            "flags": np.array([k % 3 == 0, k % 2 == 0]),
            "counts": np.array([[k, 2 * k]], np.int32)}


def _config():
    return Config(
        WriterGroupConfig(lambda s: {"positions": s["pos"], "n": np.int64(4),
                                     "labels": [np.uint8(7), np.int16(-3)]}, Once()),
        WriterGroupConfig(lambda s: Step(s["e"], np.full((2, 3, 3), s["e"][0]), s["flags"],
                                         s["counts"]), EveryNStep(3)))


def _write(path, total=8):
    with HDF5StorageWriter(path, _config(), _state(0), total) as w:
        for k in range(total):
            w.log(_state(k), k)


def test_layout_and_round_trip(tmp_path):
    """File/group attributes, group and dataset names (JAX key paths), time
    axis ceil(total / n) of EveryNStep and the values of the logged steps."""
    f = tmp_path / "t.h5"
    _write(f)
    with HDF5StorageReader(f) as r:
        assert r.list_groups() == ["group.init", "group.step"]
        assert r.attr("config_class_name") == "Config"
        assert int(r.attr("actual_steps")) == 8
        assert json.loads(r.attr("tree_def_json")) == {"init": "group.init", "step": "group.step"}
        g = r.focus_group("step")
        assert g.paths == ["array.potential_energy", "array.stress", "array.flags",
                           "array.counts"]
        d = g.read()
        assert d["potential_energy"].shape == (3, 2)  # steps 0, 3, 6
        for i, k in enumerate((0, 3, 6)):
            s = _state(k)
            np.testing.assert_array_equal(d["potential_energy"][i], s["e"])
            np.testing.assert_array_equal(d["flags"][i], s["flags"])
            np.testing.assert_array_equal(d["counts"][i], s["counts"])
            assert np.all(d["stress"][i] == s["e"][0])
        assert d["flags"].dtype == np.bool_ and d["counts"].dtype == np.int32
        one = g[1]
        np.testing.assert_array_equal(one["potential_energy"], _state(3)["e"])
        init = r.focus_group("group.init").read()
        np.testing.assert_array_equal(init["positions"], _state(0)["pos"])
        assert init["positions"].dtype == np.float32
        assert init["n"] == 4 and init["n"].dtype == np.int64
        assert init["labels"][0] == 7 and init["labels"][0].dtype == np.uint8
        assert init["labels"][1] == -3 and init["labels"][1].dtype == np.int16
        assert r.focus_group("init").paths == ["array['positions']", "array['n']",
                                               "array['labels'][0]", "array['labels'][1]"]


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_h5dump_reads_the_file(tmp_path):
    """An independent reader (the HDF5 distribution's h5dump) sees the same
    types and values: little-endian IEEE / standard integer types, h5py's
    boolean enum, variable-length UTF-8 string attributes."""
    f = tmp_path / "t.h5"
    _write(f)
    out = subprocess.run([H5DUMP, "-d", "/group.step/array.potential_energy", "-d",
                          "/group.step/array.flags", "-d", "/group.step/array.counts", str(f)],
                         capture_output=True, text=True, check=True).stdout
    assert "H5T_IEEE_F64LE" in out and "H5T_STD_I32LE" in out
    assert re.search(r"H5T_ENUM\s*{\s*H5T_STD_I8LE;\s*\"FALSE\"\s*0;\s*\"TRUE\"\s*1;", out)
    pe = re.search(r'DATASET "/group.step/array.potential_energy".*?DATA \{(.*?)\}', out, re.S)
    vals = [float(x) for x in re.findall(r"-?\d+(?:\.\d+)?", re.sub(r"\(\d+,\d+\):", "", pe.group(1)))]
    np.testing.assert_array_equal(vals, np.concatenate([_state(k)["e"] for k in (0, 3, 6)]))
    fl = re.search(r'DATASET "/group.step/array.flags".*?DATA \{(.*?)\}', out, re.S).group(1)
    assert re.findall(r"TRUE|FALSE", fl) == ["TRUE", "TRUE", "TRUE", "FALSE", "TRUE", "TRUE"]
    attrs = subprocess.run([H5DUMP, "-A", "-g", "/", str(f)], capture_output=True, text=True,
                           check=True).stdout
    assert "STRSIZE H5T_VARIABLE" in attrs and "H5T_CSET_UTF8" in attrs
    assert '"["group.init", "group.step"]"' in attrs
    assert re.search(r'ATTRIBUTE "actual_steps" {\s*DATATYPE\s+H5T_STD_I64LE', attrs)


def test_writer_errors(tmp_path):
    w = HDF5StorageWriter(tmp_path / "x.h5", _config(), _state(0), 4)
    with pytest.raises(RuntimeError):
        w.log(_state(0), 0)
    with HDF5StorageWriter(tmp_path / "y.h5", _config(), _state(0), 4) as w:
        bad = _state(0)
        bad["e"] = np.zeros(3)
        with pytest.raises(ValueError, match="leaf shape"):
            w.log(bad, 0)
    with pytest.raises(TypeError, match="WriterGroupConfig"):
        with HDF5StorageWriter(tmp_path / "z.h5", {"a": 1}, _state(0), 4):
            pass
    with HDF5StorageReader(tmp_path / "y.h5") as r, pytest.raises(KeyError):
        r.focus_group("missing")


def test_logging_frequencies():
    from kups_amd.application.logging import AtLaunchEnd

    e = EveryNStep(4)
    assert [s for s in range(10) if e.should_log(s)] == [0, 4, 8]
    assert e.leading_shape(10) == (3,) and e.dataset_index(8) == 2
    o = Once()
    assert o.should_log(0) and not o.should_log(1) and o.leading_shape(5) == ()
    a = AtLaunchEnd(4)
    assert [s for s in range(10) if a.should_log(s)] == [3, 7]
    assert a.leading_shape(10) == (2,) and a.dataset_index(7) == 1


def test_acceptance_rates_from_the_history_bits():
    """The GCMC kernel's 100-flag histories (four 32-bit words, bits 0..99)."""
    from kups_amd.application.logging import acceptance_rates

    rng = np.random.default_rng(0)
    flags = rng.random((3, 4, 100)) < rng.random((3, 4, 1))
    words = np.zeros((3, 4, 4), np.uint32)
    for b in range(100):
        words[..., b // 32] |= flags[..., b].astype(np.uint32) << np.uint32(b % 32)
    words[..., 3] |= np.uint32(0xFFFFFFF0)  # bits past 100 are not history
    np.testing.assert_allclose(acceptance_rates(words.view(np.int32)), flags.mean(-1))
