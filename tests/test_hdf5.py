"""The drivers' HDF5 checkpoints without h5py: the package's writer against
the published lookup3 checksum values and, where an interpreter with h5py is
installed on the machine (this image ships one under /opt/conda), against the
HDF5 library itself: files the writer makes are read by h5py, and files
h5py makes are read by the package's reader.  Reference layout:
Main.py:301-321 (time stepping) and 1021-1040 (continuation)."""
import glob
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from spectraldoublediffusiveconvection_amd.dropin import _hdf5


def _h5py_python():
    cands = sorted(glob.glob("/opt/conda/bin/python3.[0-9]*"))
    cands += [shutil.which(n) for n in ("python3.9", "python3.11", "python3.12")]
    for c in cands:
        if not c or not os.path.exists(c) or c.endswith("-config"):
            continue
        r = subprocess.run([c, "-c", "import h5py"], capture_output=True, timeout=120)
        if r.returncode == 0:
            return c
    return None


PY_H5 = _h5py_python()
needs_h5py = pytest.mark.skipif(PY_H5 is None, reason="no interpreter with h5py on this machine")

DUMP = r"""
import h5py, json, sys
import numpy as np
out = {}
def visit(name, obj):
    if isinstance(obj, h5py.Dataset):
        v = obj[()]
        out[name] = {"dtype": str(obj.dtype), "shape": list(obj.shape),
                     "value": (v.decode() if isinstance(v, bytes) else np.asarray(v).tolist())}
    else:
        out[name] = "group"
with h5py.File(sys.argv[1], "r") as f:
    f.visititems(visit)
print(json.dumps(out))
"""


def _tree():
    rng = np.random.default_rng(5)
    return {
        "Checkpoints": {"X_DATA": [rng.standard_normal(12) for _ in range(3)]},
        "Scalar_Data": {"KE": [0.1, 0.2, 0.3], "Time": np.linspace(0.0, 1.0, 7)},
        "Parameters": {"Ra": 2700.0, "Ra_s": 500.0, "N_r": 40, "N_fm": 512,
                       "symmetric": False, "flag": np.bool_(True), "label": "run-7"},
        "Bifurcation": dict({"s%02d" % i: [float(i), -1.5 * i] for i in range(11)},
                            Y_FOLD=[], Iterations=3),
    }


def test_lookup3_published_values():
    """Jenkins' lookup3 hashlittle test values (lookup3.c's driver)."""
    assert _hdf5.lookup3(b"") == 0xDEADBEEF
    assert _hdf5.lookup3(b"", 0xDEADBEEF) == 0xBD5B7DDE
    assert _hdf5.lookup3(b"Four score and seven years ago") == 0x17770551
    assert _hdf5.lookup3(b"Four score and seven years ago", 1) == 0xCD628161


def test_writer_reader_round_trip(tmp_path):
    fn = str(tmp_path / "t.h5")
    tree = _tree()
    _hdf5.write(fn, tree)
    with open(fn, "rb") as fh:
        assert fh.read(8) == b"\x89HDF\r\n\x1a\n"
    got = _hdf5.read(fn)
    assert set(got) == set(tree)
    np.testing.assert_array_equal(got["Checkpoints"]["X_DATA"], np.asarray(tree["Checkpoints"]["X_DATA"]))
    np.testing.assert_array_equal(got["Scalar_Data"]["Time"], tree["Scalar_Data"]["Time"])
    p = got["Parameters"]
    assert p["Ra"].dtype == np.float64 and p["Ra"][()] == 2700.0
    assert p["N_r"].dtype == np.int64 and p["N_r"][()] == 40
    assert p["symmetric"].dtype == np.bool_ and not p["symmetric"][()] and p["flag"][()]
    assert p["label"][()] == b"run-7"
    assert got["Bifurcation"]["Y_FOLD"].shape == (0,)
    assert got["Bifurcation"]["Iterations"][()] == 3
    assert len(got["Bifurcation"]) == 13


def test_reader_rejects_a_corrupted_header(tmp_path):
    fn = str(tmp_path / "t.h5")
    _hdf5.write(fn, {"G": {"x": np.arange(4.0)}})
    raw = bytearray(open(fn, "rb").read())
    raw[raw.rindex(b"OHDR") + 8] ^= 0x01      # a byte inside the root group's header
    open(fn, "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        _hdf5.read(fn)


@needs_h5py
def test_hdf5_library_reads_the_writers_file(tmp_path):
    fn = str(tmp_path / "t.h5")
    tree = _tree()
    _hdf5.write(fn, tree)
    r = subprocess.run([PY_H5, "-c", DUMP, fn], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout)
    for g in ("Checkpoints", "Scalar_Data", "Parameters", "Bifurcation"):
        assert out[g] == "group"
    x = out["Checkpoints/X_DATA"]
    assert x["dtype"] == "float64" and x["shape"] == [3, 12]
    np.testing.assert_array_equal(np.array(x["value"]), np.asarray(tree["Checkpoints"]["X_DATA"]))
    assert out["Scalar_Data/KE"]["value"] == [0.1, 0.2, 0.3]
    assert out["Parameters/Ra"] == {"dtype": "float64", "shape": [], "value": 2700.0}
    assert out["Parameters/N_fm"] == {"dtype": "int64", "shape": [], "value": 512}
    assert out["Parameters/symmetric"] == {"dtype": "bool", "shape": [], "value": False}
    assert out["Parameters/flag"]["value"] is True
    assert out["Parameters/label"]["value"] == "run-7"
    assert out["Bifurcation/Y_FOLD"]["shape"] == [0]
    assert out["Bifurcation/s10"]["value"] == [10.0, -15.0]


@needs_h5py
def test_reader_reads_a_file_made_by_the_hdf5_library(tmp_path):
    fn = str(tmp_path / "lib.h5")
    make = ("import h5py, numpy as np, sys\n"
            "with h5py.File(sys.argv[1], 'w', libver='latest') as f:\n"
            "    g = f.create_group('Checkpoints')\n"
            "    g['X_DATA'] = np.arange(24.0).reshape(2, 12)\n"
            "    f.create_group('Parameters')['symmetric'] = True\n"
            "    f['Parameters']['N_r'] = 30\n")
    r = subprocess.run([PY_H5, "-c", make, fn], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = _hdf5.read(fn)
    np.testing.assert_array_equal(got["Checkpoints"]["X_DATA"], np.arange(24.0).reshape(2, 12))
    assert got["Parameters"]["symmetric"][()] and got["Parameters"]["N_r"][()] == 30


def test_dropin_checkpoints_have_the_reference_groups(tmp_path):
    """_write_checkpoint and _write_continuation without h5py write HDF5 at
    the requested name with Checkpoints / Scalar_Data / Parameters and
    Checkpoints / Parameters / Bifurcation (Main.py:301-321, 1024-1040)."""
    from spectraldoublediffusiveconvection_amd.dropin import Main
    try:
        import h5py  # noqa: F401
        pytest.skip("h5py is importable here: the drivers use it")
    except ImportError:
        pass
    fn = str(tmp_path / "TimeStep_0.h5")
    X = [np.arange(6.0), 2 * np.arange(6.0)]
    assert Main._write_checkpoint(fn, X, {"KE": [1.0, 2.0], "Time": [0.0, 0.5]},
                                  {"Ra": 2700.0, "N_r": 40, "symmetric": True}) == fn
    t = _hdf5.read(fn)
    assert set(t) == {"Checkpoints", "Scalar_Data", "Parameters"}
    np.testing.assert_array_equal(t["Checkpoints"]["X_DATA"], np.asarray(X))
    assert t["Parameters"]["symmetric"][()]
    res = Main.result()
    res.Ra, res.KE, res.X_DATA, res.Ra_DATA, res.Iterations = [1.0, 2.0], [3.0, 4.0], X, [1.0], 2
    fc = str(tmp_path / "Continuation.h5")
    assert Main._write_continuation(fc, res, {"Ra_s": 500.0, "d": 0.35}) == fc
    t = _hdf5.read(fc)
    assert set(t) == {"Checkpoints", "Parameters", "Bifurcation"}
    assert set(t["Bifurcation"]) == set(vars(res))
    np.testing.assert_array_equal(t["Checkpoints"]["Ra_DATA"], [1.0])
    assert t["Bifurcation"]["Y_FOLD"].shape == (0,)
    if PY_H5:
        r = subprocess.run([PY_H5, "-c", DUMP, fc], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        assert json.loads(r.stdout)["Bifurcation/KE"]["value"] == [3.0, 4.0]
