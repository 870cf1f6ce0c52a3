"""Trajectory files in the reference's HDF5 layout (src/Saver.cpp:184-426,
src/Loader.cpp:102-267).  CPU only.

Pins: a write/read round trip; the written file read back by the HDF5 command-line
dumper (an independent reader); and, where the reference tree is present, the
reference's own trajectory file read with this reader against the committed fixture
values, and the header of a written file against the reference file's header (group
and dataset names, shapes, compound member names and order).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

REF_H5 = "/root/reference/TrainingData/Table_2/Wind_Yarn_1_2_Plain_Size5.h5"
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "wind_yarn12_plain5.npz")


@pytest.fixture(scope="module")
def h5(ys):
    try:
        ys.h5io.lib()
    except OSError as e:
        pytest.skip("HDF5 C library not available: %s" % e)
    return ys.h5io


def _frames(N, T, seed):
    rng = np.random.default_rng(seed)
    return {"LPos": rng.normal(size=(T, N, 3)), "LPosFix": rng.normal(size=(T, N, 3)), "LVel": rng.normal(size=(T, N, 3)),
            "EPos": rng.normal(size=(T, N, 2)), "EPosBar": rng.normal(size=(T, N, 2)), "EVel": rng.normal(size=(T, N, 2))}


@pytest.mark.parametrize("two_yarns", [False, True])
def test_round_trip(ys, h5, tmp_path, two_yarns):
    cs = ys.cloth_spec
    spec = cs.heter_spec(6, "satin", 4) if two_yarns else cs.homo_spec(5, "twill")
    fr = _frames(spec["width"] * spec["length"], 7, 1)
    p = tmp_path / "t.h5"
    h5.save_trajectory(p, spec, fr)
    size, props, got = h5.load_trajectory(p)
    assert size == (spec["width"], spec["length"])
    for k in fr:
        assert np.array_equal(got[k], fr[k]), k
    _, _, part = h5.load_trajectory(p, 2, 5)
    assert np.array_equal(part["LPos"], fr["LPos"][2:5])
    back = h5.spec_from_props(size, props, handles=spec["handles"], h=spec["h"])
    for key in ("width", "length", "L", "R", "kc", "kf", "df", "mu", "S", "handle_k", "pattern", "wind_density",
                "wind_drag", "yarns"):
        assert back[key] == spec[key], key
    assert list(back["wind_vel"]) == list(spec["wind_vel"]) and back["G"][2] == spec["G"][2]


def _dump(path, ds):
    out = subprocess.check_output([H5DUMP, "-y", "-w", "0", "-m", "%.17g", "-d", ds, str(path)]).decode()
    body = out[out.index("DATA {") + 6:]
    return np.array([float(t) for t in re.findall(r"-?\d[\d.eE+-]*", body.split("}")[0])])


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="h5dump not available")
def test_written_file_read_by_h5dump(ys, h5, tmp_path):
    spec = ys.cloth_spec.heter_spec(5, "plain", 2)
    fr = _frames(25, 4, 2)
    p = tmp_path / "d.h5"
    h5.save_trajectory(p, spec, fr)
    got = _dump(p, "/Steps/Position/Node 3 1/EPOSBar")
    assert np.array_equal(got, fr["EPosBar"][:, 3 + 1 * 5, :].ravel())
    got = _dump(p, "/Steps/Velocity/Node 0 4/LVEL")
    assert np.array_equal(got, fr["LVel"][:, 0 + 4 * 5, :].ravel())
    hdr = subprocess.check_output([H5DUMP, "-H", "-d", "/Properties/Physical properties", str(p)]).decode()
    names = re.findall(r'"(\w+)"', hdr.split("DATATYPE")[1])
    assert names == list(h5.HETER_FIELDS) + ["woven_pattern"]


@pytest.mark.skipif(not os.path.exists(REF_H5), reason="reference tree not present")
def test_reads_the_reference_file(h5):
    size, props, fr = h5.load_trajectory(REF_H5, 0, 121)
    fx = np.load(FIXTURE, allow_pickle=False)
    assert size == (5, 5)
    for k, name in (("LPos", "LPOS"), ("LPosFix", "LPOSFix"), ("EPos", "EPOS"), ("EPosBar", "EPOSBar"),
                    ("LVel", "LVEL"), ("EVel", "EVEL")):
        assert np.array_equal(fr[k], fx[name]), k
    names = [str(n) for n in fx["prop_names"]]
    assert list(props) == names
    for n, v in zip(names, fx["props"]):
        assert props[n] == v, n


@pytest.mark.skipif(not (os.path.exists(REF_H5) and os.path.exists(H5DUMP)), reason="reference tree or h5dump absent")
def test_header_matches_the_reference_file(ys, h5, tmp_path):
    def shape(path):
        txt = subprocess.check_output([H5DUMP, "-H", str(path)]).decode()
        txt = re.sub(r"\d+\.\d+e?[-+]?\d*", "X", txt)       # no values in a header, but be safe
        txt = re.sub(r"\( \d+, ", "( T, ", txt)                # the number of steps differs
        txt = re.sub(r"HDF5 \"[^\"]+\"", "HDF5 FILE", txt)
        return txt
    size, props, fr = h5.load_trajectory(REF_H5, 0, 3)
    spec = h5.spec_from_props(size, props)
    p = tmp_path / "same.h5"
    h5.save_trajectory(p, spec, fr)
    assert shape(p) == shape(REF_H5)
