"""The file boundary of the reference: inputs.txt (QMS_3D.f90:121-165, written
by master.py:16-66), geom.bin (master.py:68-77, QMS_3D.f90:229-232), and the
Fortran sequential output records read by plot.py:141-150."""
import os

import numpy as np
import pytest
from scipy.io import FortranFile

from alpineqms_amd import configs
from alpineqms_amd.qms_io import (InputsError, RecordWriter, TrajectoryError, format_inputs, material_slots,
                                  parse_inputs, read_geom, read_records, write_geom)

# The text master.py writes (master.py:23-66), with its own example values.
MASTER_TEXT = """70,70,70
0.01,0.01,0.01
0.005
0
1
z,35
yes
0,40E-6,0
0,0.05E-9,0
0.05E-9,0,0
0,0,0
1
5,0,0
1
0,0,0
1
0,0,0
2
5
1,1,1
5E7,5E7,5E7
6
1,1,1
0,0,0
geom.bin
"""


def test_parse_master_py_text():
    inp = parse_inputs(MASTER_TEXT)
    assert inp.cells == (70, 70, 70)
    assert inp.h == (0.01, 0.01, 0.01)
    assert inp.dt == 0.005 and inp.steps == 0 and inp.tt == 1
    assert (inp.slice_axis, inp.slice_loc, inp.save_all) == ("z", 35, "yes")
    assert inp.B0 == (0.0, 40e-6, 0.0)
    assert inp.G == ((0.0, 0.05e-9, 0.0), (0.05e-9, 0.0, 0.0), (0.0, 0.0, 0.0))
    assert inp.traj == ((5.0, 0.0, 0), (0.0, 0.0, 0), (0.0, 0.0, 0))
    assert [(m.id, m.mur, m.sigma) for m in inp.materials] == [(5, (1, 1, 1), (5e7, 5e7, 5e7)),
                                                              (6, (1, 1, 1), (0, 0, 0))]
    assert inp.geom_name == "geom.bin"


def test_format_parse_roundtrip_and_plot_py_line1():
    d = configs.config2(N=20, steps=3)
    txt = format_inputs(d.inputs)
    assert txt.splitlines()[0].split(",") == [str(c) for c in d.inputs.cells]   # plot.py:79
    assert float(txt.splitlines()[2]) == d.inputs.dt                             # plot.py:80
    back = parse_inputs(txt)
    assert back == d.inputs


def test_list_directed_semantics():
    # blanks as separators, extra items ignored, a record continued on the next line, slash ends a record
    t = MASTER_TEXT.replace("70,70,70", "70 70 70 99").replace("0.01,0.01,0.01", "0.01,0.01\n0.01")
    t = t.replace("geom.bin\n", "geom.bin / comment\n")
    inp = parse_inputs(t)
    assert inp.cells == (70, 70, 70) and inp.h == (0.01, 0.01, 0.01) and inp.geom_name == "geom.bin"


def test_trajectory_type_other_than_1_stops():
    t = MASTER_TEXT.split("\n")
    t[11] = "2"
    with pytest.raises(TrajectoryError, match="X DIRECTION"):
        parse_inputs("\n".join(t))


def test_short_inputs_is_an_error():
    with pytest.raises(InputsError):
        parse_inputs("\n".join(MASTER_TEXT.split("\n")[:10]))


def test_geom_roundtrip_and_short_file(tmp_path):
    g = np.arange(3 * 4 * 5, dtype=np.float64).reshape(5, 4, 3)      # (cz, cy, cx)
    path = tmp_path / "geom.bin"
    write_geom(str(path), g)
    # master.py writes data.flatten(order='F') of an (x, y, z) array: x fastest
    ref = np.transpose(g, (2, 1, 0)).flatten(order="F")
    assert np.array_equal(np.fromfile(path), ref)
    assert np.array_equal(read_geom(str(path), (3, 4, 5)), g)
    with open(path, "ab") as f:                         # trailing data is ignored
        f.write(np.ones(7).tobytes())
    assert np.array_equal(read_geom(str(path), (3, 4, 5)), g)
    with pytest.raises(InputsError):
        read_geom(str(path), (4, 4, 5))


def test_material_slots_follow_reference_lookup():
    inp = parse_inputs(MASTER_TEXT)
    geom = np.array([0.0, 5.0, 6.0, 5.9, 3.0, 6.0]).reshape(1, 1, 6)
    slots, props = material_slots(inp, geom)
    # ID 0 and IDs absent from the table are vacuum; int() truncates 5.9 -> 5
    assert slots.ravel().tolist() == [0, 1, 2, 1, 0, 2]
    assert props[0].tolist() == [1, 1, 1, 0, 0, 0]
    assert props[1].tolist() == [1, 1, 1, 5e7, 5e7, 5e7]
    with pytest.raises(InputsError):
        material_slots(inp, np.array([[[7.0]]]))
    with pytest.raises(InputsError):
        material_slots(inp, np.array([[[-1.0]]]))


def test_records_are_read_by_scipy_fortranfile(tmp_path):
    path = str(tmp_path / "Hx_all.bin")
    rng = np.random.default_rng(1)
    arrays = [rng.normal(size=(4, 3, 5)) for _ in range(3)]
    w = RecordWriter(path)
    for a in arrays:
        w.write(a)
    w.close()
    f = FortranFile(path, "r")                       # exactly how plot.py reads them
    for a in arrays:
        assert np.array_equal(f.read_reals(np.float64), a.ravel())
    f.close()
    back = read_records(path)
    assert len(back) == 3 and all(np.array_equal(b, a.ravel()) for a, b in zip(arrays, back))
    # a truncated last record is dropped, like plot.py's reader stopping
    with open(path, "ab") as fh:
        fh.write(np.int32(80).tobytes() + b"\0" * 10)
    assert len(read_records(path)) == 3
