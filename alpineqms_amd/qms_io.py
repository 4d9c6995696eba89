"""File formats of the reference's boundary.

* ``inputs.txt``: one list-directed record per line, read by
  ``QMS_3D.f90:121-165`` and written by ``master.py:16-66``.
* ``geom.bin``: raw float64 material IDs, cells_x*cells_y*cells_z values in
  Fortran order (``master.py:68-77``, read as a stream at ``QMS_3D.f90:229-232``).
* Output records: Fortran unformatted sequential files, each record
  ``[int32 nbytes][float64 payload, Fortran order][int32 nbytes]``
  (``QMS_3D.f90:408-417, 1379-1404, 2275-2300``; read by ``plot.py:141-150``
  through ``scipy.io.FortranFile``).
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np


class InputsError(ValueError):
    """Malformed inputs.txt (the reference stops with a runtime read error)."""


class TrajectoryError(InputsError):
    """Trajectory type other than 1 (QMS_3D.f90:137-151)."""


@dataclass
class Material:
    id: int
    mur: Tuple[float, float, float]
    sigma: Tuple[float, float, float]


@dataclass
class Inputs:
    cells: Tuple[int, int, int]
    h: Tuple[float, float, float]
    dt: float
    steps: int
    tt: int
    slice_axis: str
    slice_loc: int
    save_all: str
    B0: Tuple[float, float, float]
    G: Tuple[Tuple[float, float, float], ...]          # rows gx*, gy*, gz*
    traj: Tuple[Tuple[float, float, int], ...]         # per axis (v0, a, n_a)
    materials: List[Material] = field(default_factory=list)
    geom_name: str = "geom.bin"

    @property
    def nodes(self):
        return tuple(c + 1 for c in self.cells)

    def bc_vector(self) -> np.ndarray:
        return np.array(list(self.B0) + [g for row in self.G for g in row], dtype=np.float64)


# --- list-directed reading -------------------------------------------------

_TOKEN = re.compile(r"\s*(?:'([^']*)'|\"([^\"]*)\"|([^,\s/]+))\s*(,|/|$|(?=\S))")


def _tokens(line: str) -> List[str]:
    """Values of one list-directed record: separated by commas and/or blanks;
    an undelimited value ends at a blank, comma or slash, and a slash ends the
    record."""
    out: List[str] = []
    s = line.strip()
    pos = 0
    while pos < len(s):
        m = _TOKEN.match(s, pos)
        if not m or m.end() == pos:
            break
        tok = m.group(1) if m.group(1) is not None else m.group(2) if m.group(2) is not None else m.group(3)
        out.append(tok)
        pos = m.end()
        if m.group(4) == "/":
            break
    return out


class _Reader:
    """Each read consumes whole records: values missing on a record are taken
    from the following records, extra values on the last record used are
    discarded -- the list-directed behaviour of ``read(1,*)``."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def read(self, n: int, what: str) -> List[str]:
        vals: List[str] = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise InputsError(f"inputs file ended while reading {what}")
            vals.extend(_tokens(self.lines[self.pos]))
            self.pos += 1
        return vals[:n]


def _real(s: str, what: str) -> float:
    try:
        return float(s.replace("d", "e").replace("D", "E"))
    except ValueError as e:
        raise InputsError(f"bad real value {s!r} for {what}") from e


def _int(s: str, what: str) -> int:
    try:
        return int(s)
    except ValueError as e:
        raise InputsError(f"bad integer value {s!r} for {what}") from e


def parse_inputs(text: str) -> Inputs:
    """Parse inputs.txt exactly in the order of ``QMS_3D.f90:123-163``."""
    r = _Reader(text)
    v = r.read(3, "grid sizes")
    cells = tuple(_int(x, "grid size") for x in v)
    h = tuple(_real(x, "grid step") for x in r.read(3, "grid steps"))
    dt = _real(r.read(1, "del_t")[0], "del_t")
    steps = _int(r.read(1, "time_steps")[0], "time_steps")
    tt = _int(r.read(1, "save interval")[0], "save interval")
    sl = r.read(2, "output slice")
    axis = sl[0][:3].strip()
    loc = _int(sl[1], "slice location")
    save_all = r.read(1, "save_all_fields")[0][:3]
    B0 = tuple(_real(x, "background field") for x in r.read(3, "background field"))
    G = tuple(tuple(_real(x, "gradient") for x in r.read(3, "gradient row")) for _ in range(3))
    traj = []
    for axname in "XYZ":
        t = _int(r.read(1, "trajectory type")[0], "trajectory type")
        if t != 1:
            raise TrajectoryError(f"ERROR: TYPE OF TRAJECTORY IN {axname} DIRECTION IS NOT SUPPORTED")
        a = r.read(3, "trajectory")
        traj.append((_real(a[0], "initial velocity"), _real(a[1], "acceleration"), _int(a[2], "acceleration steps")))
    nmat = _int(r.read(1, "num_materials")[0], "num_materials")
    mats = []
    for _ in range(max(nmat, 0)):
        mid = _real(r.read(1, "material ID")[0], "material ID")
        mur = tuple(_real(x, "permeability") for x in r.read(3, "permeability"))
        sig = tuple(_real(x, "conductivity") for x in r.read(3, "conductivity"))
        mats.append(Material(int(mid), mur, sig))
    geom = r.read(1, "geometry file name")[0][:50]
    for c in cells:
        if c < 1:
            raise InputsError("grid sizes must be positive")
    return Inputs(cells=cells, h=h, dt=dt, steps=steps, tt=tt, slice_axis=axis, slice_loc=loc,
                  save_all=save_all, B0=B0, G=G, traj=tuple(traj), materials=mats, geom_name=geom)


def read_inputs(path: str) -> Inputs:
    with open(path, "r") as f:
        return parse_inputs(f.read())


def format_inputs(inp: Inputs) -> str:
    """inputs.txt text in the layout master.py writes (``master.py:23-66``);
    line 1 keeps its commas because ``plot.py:79`` splits on them."""
    L = []
    L.append("{},{},{}".format(*inp.cells))
    L.append("{},{},{}".format(*[repr(float(x)) for x in inp.h]))
    L.append(repr(float(inp.dt)))
    L.append(str(int(inp.steps)))
    L.append(str(int(inp.tt)))
    L.append(f"{inp.slice_axis},{int(inp.slice_loc)}")
    L.append(inp.save_all)
    L.append("{},{},{}".format(*[repr(float(x)) for x in inp.B0]))
    for row in inp.G:
        L.append("{},{},{}".format(*[repr(float(x)) for x in row]))
    for (v0, a, na) in inp.traj:
        L.append("1")
        L.append(f"{float(v0)!r},{float(a)!r},{int(na)}")
    L.append(str(len(inp.materials)))
    for m in inp.materials:
        L.append(str(int(m.id)))
        L.append("{},{},{}".format(*[repr(float(x)) for x in m.mur]))
        L.append("{},{},{}".format(*[repr(float(x)) for x in m.sigma]))
    L.append(inp.geom_name)
    return "\n".join(L) + "\n"


def write_inputs(path: str, inp: Inputs) -> None:
    with open(path, "w") as f:
        f.write(format_inputs(inp))


# --- geometry ----------------------------------------------------------------

def read_geom(path: str, cells) -> np.ndarray:
    """(cz, cy, cx) float64 IDs from a Fortran-order stream file.  A short file
    is an error (the reference's stream read hits end-of-file); trailing data
    is ignored."""
    cx, cy, cz = cells
    n = cx * cy * cz
    data = np.fromfile(path, dtype="<f8", count=n)
    if data.size < n:
        raise InputsError(f"{path}: {data.size} values, the grid needs {n}")
    return data.reshape(cz, cy, cx)


def write_geom(path: str, geom_zyx: np.ndarray) -> None:
    np.ascontiguousarray(geom_zyx, dtype="<f8").tofile(path)


def material_slots(inp: Inputs, geom_zyx: np.ndarray):
    """Map geometry IDs to material slots as ``QMS_3D.f90:234-254`` does:
    IDs are truncated with int(); slot 0 is vacuum (ID 0 and IDs missing from
    the table); an ID above the largest table ID or below 0 is an error (the
    reference indexes out of bounds).  Returns (slot array (cz,cy,cx) uint16,
    slot table (nslots, 6) of mur xyz and sigma xyz)."""
    ids = np.trunc(geom_zyx).astype(np.int64)
    maxid = max([m.id for m in inp.materials] + [0])
    if ids.size and (ids.min() < 0 or ids.max() > maxid):
        raise InputsError("geometry holds a material ID outside the material table")
    table = np.zeros((maxid + 1,), dtype=np.int64)
    for r, m in enumerate(inp.materials):
        table[m.id] = r + 1          # later duplicates win, as in the reference loop
    slots = table[ids].astype(np.uint16)
    props = np.zeros((len(inp.materials) + 1, 6), dtype=np.float64)
    props[0] = (1.0, 1.0, 1.0, 0.0, 0.0, 0.0)
    for r, m in enumerate(inp.materials):
        props[r + 1] = (*m.mur, *m.sigma)
    return slots, props


# --- Fortran unformatted sequential records ----------------------------------

class RecordWriter:
    """Appends records [int32 n][payload][int32 n] (gfortran/ifx default for
    records below 2 GiB)."""

    def __init__(self, path: str):
        self.f = open(path, "wb")

    def write(self, arr: np.ndarray) -> None:
        payload = np.ascontiguousarray(arr, dtype="<f8").tobytes()
        n = len(payload)
        if n >= 2 ** 31:
            raise ValueError("record above 2 GiB")
        marker = np.int32(n).tobytes()
        self.f.write(marker)
        self.f.write(payload)
        self.f.write(marker)
        self.f.flush()

    def close(self):
        self.f.close()


def read_records(path: str) -> List[np.ndarray]:
    """All float64 records of a sequential file (stops at a truncated record,
    like plot.py:144-148)."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos + 4 <= len(data):
        n = int(np.frombuffer(data, "<i4", 1, pos)[0])
        if n < 0 or pos + 8 + n > len(data):
            break
        out.append(np.frombuffer(data, "<f8", n // 8, pos + 4).copy())
        pos += 8 + n
    return out
