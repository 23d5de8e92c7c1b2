"""In-situ reduced diagnostics files.

The per-slice reductions run on the device inside the slice loop (hpq_set_insitu); this
module turns one step's arrays into a record of the reference's file format
(src/utils/InsituUtil.H:23-87, written by Fields::InSituWriteToFile,
src/fields/Fields.cpp:1349-1428, and the plasma and beam InSituWriteToFile,
src/particles/plasma/PlasmaParticleContainer.cpp:529-620,
src/particles/beam/BeamParticleContainer.cpp:595-690):
a JSON header describing a NumPy structured datatype ({"names": [...], "formats": [...]},
nested for groups), written once at the start of the file, then one packed little-endian
binary record per written step, appended. Files are named <prefix>/reduced_<name>.<rank>.txt
with the rank zero-padded to four digits; the reference's reader
(tools/read_insitu_diagnostics.py) and `read_file` below parse them.
"""
from __future__ import annotations

import glob
import json
import os
from typing import List, Sequence, Tuple, Union

import numpy as np

FIELD_NAMES = ["[Ex^2]", "[Ey^2]", "[Ez^2]", "[Bx^2]", "[By^2]", "[Bz^2]", "[ExmBy^2]", "[EypBx^2]",
               "[jz_beam]", "[Ez*jz_beam]"]
PLASMA_NAMES = ["sum(w)", "[x]", "[x^2]", "[y]", "[y^2]", "[ux]", "[ux^2]", "[uy]", "[uy^2]", "[uz]", "[uz^2]",
                "[ga]", "[ga^2]", "[(ga-1)*(1-vz)]"]
BEAM_NAMES = ["sum(w)", "[x]", "[x^2]", "[y]", "[y^2]", "[z]", "[z^2]", "[ux]", "[ux^2]", "[uy]", "[uy^2]",
              "[uz]", "[uz^2]", "[x*ux]", "[y*uy]", "[z*uz]", "[x*uy]", "[y*ux]", "[ux/uz]", "[uy/uz]", "[ga]",
              "[ga^2]"]

# A node is (name, numpy array or scalar) or (name, [child nodes]).
Node = Tuple[str, Union[np.ndarray, np.generic, List["Node"]]]


def _leaf(name: str, value, dtype) -> Node:
    return (name, np.ascontiguousarray(np.asarray(value, dtype=dtype).reshape(-1)))


def _format(a: np.ndarray) -> str:
    """NumPy simple datatype of a leaf (InsituUtil.H:31-45): "(size,)" prefix for
    sub-arrays of more than one element, "<" little-endian, kind letter and byte size."""
    kind = {"i": "i", "u": "u", "f": "f", "c": "c"}[a.dtype.kind]
    return (f"({a.size},)" if a.size > 1 else "") + "<" + kind + str(a.dtype.itemsize)


def header(nodes: Sequence[Node], indent: str = "") -> str:
    """JSON header text in the reference's layout (InsituUtil.H:54-78)."""
    out = [indent, "{\n", indent, '    "names": [\n']
    for i, (name, _) in enumerate(nodes):
        out += [indent, '        "', name, '"', "\n" if i + 1 == len(nodes) else ",\n"]
    out += [indent, "    ],\n", indent, '    "formats": [\n']
    for i, (_, v) in enumerate(nodes):
        if isinstance(v, list):
            out.append(header(v, "        "))
        else:
            out += [indent, '        "', _format(v), '"']
        out.append("\n" if i + 1 == len(nodes) else ",\n")
    out += [indent, "    ]\n", indent, "}"]
    return "".join(out)


def data(nodes: Sequence[Node]) -> bytes:
    """Binary record in datatype order (InsituUtil.H:81-91)."""
    parts = []
    for _, v in nodes:
        parts.append(data(v) if isinstance(v, list) else v.astype(v.dtype.newbyteorder("<"), copy=False).tobytes())
    return b"".join(parts)


def append_record(path: str, nodes: Sequence[Node]) -> None:
    """Append one record; a new or empty file gets the header first."""
    with open(path, "ab") as f:
        if f.tell() == 0:
            f.write(header(nodes).encode())
        f.write(data(nodes))


def read_file(pattern: str) -> np.ndarray:
    """All records of the files matching `pattern` as one structured array sorted by time."""
    recs = []
    for path in sorted(glob.glob(pattern)):
        with open(path, "rb") as f:
            raw = f.read()
        obj, offset = json.JSONDecoder().raw_decode(raw.decode(errors="replace"))
        recs.append(np.frombuffer(raw, dtype=np.dtype(obj), offset=offset))
    if not recs:
        raise FileNotFoundError(pattern)
    return np.sort(np.concatenate(recs), order="time")


def file_name(prefix: str, name: str, rank: int = 0) -> str:
    return os.path.join(prefix, f"reduced_{name}.{rank:04d}.txt")


def field_record(time: float, step: int, zlo: float, zhi: float, rdata: np.ndarray,
                 sums: np.ndarray) -> List[Node]:
    """Fields::InSituWriteToFile, Fields.cpp:1368-1396. rdata: (10, nslices)."""
    nz = rdata.shape[1]
    nodes: List[Node] = [_leaf("time", time, "<f8"), _leaf("step", step, "<i4"), _leaf("n_slices", nz, "<i4"),
                         _leaf("z_lo", zlo, "<f8"), _leaf("z_hi", zhi, "<f8"), _leaf("is_normalized_units", 1, "<i4")]
    nodes += [_leaf(n, rdata[i], "<f8") for i, n in enumerate(FIELD_NAMES)]
    nodes.append(("integrated", [_leaf(n, sums[i], "<f8") for i, n in enumerate(FIELD_NAMES)]))
    return nodes


def species_record(kind: str, time: float, step: int, charge: float, mass: float, zlo: float, zhi: float,
                   dxdydz: float, rdata: np.ndarray, idata: np.ndarray, sums: np.ndarray,
                   isum: int) -> List[Node]:
    """PlasmaParticleContainer.cpp:551-595 (kind "plasma") or BeamParticleContainer.cpp:617-674
    (kind "beam"). rdata: (NR, nslices); averages are the integrated sums over the total sum(w)."""
    names = PLASMA_NAMES if kind == "plasma" else BEAM_NAMES
    nz = rdata.shape[1]
    nodes: List[Node] = [_leaf("time", time, "<f8"), _leaf("step", step, "<i4"), _leaf("n_slices", nz, "<i4"),
                         _leaf("charge", charge, "<f8"), _leaf("mass", mass, "<f8"), _leaf("z_lo", zlo, "<f8"),
                         _leaf("z_hi", zhi, "<f8"), _leaf("normalized_density_factor", dxdydz, "<f8"),
                         _leaf("is_normalized_units", 1, "<i4")]
    n_avg = 13 if kind == "plasma" else len(names)
    nodes += [_leaf(names[i], rdata[i], "<f8") for i in range(1, n_avg)]
    if kind == "plasma":
        nodes.append(_leaf(names[13], rdata[13], "<f8"))
    nodes += [_leaf("sum(w)", rdata[0], "<f8"), _leaf("Np", idata, "<i4")]
    with np.errstate(divide="ignore", invalid="ignore"):
        avg = np.asarray(sums, dtype=np.float64) / np.float64(sums[0])
    nodes.append(("average", [_leaf(names[i], avg[i], "<f8") for i in range(1, n_avg)]))
    total: List[Node] = [_leaf("sum(w)", sums[0], "<f8")]
    if kind == "plasma":
        total.append(_leaf(names[13], sums[13], "<f8"))
    total.append(_leaf("Np", isum, "<i4"))
    nodes.append(("total", total))
    return nodes
