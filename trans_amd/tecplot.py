"""Tecplot ASCII files: the writer's host-side formatter (trans_tecplot_write_ascii of the
library) and a reader for ordered zones with BLOCK data packing.

The format is Tecplot's documented ASCII data format: a TITLE line, a VARIABLES line of
quoted names, one ZONE line with I, J, K, ZONETYPE=ORDERED, DATAPACKING=BLOCK and an
optional VARLOCATION=([m-n]=CELLCENTERED) range, then every variable's values in turn, i
fastest (nodal variables NI*NJ*NK values, cell-centred ones (NI-1)(NJ-1)(NK-1)).
"""
from __future__ import annotations

import ctypes as C
import re

import numpy as np

from . import capi


def write_ascii(path, x, y, z, cellvars: dict, title="", zone=""):
    """Write point coordinates x, y, z (NI x NJ x NK arrays) and cell-centred arrays
    ((NI-1) x (NJ-1) x (NK-1), insertion order) through the library's formatter."""
    lib = capi.load_library()
    x = np.asarray(x, dtype=np.float64)
    ni, nj, nk = x.shape
    pts = [np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F")) for a in (x, y, z)]
    names = list(cellvars)
    cv = np.ascontiguousarray(np.concatenate([np.asarray(cellvars[n], dtype=np.float64).ravel(order="F")
                                              for n in names])) if names else np.zeros(1)
    cn = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    dp = C.POINTER(C.c_double)
    rc = lib.trans_tecplot_write_ascii(str(path).encode(), title.encode(), zone.encode(), ni, nj, nk,
                                       *[a.ctypes.data_as(dp) for a in pts], len(names), cn, cv.ctypes.data_as(dp))
    if rc != 0:
        raise OSError(f"trans_tecplot_write_ascii failed ({rc}) for {path}")


def read_ascii(path):
    """Read a file of the format above: returns {"title", "zone", "dims": (NI, NJ, NK),
    "variables": {name: array}} with nodal arrays NI x NJ x NK and cell-centred ones
    (NI-1) x (NJ-1) x (NK-1), Fortran order."""
    with open(path) as fh:
        text = fh.read()
    head, _, body = text.partition("\nZONE")
    zone_line, _, data = body.partition("\n")
    title = re.search(r'TITLE\s*=\s*"([^"]*)"', head)
    names = re.findall(r'"([^"]*)"', head.split("VARIABLES", 1)[1])
    dims = tuple(int(re.search(rf"\b{a}\s*=\s*(\d+)", zone_line).group(1)) for a in "IJK")
    if "DATAPACKING=BLOCK" not in zone_line.replace(" ", ""):
        raise ValueError("only DATAPACKING=BLOCK is read")
    cell = set()
    m = re.search(r"VARLOCATION\s*=\s*\(\[(\d+)(?:-(\d+))?\]\s*=\s*CELLCENTERED\)", zone_line)
    if m:
        lo = int(m.group(1))
        hi = int(m.group(2) or lo)
        cell = set(range(lo, hi + 1))
    vals = np.array(data.split(), dtype=np.float64)
    out, o = {}, 0
    ni, nj, nk = dims
    for q, n in enumerate(names, start=1):
        shp = (ni - 1, nj - 1, nk - 1) if q in cell else (ni, nj, nk)
        cnt = shp[0] * shp[1] * shp[2]
        out[n] = vals[o:o + cnt].reshape(shp, order="F")
        o += cnt
    if o != vals.size:
        raise ValueError(f"{path}: {vals.size} values, the header announces {o}")
    zn = re.search(r'T\s*=\s*"([^"]*)"', zone_line)
    return {"title": title.group(1) if title else "", "zone": zn.group(1) if zn else "", "dims": dims,
            "variables": out}
