"""Host-side readers and writers of the reference's record files, for Python hosts and tests.

Fortran unformatted sequential files: every record is framed by its byte count as a 4-byte
little-endian integer before and after it.
- restart (FLDAT/FLDATb; Output/OUTPUT_SaveField.F90:14-26, Input/Init.F90:22-46): KI (int32),
  V(1:7), Rmiu, then Wt0 (Kind_dual 1) or Wt0 and Wt1 (Kind_dual 2), each over the interior
  cells with K fastest, then J, I and the component outermost (float64);
- statistics (FLfld; Output/Average_Field.F90:264-281): one record per interior cell, I fastest,
  of Xc Yc Zc and the 23 statistics.
Arrays returned and accepted use the reference layout: components first, bounds 1..N-1 of the
interior cells (shape (ncomp, NI1, NJ1, NK1), Fortran order).
"""
from __future__ import annotations

import numpy as np


def _write_record(f, payload: bytes):
    n = np.array([len(payload)], dtype="<i4").tobytes()
    f.write(n + payload + n)


def _read_record(f) -> bytes:
    head = f.read(4)
    if len(head) != 4:
        raise ValueError("truncated record marker")
    n = int(np.frombuffer(head, dtype="<i4")[0])
    body = f.read(n)
    tail = f.read(4)
    if len(body) != n or len(tail) != 4 or int(np.frombuffer(tail, dtype="<i4")[0]) != n:
        raise ValueError("bad record framing")
    return body


def _rec_order(a: np.ndarray) -> bytes:
    """(ncomp, NI1, NJ1, NK1) -> K fastest, then J, I, component outermost."""
    return np.ascontiguousarray(a, dtype="<f8").tobytes(order="C")


def _from_rec(b: bytes, ncomp: int, n1: int, n2: int, n3: int) -> np.ndarray:
    a = np.frombuffer(b, dtype="<f8")
    if a.size != ncomp * n1 * n2 * n3:
        raise ValueError("record size does not match the block")
    return np.asfortranarray(a.reshape(ncomp, n1, n2, n3))


def write_restart(path, ki: int, v, rmiu, wt0=None, wt1=None, kind_dual: int = 0):
    """v: (7, NI1, NJ1, NK1); rmiu: (NI1, NJ1, NK1); wt0/wt1 as v."""
    with open(path, "wb") as f:
        _write_record(f, np.array([ki], dtype="<i4").tobytes())
        _write_record(f, _rec_order(v))
        _write_record(f, _rec_order(np.asarray(rmiu)[None]))
        if kind_dual >= 1:
            _write_record(f, _rec_order(wt0))
        if kind_dual == 2:
            _write_record(f, _rec_order(wt1))


def read_restart(path, ni: int, nj: int, nk: int, kind_dual: int = 0) -> dict:
    """ni, nj, nk: block points."""
    n = (ni - 1, nj - 1, nk - 1)
    out = {}
    with open(path, "rb") as f:
        out["ki"] = int(np.frombuffer(_read_record(f), dtype="<i4")[0])
        out["v"] = _from_rec(_read_record(f), 7, *n)
        out["rmiu"] = _from_rec(_read_record(f), 1, *n)[0]
        if kind_dual >= 1:
            out["wt0"] = _from_rec(_read_record(f), 7, *n)
        if kind_dual == 2:
            out["wt1"] = _from_rec(_read_record(f), 7, *n)
    return out


def write_stats(path, xc, yc, zc, stats):
    """xc, yc, zc: (NI1, NJ1, NK1); stats: (23, NI1, NJ1, NK1)."""
    n1, n2, n3 = np.shape(xc)
    with open(path, "wb") as f:
        for k in range(n3):
            for j in range(n2):
                for i in range(n1):
                    r = np.concatenate([[xc[i, j, k], yc[i, j, k], zc[i, j, k]], stats[:, i, j, k]])
                    _write_record(f, r.astype("<f8").tobytes())


def read_stats(path, ni: int, nj: int, nk: int) -> dict:
    n1, n2, n3 = ni - 1, nj - 1, nk - 1
    xyz = np.zeros((3, n1, n2, n3), order="F")
    st = np.zeros((23, n1, n2, n3), order="F")
    with open(path, "rb") as f:
        for k in range(n3):
            for j in range(n2):
                for i in range(n1):
                    r = np.frombuffer(_read_record(f), dtype="<f8")
                    if r.size != 26:
                        raise ValueError("statistics record of the wrong length")
                    xyz[:, i, j, k] = r[:3]
                    st[:, i, j, k] = r[3:]
    return {"xc": xyz[0], "yc": xyz[1], "zc": xyz[2], "stats": st}
