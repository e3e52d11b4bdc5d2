"""Record files of the reference (restart FLDAT, statistics FLfld) written by the host-side
codec trans_amd/records.py and read back with Fortran's own unformatted sequential reads
(AMD flang), using the implied-do order of the reference's restart read.  The GPU tests check
that the library writes exactly these bytes."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from trans_amd import build as tb
from trans_amd import records

HERE = os.path.dirname(os.path.abspath(__file__))


def _arrays(n, seed):
    rng = np.random.Generator(np.random.MT19937(seed))
    v = np.asfortranarray(rng.normal(size=(7,) + n))
    rmiu = np.asfortranarray(rng.normal(size=n))
    wt0 = np.asfortranarray(rng.normal(size=(7,) + n))
    wt1 = np.asfortranarray(rng.normal(size=(7,) + n))
    return v, rmiu, wt0, wt1


@pytest.mark.parametrize("kind_dual", [0, 1, 2])
def test_restart_round_trip(kind_dual, tmp_path):
    n = (5, 4, 3)
    v, rmiu, wt0, wt1 = _arrays(n, 3 + kind_dual)
    p = tmp_path / "fld.dat"
    records.write_restart(p, 123, v, rmiu, wt0, wt1, kind_dual)
    r = records.read_restart(p, n[0] + 1, n[1] + 1, n[2] + 1, kind_dual)
    assert r["ki"] == 123
    assert np.array_equal(r["v"], v) and np.array_equal(r["rmiu"], rmiu)
    if kind_dual >= 1:
        assert np.array_equal(r["wt0"], wt0)
    if kind_dual == 2:
        assert np.array_equal(r["wt1"], wt1)
    # 4-byte markers around each record, K fastest within a record
    raw = p.read_bytes()
    assert np.frombuffer(raw[:4], "<i4")[0] == 4 and np.frombuffer(raw[8:12], "<i4")[0] == 4
    rec = np.frombuffer(raw[16:16 + 7 * 60 * 8], "<f8")
    assert rec[1] == v[0, 0, 0, 1] and rec[3] == v[0, 0, 1, 0] and rec[12] == v[0, 1, 0, 0]


@pytest.mark.skipif(not os.path.exists(tb.FC), reason="no Fortran compiler in this image")
@pytest.mark.parametrize("kind_dual", [0, 2])
def test_fortran_reads_the_restart_record(kind_dual, tmp_path):
    n = (6, 5, 4)
    v, rmiu, wt0, wt1 = _arrays(n, 11)
    p = tmp_path / "fld.dat"
    records.write_restart(p, 77, v, rmiu, wt0, wt1, kind_dual)
    exe = tmp_path / "rr"
    subprocess.run([tb.FC, "-O0", "-o", str(exe), os.path.join(HERE, "fortran", "read_restart.f90")],
                   check=True, cwd=tmp_path)
    out = subprocess.run([str(exe), str(p), *map(str, n), str(kind_dual)], check=True, capture_output=True,
                         text=True).stdout.split()
    assert int(out[0]) == 77
    vals = [float(x) for x in out[1:]]
    w1 = wt1 if kind_dual == 2 else np.zeros_like(wt1)
    w0 = wt0 if kind_dual >= 1 else np.zeros_like(wt0)
    want = [v[2, 1, 0, 0], v[6, -1, -1, -1], rmiu[0, -1, 1], w1[4, 0, 1, -1],
            v.sum(), rmiu.sum(), w0.sum(), w1.sum()]
    np.testing.assert_allclose(vals[:4], want[:4], rtol=0, atol=0)
    np.testing.assert_allclose(vals[4:], want[4:], rtol=1e-13)


def test_stats_round_trip(tmp_path):
    n = (4, 3, 2)
    rng = np.random.Generator(np.random.MT19937(5))
    xc, yc, zc = (rng.normal(size=n) for _ in range(3))
    st = rng.normal(size=(23,) + n)
    p = tmp_path / "fld.sta"
    records.write_stats(p, xc, yc, zc, st)
    r = records.read_stats(p, n[0] + 1, n[1] + 1, n[2] + 1)
    assert np.array_equal(r["stats"], st) and np.array_equal(r["xc"], xc) and np.array_equal(r["zc"], zc)
    assert p.stat().st_size == 24 * (8 + 26 * 8)
    # records in I-fastest cell order: the second record is cell (2, 1, 1)
    raw = p.read_bytes()
    assert np.frombuffer(raw[216 + 4:216 + 12], "<f8")[0] == xc[1, 0, 0]
