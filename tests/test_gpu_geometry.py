"""GPU test of the device geometry pre-processing: the brute-force wall distance
(trans_geom_wall_distance, Geom/CalcDst.F90:121-158) equals the host computation bitwise,
on the flat plate's blocks (panels from every wall patch) and on a random cloud of cells and
panels larger than one LDS tile."""
import time

import numpy as np
import pytest

from trans_amd import cases, mesh, solver

pytestmark = pytest.mark.gpu


def _host_min(cells, panels):
    best = np.full(cells.shape[0], 1.0e20)
    for q in range(0, panels.shape[0], 4096):
        d2 = ((cells[:, None, :] - panels[None, q:q + 4096, :]) ** 2).sum(axis=2).min(axis=1)
        best = np.minimum(best, d2)
    return np.sqrt(np.maximum(best, 1e-20))


def test_wall_distance_random_cloud():
    print("case: 20000 cells x 3000 panels, random", flush=True)
    rng = np.random.Generator(np.random.MT19937(7))
    cells = rng.uniform(-1, 1, (20000, 3))
    panels = rng.uniform(-1, 1, (3000, 3))
    got = mesh.wall_distance_device(cells, panels)
    assert np.array_equal(got, _host_min(cells, panels))


def test_wall_distance_flat_plate_blocks():
    print("case: flat plate 8 blocks, host vs device wall distance", flush=True)
    host, org, glob = cases.flat_plate()
    t0 = time.perf_counter()
    solver.prepare_case(host, org, glob)
    t1 = time.perf_counter()
    dev, org, glob = cases.flat_plate()
    solver.prepare_case(dev, org, glob, device_geometry=True)
    t2 = time.perf_counter()
    for a, b in zip(host, dev):
        assert np.any(a["dst"] != 1.0e3)
        assert np.array_equal(a["dst"], b["dst"])
    print("case set-up: host", round(t1 - t0, 3), "s, device wall distance", round(t2 - t1, 3), "s", flush=True)


def test_no_panels():
    got = mesh.wall_distance_device(np.zeros((5, 3)), np.zeros((0, 3)))
    assert np.all(got == 1.0e3)
