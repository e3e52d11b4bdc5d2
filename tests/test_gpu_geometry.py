"""GPU test of the device geometry pre-processing: the brute-force wall distance
(trans_geom_wall_distance, Geom/CalcDst.F90:121-158) equals the host computation bitwise,
on the flat plate's blocks (panels from every wall patch) and on a random cloud of cells and
panels larger than one LDS tile; the O(N) metrics (trans_geom_metrics, Geom/CalcDst.F90:170-632)
give every geometry array of the prepared case bitwise (thtc, from the device atan2, within
one rounding)."""
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


@pytest.mark.parametrize("case", ["plate", "bump", "thin"])
def test_metrics_equal_host(case):
    """Every array the case set-up produces (centres, SD, Grad, Alagm, aleng, aseng, Vol, rad,
    ghost copies, wall distance, and the ghost geometry built on top of them) is the same with
    the metrics computed on the device.  'thin' has two cells in k, which switches the
    neighbour-distance candidates (CalcDst.F90:430-476)."""
    print(f"case: {case}, host vs device metrics", flush=True)

    def build():
        if case == "plate":
            return cases.flat_plate(nblk=3, ni_total=49, nj=33, nk=4, lo=9, hi=25)
        if case == "bump":
            return cases.bump_channel(33, 17, 5)
        return cases.bump_channel(17, 9, 3)

    host, org, glob = build()
    solver.prepare_case(host, org, glob)
    dev, org, glob = build()
    solver.prepare_case(dev, org, glob, device_geometry=True)
    for a, b in zip(host, dev):
        assert set(a.arrays) == set(b.arrays)
        for nm in a.arrays:
            x, y = a[nm], b[nm]
            if nm == "thtc":
                np.testing.assert_allclose(y, x, rtol=0, atol=4e-15)  # a few ulps of pi
            else:
                assert np.array_equal(x, y), nm


def test_metrics_reject_negative_volume():
    """A block whose i direction is reversed against its own orientation check fails as
    CalcDst stops: no device fault, TRANS_EINVAL mapped to ValueError."""
    blocks, org, glob = cases.bump_channel(17, 9, 5)
    b = blocks[0]
    b.xx[...] = b.xx[::-1, :, :].copy(order="F")
    with pytest.raises(ValueError):
        mesh.compute_geometry(blocks, device=True)
