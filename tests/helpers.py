"""Shared test helpers: generated cases, seeded perturbations and comparisons."""
import numpy as np

from trans_amd import cases, mesh, solver
from trans_amd.mesh import O, sl

SEED = 20261202


def small_bump(ni=33, nj=17, nk=5):
    blocks, org, glob = cases.bump_channel(ni, nj, nk)
    prm = solver.prepare_case(blocks, org, glob)
    return blocks, org, glob, prm


def perturb(blocks, prm, amp=0.01, seed=SEED, turb=False):
    """Seeded +-amp perturbation of rho, u, v, w, p (and k, omega) on every cell incl. ghosts,
    keeping rho*E and T consistent, so that every limiter / entropy-fix branch is hit."""
    rng = np.random.Generator(np.random.MT19937(seed))
    for b in blocks:
        V, PP, T = b["v"], b["pp"], b["t"]
        shp = PP.shape
        r = V[0] * (1 + amp * rng.uniform(-1, 1, shp))
        u = V[1] / V[0] * (1 + amp * rng.uniform(-1, 1, shp)) + amp * rng.uniform(-1, 1, shp)
        v = V[2] / V[0] + amp * rng.uniform(-1, 1, shp)
        w = V[3] / V[0] + amp * rng.uniform(-1, 1, shp)
        p = PP * (1 + amp * rng.uniform(-1, 1, shp))
        V[0], V[1], V[2], V[3] = r, r * u, r * v, r * w
        PP[...] = p
        rad = b["rad"]
        V[4] = p / 0.4 + 0.5 * r * (u * u + v * v + w * w) - 0.5 * (prm.omega1 * rad) ** 2 * r
        T[...] = prm.rxm2 * p / r
        if turb:
            V[5] *= (1 + amp * rng.uniform(-1, 1, shp))
            V[6] *= (1 + amp * rng.uniform(-1, 1, shp))
            b["rmiu"][...] *= (1 + amp * rng.uniform(-1, 1, shp))


def interior(a, b, ncomp_slice=None):
    """Interior cells 1..N-1 of an array in reference layout."""
    I, J, K = sl(1, b.ni - 1), sl(1, b.nj - 1), sl(1, b.nk - 1)
    if a.ndim == 4:
        return a[ncomp_slice if ncomp_slice is not None else slice(None), I, J, K]
    return a[I, J, K]


def relerr(x, ref):
    """max |x - ref| / max |ref| per leading component (or overall for scalars)."""
    if x.ndim == 4:
        return np.array([np.max(np.abs(x[c] - ref[c])) / max(np.max(np.abs(ref[c])), 1e-300)
                         for c in range(x.shape[0])])
    return np.max(np.abs(x - ref)) / max(np.max(np.abs(ref)), 1e-300)
