"""Python host of the C-ABI: the calls the reference's driver makes per iteration
(main.F90:48-211), issued to libtrans_hip.so through ctypes."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi, mesh


class TransError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__(f"trans_hip error {rc}: {msg}")
        self.rc = rc


def stage_performance(inlet, outlet, gamma=1.4):
    """Total-pressure ratio and isentropic efficiency between two plane integrals, from the
    absolute-frame totals: pi = p0_out / p0_in, eta = (pi^((gamma-1)/gamma) - 1) / (T0_out /
    T0_in - 1)."""
    pi = outlet["p0_abs_mass_avg"] / inlet["p0_abs_mass_avg"]
    tau = outlet["t0_abs_mass_avg"] / inlet["t0_abs_mass_avg"]
    return {"pressure_ratio": pi, "temperature_ratio": tau,
            "efficiency": (pi ** ((gamma - 1.0) / gamma) - 1.0) / (tau - 1.0) if tau != 1.0 else float("nan")}


class Solver:
    def __init__(self, blocks, params: capi.TransParams, ki: int = 1):
        self.lib = capi.load_library()
        self.blocks = blocks
        self.params = params
        n = len(blocks)
        self._patches = [mesh.patch_structs(b) for b in blocks]
        descs = (capi.TransBlockDesc * n)()
        for i, b in enumerate(blocks):
            descs[i].ni, descs[i].nj, descs[i].nk = b.ni, b.nj, b.nk
            descs[i].npatch = len(b.patches)
            descs[i].patches = C.cast(self._patches[i], C.POINTER(capi.TransPatch))
        self.ctx = C.c_void_p()
        rc = self.lib.trans_hip_create(C.byref(params), n, descs, C.byref(self.ctx))
        self._check(rc)
        for i, b in enumerate(blocks):
            self._check(self.lib.trans_hip_set_geometry(self.ctx, i, C.byref(mesh.geom_view(b))))
            self._check(self.lib.trans_hip_set_state(self.ctx, i, C.byref(mesh.state_view(b))))
        self.set_iteration(ki)

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.trans_hip_last_error(self.ctx).decode() if self.ctx else ""
            raise TransError(rc, msg)

    def close(self):
        if self.ctx:
            self.lib.trans_hip_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- per-routine drop-ins
    def set_iteration(self, ki):
        self.ki = ki
        self._check(self.lib.trans_hip_set_iteration(self.ctx, ki))

    def set_params(self, params):
        self.params = params
        self._check(self.lib.trans_hip_set_params(self.ctx, C.byref(params)))

    def spectral_radius(self):
        self._check(self.lib.trans_hip_spectral_radius(self.ctx))

    def bc(self, sweep):
        self._check(self.lib.trans_hip_bc(self.ctx, sweep))

    def space(self):
        self._check(self.lib.trans_hip_space(self.ctx))

    def residual(self):
        self._check(self.lib.trans_hip_residual(self.ctx))

    def lusgs_update(self):
        res = (capi.TransResid * len(self.blocks))()
        self._check(self.lib.trans_hip_lusgs_update(self.ctx, res))
        return [r.as_dict() for r in res]

    def rk_update(self, itr_chld):
        res = (capi.TransResid * len(self.blocks))()
        self._check(self.lib.trans_hip_rk_update(self.ctx, itr_chld, res))
        return [r.as_dict() for r in res]

    def plane_integrals(self, blk, direction, index):
        """Mass flow, area and mass-averaged p0, T0 over face plane `index` of direction
        `direction` (1 i, 2 j, 3 k) of block `blk` (0-based)."""
        out = capi.TransPlaneIntegrals()
        self._check(self.lib.trans_hip_plane_integrals(self.ctx, blk, direction, index, C.byref(out)))
        return out.as_dict()

    def update_u(self):
        self._check(self.lib.trans_hip_update_u(self.ctx))

    def iterate(self, n, want_resid=True):
        res = (capi.TransResid * len(self.blocks))() if want_resid else None
        self._check(self.lib.trans_hip_iterate(self.ctx, n, res))
        self.ki += n
        return [r.as_dict() for r in res] if want_resid else None

    def profile(self, on=True):
        self._check(self.lib.trans_hip_profile(self.ctx, 1 if on else 0))

    def profile_stats(self):
        """{kernel name: (total device ms, launches)} since profiling was switched on."""
        n = self.lib.trans_hip_profile_count(self.ctx)
        if n < 0:
            self._check(n)
        out = {}
        buf = C.create_string_buffer(128)
        for i in range(n):
            ms, cnt = C.c_double(), C.c_long()
            self._check(self.lib.trans_hip_profile_get(self.ctx, i, buf, 128, C.byref(ms), C.byref(cnt)))
            out[buf.value.decode()] = (ms.value, cnt.value)
        return out

    def sync(self):
        self._check(self.lib.trans_hip_sync(self.ctx))

    def autotune(self, iters=10):
        """Times every launch configuration and keeps the fastest; returns
        (chosen configuration, list of all timings)."""
        out = (capi.TransTuneEntry * 64)()
        n = C.c_int()
        self._check(self.lib.trans_hip_autotune(self.ctx, iters, out, 64, C.byref(n)))
        return self.tuning(), [out[q].as_dict() for q in range(min(n.value, 64))]

    def set_tuning(self, sweep, streams, graph, tile_a=16, tile_b=16, chunk=8):
        t = capi.TransTuneEntry(sweep, streams, graph, 0.0, tile_a, tile_b, chunk)
        self._check(self.lib.trans_hip_set_tuning(self.ctx, C.byref(t)))

    def tuning(self):
        t = capi.TransTuneEntry()
        self._check(self.lib.trans_hip_get_tuning(self.ctx, C.byref(t)))
        return t.as_dict()

    # ---- restart records and time statistics
    def write_restart(self, blk, ki, path):
        self._check(self.lib.trans_hip_write_restart(self.ctx, blk, ki, str(path).encode()))

    def read_restart(self, blk, path):
        ki = C.c_int()
        self._check(self.lib.trans_hip_read_restart(self.ctx, blk, str(path).encode(), C.byref(ki)))
        return ki.value

    def stats_enable(self, ki_aver, ki_statis):
        self._check(self.lib.trans_hip_stats_enable(self.ctx, ki_aver, ki_statis))

    def average_fields(self):
        self._check(self.lib.trans_hip_average_fields(self.ctx))

    def get_stats(self, blk):
        a = np.zeros((capi.NSTAT,) + self.blocks[blk].shape(), order="F")
        v = capi.fview(a)
        self._check(self.lib.trans_hip_get_stats(self.ctx, blk, C.byref(v)))
        return a

    def set_stats(self, blk, a):
        a = np.asfortranarray(a, dtype=np.float64)
        v = capi.fview(a)
        self._check(self.lib.trans_hip_set_stats(self.ctx, blk, C.byref(v)))

    def write_stats(self, blk, xc, yc, zc, path):
        keep = [np.asfortranarray(x, dtype=np.float64) for x in (xc, yc, zc)]
        views = [capi.fview(x) for x in keep]
        self._check(self.lib.trans_hip_write_stats(self.ctx, blk, *[C.byref(v) for v in views], str(path).encode()))

    def write_tecplot(self, blk, path, scales=None):
        """Tecplot ASCII file of block blk's instantaneous field (point coordinates from the
        block, 31 cell-centred variables computed on the device); scales = (Roref, Vref, Pref,
        Tinf) or None for nondimensional values."""
        b = self.blocks[blk]
        keep = [np.asfortranarray(a, dtype=np.float64) for a in (b.xx, b.yy, b.zz)]
        views = [capi.fview(a) for a in keep]
        sc = None if scales is None else (C.c_double * 4)(*scales)
        self._check(self.lib.trans_hip_write_tecplot(self.ctx, blk, *[C.byref(v) for v in views], sc,
                                                     str(path).encode()))

    def read_stats(self, blk, path):
        self._check(self.lib.trans_hip_read_stats(self.ctx, blk, str(path).encode()))

    # ---- data out
    def set_state(self, blk, arrays):
        """Upload v, pp, t, rmiu, wt0, wt1 and wt2 of block blk from `arrays` (reference
        layout, bounds -2:N+2, as get_state returns them)."""
        keep = {}
        sv = capi.TransStateView()
        for nm in ("v", "pp", "t", "rmiu", "wt0", "wt1", "wt2"):
            keep[nm] = np.asfortranarray(arrays[nm], dtype=np.float64)
            setattr(sv, nm, capi.fview(keep[nm]))
        self._check(self.lib.trans_hip_set_state(self.ctx, blk, C.byref(sv)))

    def get_state(self, blk):
        b = self.blocks[blk]
        out = {}
        shp = b.shape()
        for nm, nc in (("v", 7), ("pp", 1), ("t", 1), ("rmiu", 1), ("wt0", 7), ("wt1", 7), ("wt2", 7)):
            out[nm] = np.zeros(shp if nc == 1 else (nc,) + shp, order="F")
        sv = capi.TransStateView()
        for nm in out:
            setattr(sv, nm, capi.fview(out[nm]))
        self._check(self.lib.trans_hip_get_state(self.ctx, blk, C.byref(sv)))
        return out

    def get_field(self, blk, which):
        b = self.blocks[blk]
        nc = {capi.FIELD_D: 7, capi.FIELD_RDS: 3, capi.FIELD_DTM: 1, capi.FIELD_DQDXYZ: 18,
              capi.FIELD_SHOCK: 1, capi.FIELD_FSCHEME: 1}[which]
        shp = b.shape()
        a = np.zeros(shp if nc == 1 else (nc,) + shp, order="F")
        v = capi.fview(a)
        self._check(self.lib.trans_hip_get_field(self.ctx, blk, which, C.byref(v)))
        return a


def prepare_case(blocks, org, glob, implicit_fp32=0, device_geometry=False):
    """Connectivity, geometry and initial field for a generated case (device_geometry: the
    wall distances by the HIP library's kernel)."""
    from .params import make_params
    prm = make_params(org, glob, implicit_fp32=implicit_fp32)
    mesh.build_connectivity(blocks, nblade=glob.nblade, npassage=glob.Npassage)
    mesh.attach_wall_panels(blocks)
    mesh.compute_geometry(blocks, device=device_geometry)
    mesh.initial_state(blocks, prm, q_dst=org.Q_Dst, if_ns_uniform=org.IF_NS_Uniform)
    return prm
