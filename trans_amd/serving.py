"""Python host of the serving entry points (trans_srv_* in include/trans_hip.h).

A Server admits steady solves by HBM footprint, keeps their state resident on the device and
interleaves their iterations; the scheduling itself runs in the library (serve.h).  This
module only marshals blocks into requests and results back into arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi, mesh
from .solver import TransError

STATE_NAMES = {capi.SRV_QUEUED: "queued", capi.SRV_RUNNING: "running", capi.SRV_CONVERGED: "converged",
               capi.SRV_MAX_ITER: "max_iter", capi.SRV_CANCELLED: "cancelled", capi.SRV_FAILED: "failed"}


def _descs(blocks):
    patches = [mesh.patch_structs(b) for b in blocks]
    descs = (capi.TransBlockDesc * len(blocks))()
    for i, b in enumerate(blocks):
        descs[i].ni, descs[i].nj, descs[i].nk = b.ni, b.nj, b.nk
        descs[i].npatch = len(b.patches)
        descs[i].patches = C.cast(patches[i], C.POINTER(capi.TransPatch))
    return descs, patches


def footprint(blocks, params: capi.TransParams) -> int:
    """Device bytes a context for these blocks would allocate (no device needed)."""
    lib = capi.load_library()
    descs, _keep = _descs(blocks)
    out = C.c_size_t()
    rc = lib.trans_hip_footprint(C.byref(params), len(blocks), descs, C.byref(out))
    if rc:
        raise TransError(rc, "footprint: the solver rejects these inputs")
    return out.value


class Server:
    def __init__(self, hbm_budget: int, max_active: int = 8, chunk: int = 20):
        self.lib = capi.load_library()
        cfg = capi.TransSrvConfig(hbm_budget, max_active, chunk)
        self.srv = C.c_void_p()
        self._check(self.lib.trans_srv_create(C.byref(cfg), C.byref(self.srv)))
        self.shapes = {}

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.trans_srv_last_error(self.srv).decode() if self.srv else ""
            raise TransError(rc, msg)

    def submit(self, blocks, params: capi.TransParams, target_drop: float = 1e-8, max_iter: int = 20000) -> int:
        """Queue one steady solve; the library copies the inputs before returning."""
        n = len(blocks)
        descs, _keep = _descs(blocks)
        geom = (capi.TransGeomView * n)(*[mesh.geom_view(b) for b in blocks])
        state = (capi.TransStateView * n)(*[mesh.state_view(b) for b in blocks])
        req = capi.TransSrvRequest(params, n, descs, geom, state, target_drop, max_iter)
        rid = C.c_int()
        self._check(self.lib.trans_srv_submit(self.srv, C.byref(req), C.byref(rid)))
        self.shapes[rid.value] = [b.shape() for b in blocks]
        return rid.value

    def progress(self, wait: bool = True) -> int:
        """One scheduling round; returns the number of queued or running solves."""
        n = C.c_int()
        self._check(self.lib.trans_srv_progress(self.srv, 1 if wait else 0, C.byref(n)))
        return n.value

    def run(self):
        """Schedule until every submitted solve has finished."""
        while self.progress(wait=True):
            pass

    def poll(self, rid: int) -> dict:
        st = capi.TransSrvStatus()
        self._check(self.lib.trans_srv_poll(self.srv, rid, C.byref(st)))
        d = st.as_dict()
        d["state_name"] = STATE_NAMES.get(d["state"], "?")
        return d

    def cancel(self, rid: int):
        self._check(self.lib.trans_srv_cancel(self.srv, rid))

    def result(self, rid: int) -> list:
        """Final state of every block of a finished solve (reference layout, -2:N+2)."""
        out = []
        for blk, shp in enumerate(self.shapes[rid]):
            arrs = {}
            for nm, nc in (("v", 7), ("pp", 1), ("t", 1), ("rmiu", 1), ("wt0", 7), ("wt1", 7), ("wt2", 7)):
                arrs[nm] = np.zeros(shp if nc == 1 else (nc,) + shp, order="F")
            sv = capi.TransStateView()
            for nm in arrs:
                setattr(sv, nm, capi.fview(arrs[nm]))
            self._check(self.lib.trans_srv_result(self.srv, rid, blk, C.byref(sv)))
            out.append(arrs)
        return out

    def release(self, rid: int):
        self._check(self.lib.trans_srv_release(self.srv, rid))

    def close(self):
        if self.srv:
            self.lib.trans_srv_destroy(self.srv)
            self.srv = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
