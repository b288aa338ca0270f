"""Python handle of the request server (AFS_* in include/afc_hip.h).

Requests carry host arrays; the library keeps arrays resident on the GPU,
admits requests by HBM budget and batches them continuously. The numpy
buffers handed to a submit call must stay alive until its wait returns.
"""
from __future__ import annotations

import ctypes
from ctypes import byref, c_double, c_uint32, c_uint64, c_void_p

import numpy as np

from . import AFCError, dll
from .nativemethods import AFH_PointSet, AFS_Config

OK, ERR_BUDGET = 0, 5


def _err(what, rc):
    if rc != 0:
        raise AFCError(f"{what} failed ({rc}): {dll().AFC_GetLastError().decode()}")


class Server:
    def __init__(self, hbm_budget=1 << 30, pinned_budget=256 << 20, n_streams=4, max_gspat_batch=256,
                 gspat_precision=3, gspat_iterations=100, gspat_cus=0, gspat_graphs=False):
        cfg = AFS_Config(hbm_budget, pinned_budget, n_streams, max_gspat_batch, gspat_precision, gspat_iterations,
                         gspat_cus, 1 if gspat_graphs else 0)
        self.handle = c_void_p()
        _err("AFS_Create", dll().AFS_Create(byref(cfg), byref(self.handle)))
        self._keep = {}

    def close(self):
        if self.handle:
            dll().AFS_Destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_array(self, source_type, sources, wavenums, attens) -> int:
        src = np.ascontiguousarray(sources, np.float32)
        k = np.ascontiguousarray(wavenums, np.float32)
        a = np.ascontiguousarray(attens, np.float32)
        aid = c_uint32()
        _err("AFS_AddArray", dll().AFS_AddArray(self.handle, src.ctypes.data_as(c_void_p), k.ctypes.data_as(c_void_p),
                                               a.ctypes.data_as(c_void_p), k.shape[0], source_type, byref(aid)))
        return aid.value

    def submit_field(self, array_id: int, pointset: AFH_PointSet, field_type: int, out: np.ndarray) -> int:
        t = c_uint64()
        rc = dll().AFS_SubmitField(self.handle, array_id, byref(pointset), field_type, out.ctypes.data_as(c_void_p),
                                   byref(t))
        _err("AFS_SubmitField", rc)
        self._keep[t.value] = (out, pointset)
        return t.value

    def submit_gspat(self, array_id: int, foci: np.ndarray, amps: np.ndarray, out_amp: np.ndarray,
                     out_phase: np.ndarray) -> int:
        f = np.ascontiguousarray(foci, np.float32)
        a = np.ascontiguousarray(amps, np.float32)
        t = c_uint64()
        rc = dll().AFS_SubmitGSPAT(self.handle, array_id, f.ctypes.data_as(c_void_p), a.ctypes.data_as(c_void_p),
                                   a.shape[0], out_amp.ctypes.data_as(c_void_p), out_phase.ctypes.data_as(c_void_p),
                                   byref(t))
        _err("AFS_SubmitGSPAT", rc)
        self._keep[t.value] = (out_amp, out_phase)
        return t.value

    def submit_gspat_batch(self, array_ids, foci, amps, out_amp, out_phase) -> list:
        """Several GS-PAT requests at once (AFS_SubmitGSPATBatch): request i is
        (array_ids[i], foci[i], amps[i]) with results in out_amp[i] /
        out_phase[i]. Returns the tickets in order."""
        n = len(array_ids)
        fs = [np.ascontiguousarray(f, np.float32) for f in foci]
        as_ = [np.ascontiguousarray(a, np.float32) for a in amps]
        ids = np.ascontiguousarray(array_ids, np.uint32)
        m = np.array([a.shape[0] for a in as_], np.uint32)
        ptr = lambda xs: np.array([x.ctypes.data for x in xs], np.uint64)
        pf, pa, po, pp = ptr(fs), ptr(as_), ptr(out_amp), ptr(out_phase)
        tickets = np.zeros(n, np.uint64)
        rc = dll().AFS_SubmitGSPATBatch(self.handle, n, ids.ctypes.data_as(c_void_p), pf.ctypes.data_as(c_void_p),
                                        pa.ctypes.data_as(c_void_p), m.ctypes.data_as(c_void_p),
                                        po.ctypes.data_as(c_void_p), pp.ctypes.data_as(c_void_p),
                                        tickets.ctypes.data_as(c_void_p))
        _err("AFS_SubmitGSPATBatch", rc)
        out = [int(t) for t in tickets]
        for t, oa, op in zip(out, out_amp, out_phase):
            self._keep[t] = (oa, op)
        return out

    def wait(self, ticket: int) -> float:
        lat = c_double()
        rc = dll().AFS_Wait(self.handle, ticket, byref(lat))
        self._keep.pop(ticket, None)
        _err("AFS_Wait", rc)
        return lat.value

    def graph_counters(self):
        """(graph launches, graph builds) of the GS-PAT bucket chains."""
        n, b = c_uint64(), c_uint64()
        _err("AFS_GraphCounters", dll().AFS_GraphCounters(self.handle, byref(n), byref(b)))
        return n.value, b.value

    def stats(self, kind=-1, reset=False):
        """(p50 ms, p99 ms, completed) of field (0), GS-PAT (1) or all (-1) requests."""
        p50, p99, n = c_double(), c_double(), c_uint64()
        _err("AFS_Stats", dll().AFS_Stats(self.handle, kind, byref(p50), byref(p99), byref(n), 1 if reset else 0))
        return p50.value, p99.value, n.value
