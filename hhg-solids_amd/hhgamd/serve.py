"""Client of the continuous-batching solve server (hhg_server_* C ABI).

``Server(...)`` starts the native server: host preparation threads, HBM
admission, one HIP stream per running solve. ``submit(param)`` takes the
reference's pySWE parameter dictionary (plus ``precision``) and returns a
request id; ``result(id)`` waits and returns ``(t, J)`` with ``J`` of shape
(nt, 3). ``run_stream`` replays a timed arrival list and reports latency
percentiles and throughput.
"""
import ctypes as ct
import threading
import time

import numpy as np

from . import lib as _lib

STATES = {0: "queued", 1: "preparing", 2: "prepared", 3: "running", 4: "done", 5: "failed"}
_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C")
_C = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C")


def _bind(L):
    V, I, D, P, LL = ct.c_void_p, ct.c_int, ct.c_double, ct.c_char_p, ct.c_longlong
    sig = {
        "hhg_server_new": ([D, I, I], V), "hhg_server_delete": ([V], None),
        "hhg_server_submit": ([V, P, V, I], LL), "hhg_server_wait": ([V, LL, D], I),
        "hhg_server_state": ([V, LL], I), "hhg_server_nt": ([V, LL], I),
        "hhg_server_result": ([V, LL, _D, _C, _C, _C], I), "hhg_server_timing": ([V, LL, _D], I),
        "hhg_server_release": ([V, LL], I), "hhg_server_stats": ([V, _D], I),
        "hhg_server_error": ([V, LL], P), "hhg_server_set_policy": ([V, I, D], I),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    return L


class Server:
    def __init__(self, hbm_budget_gb=64.0, max_inflight=8, host_threads=4, policy="least-work", aging_s=2.0):
        self.lib = _bind(_lib.bind(_lib.load()))
        self.ptr = self.lib.hhg_server_new(float(hbm_budget_gb), int(max_inflight), int(host_threads))
        if not self.ptr:
            raise RuntimeError("hhg_server_new: " + self.lib.hhg_last_error().decode())
        self._lock = threading.Lock()
        code = {"fifo": 0, "least-work": 1}.get(policy, -1)
        if self.lib.hhg_server_set_policy(self.ptr, code, float(aging_s)) != 0:
            err = self.lib.hhg_last_error().decode()
            self.close()
            raise RuntimeError("hhg_server_set_policy: " + err)

    def submit(self, param):
        L = self.lib
        st = L.Settings_new()
        try:
            for k in ("nr1", "nr2", "nr3"):
                getattr(L, "Settings_set_" + k)(st, int(param.get(k, 1)))
            for k in ("tmax", "dt", "intensity", "lambda", "tmax_field"):
                getattr(L, "Settings_set_" + k)(st, float(param[k]))
            L.Settings_set_pol_vec(st, np.ascontiguousarray(param["pol_vec"], dtype=np.float64))
            L.Settings_set_phi_vec(st, np.ascontiguousarray(param["phi_vec"], dtype=np.float64))
            prec = {"fp64": 0, "fp32": 1}.get(param.get("precision", "fp64"), -1)
            rid = L.hhg_server_submit(self.ptr, param["path_tb"].encode(), st, prec)
        finally:
            L.Settings_delete(st)
        if rid < 0:
            raise RuntimeError("hhg_server_submit: " + L.hhg_last_error().decode())
        return rid

    def state(self, rid):
        return STATES.get(self.lib.hhg_server_state(self.ptr, rid), "unknown")

    def wait(self, rid, timeout=-1.0):
        return STATES.get(self.lib.hhg_server_wait(self.ptr, rid, float(timeout)), "unknown")

    def result(self, rid, timeout=-1.0):
        st = self.wait(rid, timeout)
        if st != "done":
            raise RuntimeError(f"request {rid} {st}: {self.lib.hhg_server_error(self.ptr, rid).decode()}")
        nt = self.lib.hhg_server_nt(self.ptr, rid)
        t = np.zeros(nt)
        J = [np.zeros(nt, dtype=np.complex128) for _ in range(3)]
        if self.lib.hhg_server_result(self.ptr, rid, t, *J) != 0:
            raise RuntimeError(self.lib.hhg_last_error().decode())
        return t, np.stack(J, axis=1)

    def error(self, rid):
        return self.lib.hhg_server_error(self.ptr, rid).decode()

    def timing(self, rid):
        out = np.zeros(7)
        self.lib.hhg_server_timing(self.ptr, rid, out)
        keys = ("submit", "prepared", "admitted", "done", "device_ms", "ksteps", "bytes")
        return dict(zip(keys, out.tolist()))

    def release(self, rid):
        self.lib.hhg_server_release(self.ptr, rid)

    def stats(self):
        out = np.zeros(10)
        self.lib.hhg_server_stats(self.ptr, out)
        keys = ("submitted", "completed", "failed", "inflight", "queued", "peak_inflight", "ksteps",
                "bytes_in_use", "peak_bytes")
        return dict(zip(keys, out[:9].tolist()))

    def close(self):
        if self.ptr:
            self.lib.hhg_server_delete(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_stream(server, arrivals, keep_results=False, keep_first=None):
    """Submit ``[(t_arrival_s, param)]`` at their arrival times (relative to
    now), wait for all, and return per-request records and a summary. With
    ``keep_results`` each completed record (only the first ``keep_first`` when
    given) also holds its ``t`` and ``J`` (fetched after the wait; the
    latencies come from the server's clock)."""
    ids = []
    t0 = time.perf_counter()
    for ta, p in arrivals:
        delay = ta - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(server.submit(p))
    recs = []
    for rid, (ta, p) in zip(ids, arrivals):
        st = server.wait(rid)
        tm = server.timing(rid)
        recs.append({"id": rid, "state": st, "latency_s": tm["done"] - tm["submit"],
                     "queue_s": tm["admitted"] - tm["submit"] if tm["admitted"] > 0 else None,
                     "device_ms": tm["device_ms"], "ksteps": tm["ksteps"], "nb_grid": (p["nr1"], p["nr2"]),
                     "t_done": tm["done"], "t_submit": tm["submit"]})
        if keep_results and st == "done" and (keep_first is None or len(recs) <= keep_first):
            recs[-1]["t"], recs[-1]["J"] = server.result(rid)
    ok = [r for r in recs if r["state"] == "done"]
    lat = np.array([r["latency_s"] for r in ok]) if ok else np.zeros(1)
    span = (max(r["t_done"] for r in ok) - min(r["t_submit"] for r in recs)) if ok else 0.0
    summary = {"requests": len(recs), "completed": len(ok), "failed": len(recs) - len(ok),
               "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
               "ksteps_per_s": float(sum(r["ksteps"] for r in ok) / span) if span > 0 else 0.0,
               "span_s": span, "peak_inflight": server.stats()["peak_inflight"]}
    for rid in ids:
        server.release(rid)
    return recs, summary
