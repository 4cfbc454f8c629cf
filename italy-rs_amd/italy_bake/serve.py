"""Served bakes over the C ABI (ib_server_*): submit / wait / fetch / release, plus an open-loop
Poisson load generator that reports throughput and latency percentiles."""
import ctypes
import time

import numpy as np

from . import _lib
from ._lib import BakeError, Grid

STATE_NAMES = {0: "queued", 1: "running", 2: "done", 3: "failed"}


class Request(ctypes.Structure):
    _fields_ = [("xyz", ctypes.c_void_p), ("ntri", ctypes.c_uint64), ("sdf", Grid), ("band_voxels", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("occ", Grid), ("mesh_key", ctypes.c_uint64)]


class TicketInfo(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int32), ("status", ctypes.c_int32), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("nbricks", ctypes.c_uint64), ("resident_bytes", ctypes.c_uint64), ("worker", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("budget", "in_use", "cached", "pending", "ready", "running_batches", "batches", "completed",
                 "mesh_hits", "mesh_builds", "meshes_resident", "mesh_bytes", "evictions", "failed",
                 "alloc_failures")]


def _bind(lib):
    if getattr(lib, "_serve_bound", False):
        return lib
    vp, u64 = ctypes.c_void_p, ctypes.c_uint64
    lib.ib_server_create.argtypes = [ctypes.c_int, u64, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(vp)]
    lib.ib_server_destroy.argtypes = [vp]
    lib.ib_server_destroy.restype = None
    lib.ib_server_last_error.restype = ctypes.c_char_p
    lib.ib_server_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(u64)]
    lib.ib_server_info.argtypes = [vp, u64, ctypes.POINTER(TicketInfo)]
    lib.ib_server_wait.argtypes = [vp, u64, ctypes.c_double]
    lib.ib_server_drain.argtypes = [vp, ctypes.c_double]
    lib.ib_server_fetch.argtypes = [vp, u64, vp, u64, vp, vp]
    lib.ib_server_device_result.argtypes = [vp, u64, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp)]
    lib.ib_server_release.argtypes = [vp, u64]
    lib.ib_server_stats.argtypes = [vp, ctypes.POINTER(Counters)]
    lib._serve_bound = True
    return lib


class Server:
    def __init__(self, device=0, hbm_budget_bytes=0, workers=3, max_batch=64):
        self.lib = _bind(_lib.load())
        h = ctypes.c_void_p()
        self._check(self.lib.ib_server_create(int(device), int(hbm_budget_bytes), int(workers), int(max_batch),
                                              ctypes.byref(h)))
        self._h = h
        self._keep = {}  # soups must stay alive until their ticket is done

    def _check(self, st):
        if st != 0:
            raise BakeError(st, self.lib.ib_server_last_error().decode(errors="replace"))

    def close(self):
        if self._h:
            self.lib.ib_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, xyz, sdf_grid=None, band_voxels=0, occ_grid=None, mesh_key=0):
        """mesh_key != 0 names the soup: its LBVH stays resident and later requests with the key reuse it."""
        xyz = np.ascontiguousarray(xyz, np.float32)
        r = Request(ctypes.c_void_p(xyz.ctypes.data) if xyz.size else None, xyz.size // 9,
                    sdf_grid if sdf_grid is not None else Grid(), int(band_voxels), 0,
                    occ_grid if occ_grid is not None else Grid(), int(mesh_key))
        t = ctypes.c_uint64()
        self._check(self.lib.ib_server_submit(self._h, ctypes.byref(r), ctypes.byref(t)))
        self._keep[t.value] = (xyz, sdf_grid, band_voxels, occ_grid)
        return t.value

    def info(self, ticket):
        i = TicketInfo()
        self._check(self.lib.ib_server_info(self._h, ticket, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in TicketInfo._fields_ if k != "reserved"}

    def wait(self, ticket, timeout_s=600.0):
        self._check(self.lib.ib_server_wait(self._h, ticket, float(timeout_s)))

    def drain(self, timeout_s=3600.0):
        self._check(self.lib.ib_server_drain(self._h, float(timeout_s)))

    def fetch(self, ticket):
        """(distances, brick index or None, occupancy words or None) on the host."""
        xyz, g, band, og = self._keep[ticket]
        info = self.info(ticket)
        idx = occ = None
        if band > 0:
            dist = np.empty((max(info["nbricks"], 1), 8, 8, 8), np.float32)
            idx = np.empty((g.nz // 8, g.ny // 8, g.nx // 8), np.int32)
        else:
            dist = np.empty(g.shape(), np.float32) if g is not None else np.empty(0, np.float32)
        if og is not None:
            occ = np.empty(((og.nx + 31) // 32) * og.ny * og.nz, np.uint32)
        self._check(self.lib.ib_server_fetch(self._h, ticket, ctypes.c_void_p(dist.ctypes.data), dist.size,
                                             ctypes.c_void_p(idx.ctypes.data) if idx is not None else None,
                                             ctypes.c_void_p(occ.ctypes.data) if occ is not None else None))
        if band > 0:
            dist = dist[: info["nbricks"]]
        return dist, idx, occ

    def release(self, ticket):
        self._check(self.lib.ib_server_release(self._h, ticket))
        self._keep.pop(ticket, None)

    def counters(self):
        c = Counters()
        self._check(self.lib.ib_server_stats(self._h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in Counters._fields_}


def open_loop(server, requests, rate_hz, duration_s, seed=0):
    """Submit `requests` (list of (xyz, sdf_grid, band_voxels, occ_grid[, mesh_key])) drawn uniformly at random
    at Poisson arrival times of `rate_hz` for `duration_s`.  Like a client that fetches as soon as
    a result is ready, finished tickets are released while later ones are still arriving.
    Returns throughput and end-to-end latency percentiles (submit -> result resident on the device)."""
    rng = np.random.default_rng(seed)
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t > duration_s:
            break
        arrivals.append((t, int(rng.integers(len(requests)))))
    outstanding, lat, kinds = [], [], []
    failed = 0

    def reap(block):
        nonlocal failed
        keep = []
        for tk, k in outstanding:
            i = server.info(tk)
            if i["state"] < 2 and not block:
                keep.append((tk, k))
                continue
            if i["state"] < 2:
                server.wait(tk)
                i = server.info(tk)
            if i["state"] == 2:
                lat.append(i["t_done"] - i["t_submit"])
                kinds.append(k)
            else:
                failed += 1
            server.release(tk)
        outstanding[:] = keep

    t0 = time.perf_counter()
    for n, (at, k) in enumerate(arrivals):
        delay = t0 + at - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        xyz, g, band, og = requests[k][:4]
        key = requests[k][4] if len(requests[k]) > 4 else 0
        outstanding.append((server.submit(xyz, g, band, og, key), k))
        if n % 16 == 15:
            reap(False)
    reap(True)
    t1 = time.perf_counter()
    lat_a = np.array(lat) if lat else np.zeros(1)
    out = {
        "requests": len(arrivals), "failed": failed, "rate_hz": rate_hz, "duration_s": duration_s,
        "wall_s": t1 - t0, "meshes_per_s": (len(arrivals) - failed) / (t1 - t0),
        "latency_s": {"p50": float(np.percentile(lat_a, 50)), "p95": float(np.percentile(lat_a, 95)),
                      "p99": float(np.percentile(lat_a, 99)), "max": float(lat_a.max()),
                      "mean": float(lat_a.mean())},
        "batches": server.counters()["batches"],
    }
    out["_latencies"] = list(zip(kinds, lat))
    return out


def closed_loop(server, requests, n, concurrency, seed=0):
    """Saturation throughput: keep `concurrency` requests (drawn uniformly at random from `requests`, as in
    open_loop) in flight, submitting the next one as each finishes, until n have finished.  Returns
    meshes/s over the run and the latency percentiles (submit -> result resident on the device)."""
    rng = np.random.default_rng(seed)
    draws = rng.integers(len(requests), size=n)
    outstanding, lat = [], []
    failed = submitted = 0

    def submit():
        nonlocal submitted
        xyz, g, band, og = requests[draws[submitted]][:4]
        key = requests[draws[submitted]][4] if len(requests[draws[submitted]]) > 4 else 0
        outstanding.append(server.submit(xyz, g, band, og, key))
        submitted += 1

    t0 = time.perf_counter()
    while submitted < min(concurrency, n):
        submit()
    while outstanding:
        keep = []
        for tk in outstanding:
            i = server.info(tk)
            if i["state"] < 2:
                keep.append(tk)
                continue
            if i["state"] == 2:
                lat.append(i["t_done"] - i["t_submit"])
            else:
                failed += 1
            server.release(tk)
        outstanding[:] = keep
        while submitted < n and len(outstanding) < concurrency:
            submit()
        if len(outstanding) == len(keep) and outstanding:
            time.sleep(0.0002)
    t1 = time.perf_counter()
    lat_a = np.array(lat) if lat else np.zeros(1)
    return {"requests": n, "concurrency": concurrency, "failed": failed, "wall_s": t1 - t0,
            "meshes_per_s": (n - failed) / (t1 - t0),
            "latency_s": {"p50": float(np.percentile(lat_a, 50)), "p95": float(np.percentile(lat_a, 95)),
                          "p99": float(np.percentile(lat_a, 99)), "max": float(lat_a.max())}}


def capacity(server, requests, n, concurrency, rates, duration_s, seed=0):
    """Served capacity: the closed-loop saturation throughput, then open-loop Poisson runs at the given
    fractions of it.  A rate is sustained when the run drains within 10% of its arrival window (the queue
    does not grow); `sustained_meshes_per_s` is the highest sustained rate of the sweep."""
    sat = closed_loop(server, requests, n, concurrency, seed)
    sweep = []
    for f in rates:
        r = open_loop(server, requests, rate_hz=f * sat["meshes_per_s"], duration_s=duration_s, seed=seed + 1)
        r.pop("_latencies", None)
        r["fraction_of_saturation"] = f
        r["sustained"] = r["failed"] == 0 and r["wall_s"] <= 1.1 * duration_s
        sweep.append(r)
    ok = [r["rate_hz"] for r in sweep if r["sustained"]]
    return {"saturation": sat, "sweep": sweep, "capacity_meshes_per_s": sat["meshes_per_s"],
            "sustained_meshes_per_s": max(ok) if ok else None}
