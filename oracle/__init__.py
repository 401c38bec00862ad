"""ctypes bindings of the oracle (CPU restatement) -- TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
use this module, and only as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class Block(C.Structure):
    _fields_ = [("nreal", C.c_int), ("nall", C.c_int), ("nface", C.c_int),
                ("face2cell", _ip), ("S", _dp), ("area", _dp), ("fcenter", _dp), ("rL", _dp),
                ("rR", _dp), ("cell_off", _ip), ("cell2face", _ip), ("cell2cell", _ip),
                ("volume", _dp), ("lsqw", _dp), ("n_bnd", C.c_int), ("n_real_bnd", C.c_int),
                ("bnd_type", _ip), ("bnd_off", _ip), ("bnd_ghosts", _ip), ("n_wall", C.c_int),
                ("wall_faces", _ip), ("W", _dp), ("conv", _dp), ("diss", _dp), ("W0", _dp),
                ("grad", _dp), ("lim", _dp), ("dt", _dp), ("specrad", _dp)]


class Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("aoa_deg", C.c_double),
                ("mach", C.c_double), ("cmac", C.c_double), ("omega", C.c_double),
                ("k", C.c_double), ("criterion", C.c_double), ("stages", C.c_int),
                ("order", C.c_int), ("flux", C.c_int), ("gradient", C.c_int),
                ("limiter", C.c_int), ("max_iter", C.c_int), ("smoothing", C.c_int)]


class Interface(C.Structure):
    _fields_ = [("dst_block", C.c_int), ("origin_block", C.c_int), ("n", C.c_int),
                ("dst_cells", _ip), ("origin_cells", _ip)]


class Monitor(C.Structure):
    _fields_ = [("iter", C.c_int), ("stop", C.c_int), ("rms", C.c_double * 5),
                ("drop", C.c_double * 5), ("cl", C.c_double), ("cd", C.c_double),
                ("cmx", C.c_double), ("cmy", C.c_double), ("cmz", C.c_double)]

    def as_dict(self):
        return {"iter": self.iter, "stop": self.stop, "rms": list(self.rms), "drop": list(self.drop),
                "cl": self.cl, "cd": self.cd, "cmx": self.cmx, "cmy": self.cmy, "cmz": self.cmz}


_lib = None


def build():
    """Compile liboracle.so (plain C, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        bp = C.POINTER(Block)
        for name in ("e3do_save_w0", "e3do_spectral_radius", "e3do_timestep", "e3do_flux_conv_central",
                     "e3do_flux_conv_ausm", "e3do_venkatakrishnan", "e3do_barth_jespersen"):
            getattr(L, name).argtypes = [bp, C.c_double]
        for name in ("e3do_green_gauss", "e3do_least_squares"):
            getattr(L, name).argtypes = [bp]
        for name in ("e3do_flux_diss_roe", "e3do_flux_diss_ausm", "e3do_update"):
            getattr(L, name).argtypes = [bp, C.c_double, C.c_double]
        L.e3do_freestream.argtypes = [C.POINTER(Params), _dp]
        L.e3do_init_flow.argtypes = [bp, _dp]
        L.e3do_bc.argtypes = [bp, _dp, C.c_double]
        L.e3do_rk_coefficients.argtypes = [C.c_int, _dp, _dp]
        L.e3do_residual.argtypes = [bp, C.POINTER(Params), C.c_int]
        L.e3do_runge_kutta.argtypes = [bp, C.POINTER(Params), _dp]
        L.e3do_convergence.argtypes = [bp, _dp]
        L.e3do_aerodynamic.argtypes = [bp, C.POINTER(Params), _dp]
        L.e3do_exchange.argtypes = [bp, C.POINTER(Interface), C.c_int]
        L.e3do_iterate.argtypes = [bp, C.c_int, C.POINTER(Interface), C.c_int, C.POINTER(Params),
                                   C.c_int, _ip, _dp, C.POINTER(Monitor)]
        L.e3do_iterate.restype = C.c_int
        L.e3do_iterate_mt.argtypes = [bp, C.c_int, C.POINTER(Interface), C.c_int, C.POINTER(Params),
                                      C.c_int, _ip, _dp, C.POINTER(Monitor), C.c_int]
        L.e3do_iterate_mt.restype = C.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class OracleBlock:
    """One block: geometry arrays (reference numbering) plus owned state arrays."""

    def __init__(self, g):
        self.g = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in g.items()}
        g = self.g
        na, nr = g["nall"], g["nreal"]
        self.nall, self.nreal = na, nr
        self.W = np.zeros(5 * na)
        self.conv = np.zeros(5 * na)
        self.diss = np.zeros(5 * na)
        self.W0 = np.zeros(5 * na)
        self.grad = np.zeros(15 * na)
        self.lim = np.zeros(5 * na)
        self.dt = np.zeros(max(nr, 1))
        self.specrad = np.zeros(max(nr, 1))
        for k in ("face2cell", "cell_off", "cell2face", "cell2cell", "bnd_type", "bnd_off", "bnd_ghosts",
                  "wall_faces"):
            g[k] = np.ascontiguousarray(g[k], dtype=np.int32)
        for k in ("S", "area", "fcenter", "rL", "rR", "volume", "lsqw"):
            g[k] = np.ascontiguousarray(g[k], dtype=np.float64)
        self.struct = Block(nr, na, g["nface"], _p(g["face2cell"], _ip), _p(g["S"], _dp), _p(g["area"], _dp),
                            _p(g["fcenter"], _dp), _p(g["rL"], _dp), _p(g["rR"], _dp),
                            _p(g["cell_off"], _ip), _p(g["cell2face"], _ip), _p(g["cell2cell"], _ip),
                            _p(g["volume"], _dp), _p(g["lsqw"], _dp), len(g["bnd_type"]),
                            g["n_real_boundaries"], _p(g["bnd_type"], _ip), _p(g["bnd_off"], _ip),
                            _p(g["bnd_ghosts"], _ip), len(g["wall_faces"]), _p(g["wall_faces"], _ip),
                            _p(self.W, _dp), _p(self.conv, _dp), _p(self.diss, _dp), _p(self.W0, _dp),
                            _p(self.grad, _dp), _p(self.lim, _dp), _p(self.dt, _dp), _p(self.specrad, _dp))

    @property
    def ref(self):
        return C.byref(self.struct)


def params(mach, aoa_deg=0.0, flux=0, order=1, gradient=0, limiter=1, omega=1e-16, k=5.0, gamma=1.4,
           cfl=1.5, stages=5, cmac=1.0, max_iter=200, criterion=-8.0, smoothing=0):
    return Params(gamma, cfl, aoa_deg, mach, cmac, omega, k, criterion, stages, order, flux, gradient,
                  limiter, max_iter, smoothing)


def freestream(p):
    fs = np.zeros(5)
    lib().e3do_freestream(C.byref(p), _p(fs, _dp))
    return fs


class Solver:
    """Multi-block oracle solver: Solver::solve semantics on OracleBlocks."""

    def __init__(self, blocks_geom, p, interfaces=True):
        """interfaces=False drops the block interfaces (their ghosts keep their
        values): used to time one block of a multi-block mesh on its own."""
        self.blocks = [OracleBlock(g) for g in blocks_geom]
        self.p = p
        self.fs = freestream(p)
        self.arr = (Block * len(self.blocks))(*[b.struct for b in self.blocks])
        # re-point the array copies at the same state buffers (structs are copied by value)
        self._itf_keep = []
        itfs = []
        for d, g in enumerate(blocks_geom if interfaces else []):
            # destination ghosts are numbered from GhostCount in topology order
            # (DataStructure/src/CompleteMesh.cpp:240-250)
            off = np.asarray(g["itf_off"])
            pos = 0
            for i in range(len(g["itf_origin_block"])):
                n = int(off[i + 1] - off[i])
                dst = np.arange(g["ghost_count"] + pos, g["ghost_count"] + pos + n, dtype=np.int32)
                org = np.ascontiguousarray(g["itf_origin_cells"][off[i]:off[i + 1]], dtype=np.int32)
                pos += n
                self._itf_keep += [dst, org]
                itfs.append(Interface(d, int(g["itf_origin_block"][i]), n, _p(dst, _ip), _p(org, _ip)))
        self.itf = (Interface * max(len(itfs), 1))(*itfs)
        self.n_itf = len(itfs)
        self.it = C.c_int(0)
        self.rms0 = np.zeros(5)

    def init_freestream(self):
        for i, b in enumerate(self.blocks):
            lib().e3do_init_flow(C.byref(self.arr[i]), _p(self.fs, _dp))

    def iterate(self, n, threads=1):
        """n iterations of Solver::solve; threads > 1 spreads each iteration's
        blocks over that many threads in contiguous ranges (same results)."""
        hist = (Monitor * max(n, 1))()
        done = lib().e3do_iterate_mt(self.arr, len(self.blocks), self.itf, self.n_itf, C.byref(self.p), n,
                                     C.byref(self.it), _p(self.rms0, _dp), hist, threads)
        return [hist[i].as_dict() for i in range(done)]


def _log10(x):
    import math
    if x > 0:
        return math.log10(x)
    return -math.inf if x == 0 else math.nan


def _blocks_loop(self, n, stage_iter=None, on_stage=None, on_iter=None):
    """Solver::solve driven from Python through the oracle's primitives, with hooks.

    Same loop as e3do_iterate (Solver.cpp:36-98): per block saveW0, spectral
    radius, time step, then per stage residual (hook) / update / BCs, then the
    block monitors; after all blocks the exchange and the global sums.
    Returns the monitor rows."""
    L = lib()
    p = self.p
    a = np.zeros(5)
    bt = np.zeros(5)
    L.e3do_rk_coefficients(p.stages, _p(a, _dp), _p(bt, _dp))
    if self.it.value == 0:
        L.e3do_exchange(self.arr, self.itf, self.n_itf)
    rows = []
    for _ in range(n):
        it = self.it.value
        brms = np.zeros((len(self.blocks), 10))
        for i in range(len(self.blocks)):
            b = C.byref(self.arr[i])
            L.e3do_save_w0(b, p.gamma)
            L.e3do_spectral_radius(b, p.gamma)
            L.e3do_timestep(b, p.cfl)
            for s in range(p.stages):
                L.e3do_residual(b, C.byref(p), s)
                if on_stage is not None and it == stage_iter:
                    on_stage(it, s, i, self.blocks[i])
                L.e3do_update(b, p.gamma, a[s])
                L.e3do_bc(b, _p(self.fs, _dp), p.gamma)
            out = np.zeros(5)
            L.e3do_convergence(b, _p(out, _dp))
            brms[i, :5] = out
            L.e3do_aerodynamic(b, C.byref(p), _p(out, _dp))
            brms[i, 5:] = out
        L.e3do_exchange(self.arr, self.itf, self.n_itf)
        tot = np.zeros(10)
        for i in range(len(self.blocks)):  # sequential sum in block order, as getGlobal
            for q in range(10):
                tot[q] += brms[i, q]
        if it == 0:
            self.rms0[:] = tot[:5]
        drop = np.array([_log10(tot[v]) - _log10(self.rms0[v]) for v in range(5)])  # C libm log10
        stop = int(it + 1 == p.max_iter or drop[0] <= p.criterion)
        rows.append({"iter": it, "stop": stop, "rms": tot[:5].tolist(), "drop": drop.tolist(), "cl": tot[5],
                     "cd": tot[6], "cmx": tot[7], "cmy": tot[8], "cmz": tot[9], "block": brms})
        self.it.value = it + 1
        if on_iter is not None:
            on_iter(it, self.blocks)
        if stop:
            break
    return rows


Solver.run = _blocks_loop
