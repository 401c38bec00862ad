"""euler3d_amd: MI355X residual-and-update path of euler3D behind a C ABI.

This package is a thin ctypes layer over ``libeuler3d_amd.so`` (declared in
``include/euler3d.h``). It loads the in-tree library and raises if it is
missing: there is no CPU fallback for any operator.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libeuler3d_amd.so")

E3D_OK = 0
FLUX_ROE, FLUX_AUSM = 0, 1
GRAD_GREEN_GAUSS, GRAD_LEAST_SQUARES = 0, 1
LIM_BARTH_JESPERSEN, LIM_VENKATAKRISHNAN = 0, 1
ARR_W, ARR_W0, ARR_CONV, ARR_DISS, ARR_GRAD, ARR_LIM, ARR_DT, ARR_SPECRAD = range(8)
ORDER_REFERENCE, ORDER_RCM, ORDER_MORTON = 0, 1, 2
TILES_AUTO, TILES_LINEAR, TILES_XCD_CONTIGUOUS, TILES_XCD_CHUNKS = -1, 0, 1, 2
GEOMETRY_AUTO, GEOMETRY_STREAMED = -1, 0
BC_WALL, BC_FARFIELD, BC_SYMMETRY, BC_CONNEXION = range(4)

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class BumpDesc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("h", C.c_double),
                ("brick", C.c_int), ("px", C.c_int), ("py", C.c_int), ("pz", C.c_int)]


class BlockView(C.Structure):
    _fields_ = [("nreal", C.c_int), ("nall", C.c_int), ("nface", C.c_int), ("nnode", C.c_int),
                ("n_boundaries", C.c_int), ("n_real_boundaries", C.c_int), ("n_wall_faces", C.c_int),
                ("face2cell", _ip), ("face_nnode", _ip), ("face_nodes", _ip),
                ("S", _dp), ("area", _dp), ("fcenter", _dp), ("rL", _dp), ("rR", _dp),
                ("cell_off", _ip), ("cell2face", _ip), ("cell2cell", _ip),
                ("volume", _dp), ("ccenter", _dp), ("lsqw", _dp),
                ("bnd_type", _ip), ("bnd_off", _ip), ("bnd_ghosts", _ip),
                ("wall_faces", _ip), ("ghost_count", C.c_int), ("n_interfaces", C.c_int),
                ("itf_origin_block", _ip), ("itf_off", _ip), ("itf_origin_cells", _ip),
                ("cell_node_off", _ip), ("cell_nodes", _ip), ("node_xyz", _dp)]


class Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("cfl", C.c_double), ("aoa_deg", C.c_double),
                ("mach", C.c_double), ("cmac", C.c_double), ("stages", C.c_int),
                ("max_iter", C.c_int), ("criterion", C.c_double), ("smoothing", C.c_int),
                ("flux", C.c_int), ("order", C.c_int), ("gradient", C.c_int),
                ("limiter", C.c_int), ("omega", C.c_double), ("k", C.c_double)]


class Monitor(C.Structure):
    _fields_ = [("iter", C.c_int), ("stop", C.c_int), ("rms", C.c_double * 5),
                ("drop", C.c_double * 5), ("cl", C.c_double), ("cd", C.c_double),
                ("cmx", C.c_double), ("cmy", C.c_double), ("cmz", C.c_double)]

    def as_dict(self):
        return {"iter": self.iter, "stop": self.stop, "rms": list(self.rms), "drop": list(self.drop),
                "cl": self.cl, "cd": self.cd, "cmx": self.cmx, "cmy": self.cmy, "cmz": self.cmz}


class Instance(C.Structure):
    _fields_ = [("mach", C.c_double), ("aoa_deg", C.c_double), ("cmac", C.c_double), ("criterion", C.c_double),
                ("max_iter", C.c_int), ("rms_ref", C.c_double * 5)]


class BatchStatus(C.Structure):
    _fields_ = [("iterations", C.c_int), ("stopped", C.c_int), ("last", Monitor), ("stop_seconds", C.c_double)]


class E3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"euler3d_amd error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load the in-tree shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        sig = {
            "e3d_last_error": (C.c_char_p, [C.c_void_p]),
            "e3d_version": (C.c_char_p, []),
            "e3d_mesh_read": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
            "e3d_mesh_bump": (C.c_int, [C.POINTER(BumpDesc), C.POINTER(C.c_void_p)]),
            "e3d_mesh_write": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p]),
            "e3d_mesh_nblocks": (C.c_int, [C.c_void_p]),
            "e3d_mesh_block": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(BlockView)]),
            "e3d_mesh_free": (None, [C.c_void_p]),
            "e3d_mesh_layout": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _ip, _ip]),
            "e3d_read_input": (C.c_int, [C.c_char_p, C.POINTER(Params), C.c_char_p, C.c_int]),
            "e3d_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
            "e3d_destroy": (None, [C.c_void_p]),
            "e3d_set_params": (C.c_int, [C.c_void_p, C.POINTER(Params)]),
            "e3d_upload_mesh": (C.c_int, [C.c_void_p, C.c_void_p]),
            "e3d_nblocks": (C.c_int, [C.c_void_p]),
            "e3d_init_freestream": (C.c_int, [C.c_void_p]),
            "e3d_get_array": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
            "e3d_set_array": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp]),
            "e3d_flux_diss": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double]),
            "e3d_flux_conv": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "e3d_gradients": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "e3d_limiters": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double]),
            "e3d_bc_marker": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "e3d_residual": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "e3d_update": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
            "e3d_exchange": (C.c_int, [C.c_void_p]),
            "e3d_monitors": (C.c_int, [C.c_void_p, C.POINTER(Monitor)]),
            "e3d_iterate": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Monitor)]),
            "e3d_request_stop": (C.c_int, [C.c_void_p]),
            "e3d_reset_monitors": (C.c_int, [C.c_void_p]),
            "e3d_monitor_state": (C.c_int, [C.c_void_p, _ip, _dp]),
            "e3d_set_monitor_state": (C.c_int, [C.c_void_p, C.c_int, _dp]),
            "e3d_kernel_time": (C.c_int, [C.c_void_p, C.c_char_p, _dp, _ip]),
            "e3d_enable_timing": (C.c_int, [C.c_void_p, C.c_int]),
            "e3d_sync": (C.c_int, [C.c_void_p]),
            "e3d_stream_bandwidth": (C.c_int, [C.c_int, C.c_long, C.c_int, C.c_int, C.c_int, _dp]),
            "e3d_upload_batch": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(Instance), C.c_int]),
            "e3d_batch_size": (C.c_int, [C.c_void_p]),
            "e3d_batch_blocks": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip]),
            "e3d_batch_run": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(BatchStatus)]),
            "e3d_batch_history": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(Monitor), C.c_int]),
            "e3d_set_tile_order": (C.c_int, [C.c_void_p, C.c_int]),
            "e3d_set_tile_cells": (C.c_int, [C.c_void_p, C.c_int]),
            "e3d_set_slot_geometry": (C.c_int, [C.c_void_p, C.c_int]),
            "e3d_mesh_tet_geometry": (C.c_int, [C.c_void_p, C.c_int, _dp, C.POINTER(C.c_uint)]),
            "e3d_slot_geometry_nodes": (C.c_int, [C.c_void_p]),
            "e3d_set_ordering": (C.c_int, [C.c_void_p, C.c_int]),
            "e3d_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_ulonglong,
                                         C.POINTER(C.c_void_p)]),
            "e3d_mlp_destroy": (None, [C.c_void_p]),
            "e3d_mlp_param_count": (C.c_long, [C.c_void_p]),
            "e3d_mlp_get_params": (C.c_int, [C.c_void_p, _fp]),
            "e3d_mlp_set_params": (C.c_int, [C.c_void_p, _fp]),
            "e3d_mlp_set_data": (C.c_int, [C.c_void_p, _fp, _fp, C.c_long]),
            "e3d_mlp_gradients": (C.c_int, [C.c_void_p, _ip, C.c_int, _fp, _fp]),
            "e3d_mlp_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_int, C.c_ulonglong, _fp]),
            "e3d_mlp_predict": (C.c_int, [C.c_void_p, _fp, C.c_long, _fp]),
            "e3d_mlp_predict_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p]),
            "e3d_mlp_time_step": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp]),
        }
        for name in ("e3d_save_w0", "e3d_spectral_radius", "e3d_timestep", "e3d_interpolation",
                     "e3d_smooth_residual", "e3d_bc", "e3d_runge_kutta"):
            sig[name] = (C.c_int, [C.c_void_p, C.c_int])
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, ctx=None):
    if rc < 0:
        msg = lib().e3d_last_error(ctx)
        raise E3DError(rc, msg.decode() if msg else "")
    return rc


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


class Mesh:
    """Host multi-block mesh built in the reference's numbering."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def bump(cls, nx, ny, nz, h=0.1, brick=4, parts=(1, 1, 1)):
        d = BumpDesc(nx, ny, nz, h, brick, parts[0], parts[1], parts[2])
        out = C.c_void_p()
        _check(lib().e3d_mesh_bump(C.byref(d), C.byref(out)))
        return cls(out.value)

    @classmethod
    def read(cls, topology_path):
        out = C.c_void_p()
        _check(lib().e3d_mesh_read(os.fsencode(topology_path), C.byref(out)))
        return cls(out.value)

    def write(self, directory, stem):
        _check(lib().e3d_mesh_write(self._h, os.fsencode(directory), stem.encode()))
        return os.path.join(directory, stem + ".top")

    @property
    def nblocks(self):
        return _check(lib().e3d_mesh_nblocks(self._h))

    def block(self, b):
        v = BlockView()
        _check(lib().e3d_mesh_block(self._h, b, C.byref(v)))
        F, NA, NR = v.nface, v.nall, v.nreal
        out = {"nreal": NR, "nall": NA, "nface": F, "nnode": v.nnode,
               "n_real_boundaries": v.n_real_boundaries, "ghost_count": v.ghost_count}
        out["face2cell"] = _arr(v.face2cell, 2 * F, np.int32)
        out["face_nnode"] = _arr(v.face_nnode, F, np.int32)
        out["face_nodes"] = _arr(v.face_nodes, 4 * F, np.int32)
        for k, n in (("S", 3 * F), ("area", F), ("fcenter", 3 * F), ("rL", 3 * F), ("rR", 3 * F)):
            out[k] = _arr(getattr(v, k), n, np.float64)
        out["cell_off"] = _arr(v.cell_off, NA + 1, np.int32)
        nslot = int(out["cell_off"][-1])
        out["cell2face"] = _arr(v.cell2face, nslot, np.int32)
        out["cell2cell"] = _arr(v.cell2cell, nslot, np.int32)
        out["volume"] = _arr(v.volume, NA, np.float64)
        out["ccenter"] = _arr(v.ccenter, 3 * NA, np.float64)
        out["lsqw"] = _arr(v.lsqw, 3 * int(out["cell_off"][NR]), np.float64)
        out["bnd_type"] = _arr(v.bnd_type, v.n_boundaries, np.int32)
        out["bnd_off"] = _arr(v.bnd_off, v.n_boundaries + 1, np.int32)
        out["bnd_ghosts"] = _arr(v.bnd_ghosts, int(out["bnd_off"][-1]), np.int32)
        out["wall_faces"] = _arr(v.wall_faces, v.n_wall_faces, np.int32)
        out["itf_origin_block"] = _arr(v.itf_origin_block, v.n_interfaces, np.int32)
        out["itf_off"] = _arr(v.itf_off, v.n_interfaces + 1, np.int32)
        out["itf_origin_cells"] = _arr(v.itf_origin_cells, int(out["itf_off"][-1]), np.int32)
        out["cell_node_off"] = _arr(v.cell_node_off, NA + 1, np.int32)
        out["cell_nodes"] = _arr(v.cell_nodes, int(out["cell_node_off"][-1]), np.int32)
        out["node_xyz"] = _arr(v.node_xyz, 3 * v.nnode, np.float64)
        return out

    def layout(self, b, mode):
        """(cell_pos, face_pos): device positions of block b's real cells and faces under
        a memory order (ORDER_REFERENCE, ORDER_RCM, ORDER_MORTON)."""
        v = self.block(b)
        cp = np.empty(v["nreal"], np.int32)
        fp = np.empty(v["nface"], np.int32)
        _check(lib().e3d_mesh_layout(self._h, b, mode, cp.ctypes.data_as(_ip), fp.ctypes.data_as(_ip)))
        return cp, fp

    def tet_geometry(self, b):
        """(node_xyz [nreal x 12], slot_code [nreal]) of block b: the per-cell node copies and
        slot codes from which the Green-Gauss kernel rebuilds S and r on a tet mesh; raises
        E3DError if the block holds other cells or a rebuilt slot differs from the stored metrics."""
        n = self.block(b)["nreal"]
        xyz = np.empty((n, 12), np.float64)
        code = np.empty(n, np.uint32)
        _check(lib().e3d_mesh_tet_geometry(self._h, b, xyz.ctypes.data_as(_dp),
                                           code.ctypes.data_as(C.POINTER(C.c_uint))))
        return xyz, code

    def close(self):
        if self._h and self._h.value:
            lib().e3d_mesh_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_input(path):
    p = Params()
    buf = C.create_string_buffer(4096)
    _check(lib().e3d_read_input(os.fsencode(path), C.byref(p), buf, 4096))
    return p, buf.value.decode()


def make_params(mach, aoa_deg=0.0, flux=FLUX_ROE, order=1, gradient=GRAD_GREEN_GAUSS,
                limiter=LIM_VENKATAKRISHNAN, omega=1e-16, k=5.0, gamma=1.4, cfl=1.5, stages=5,
                cmac=1.0, max_iter=200, criterion=-8.0, smoothing=0):
    return Params(gamma, cfl, aoa_deg, mach, cmac, stages, max_iter, criterion, smoothing, flux, order,
                  gradient, limiter, omega, k)


_ARR_SHAPE = {ARR_W: (5, "nall"), ARR_W0: (5, "nreal"), ARR_CONV: (5, "nreal"),
              ARR_DISS: (5, "nreal"), ARR_GRAD: (15, "nall"), ARR_LIM: (5, "nall"),
              ARR_DT: (1, "nreal"), ARR_SPECRAD: (1, "nreal")}


class Context:
    """Device context: one GPU, one stream, one uploaded multi-block mesh."""

    def __init__(self, device=0):
        out = C.c_void_p()
        _check(lib().e3d_create(device, C.byref(out)))
        self._h = C.c_void_p(out.value)
        self.sizes = []

    def _c(self, rc):
        return _check(rc, self._h)

    def set_params(self, params):
        self.params = params
        self._c(lib().e3d_set_params(self._h, C.byref(params)))

    def upload(self, mesh):
        self._c(lib().e3d_upload_mesh(self._h, mesh._h))
        self.sizes = []
        for b in range(mesh.nblocks):
            v = BlockView()
            _check(lib().e3d_mesh_block(mesh._h, b, C.byref(v)))
            self.sizes.append({"nreal": v.nreal, "nall": v.nall})

    def upload_batch(self, meshes, instances):
        """Uploads meshes as one batch; instances: dicts with mach, aoa_deg, cmac,
        criterion, max_iter (one per mesh)."""
        n = len(meshes)
        hs = (C.c_void_p * n)(*[m._h.value for m in meshes])
        ins = (Instance * n)(*[Instance(d["mach"], d.get("aoa_deg", 0.0), d.get("cmac", 1.0),
                                        d.get("criterion", -8.0), d.get("max_iter", 1000),
                                        (C.c_double * 5)(*d.get("rms_ref", [0.0] * 5))) for d in instances])
        self._c(lib().e3d_upload_batch(self._h, hs, ins, n))
        self.sizes = []
        for mesh in meshes:
            for b in range(mesh.nblocks):
                v = BlockView()
                _check(lib().e3d_mesh_block(mesh._h, b, C.byref(v)))
                self.sizes.append({"nreal": v.nreal, "nall": v.nall})

    def batch_blocks(self, i):
        first, count = C.c_int(), C.c_int()
        self._c(lib().e3d_batch_blocks(self._h, i, C.byref(first), C.byref(count)))
        return list(range(first.value, first.value + count.value))

    def batch_run(self, max_iterations, use_graph=True):
        """Runs the batch; returns (iterations launched, per-instance status dicts)."""
        n = lib().e3d_batch_size(self._h)
        st = (BatchStatus * n)()
        done = self._c(lib().e3d_batch_run(self._h, max_iterations, 1 if use_graph else 0, st))
        return done, [{"iterations": x.iterations, "stopped": x.stopped, "last": x.last.as_dict(),
                       "stop_seconds": x.stop_seconds} for x in st]

    def batch_history(self, i, max_rows):
        """Monitor rows of instance i, at most max_rows (from iteration 0)."""
        rows = (Monitor * max(max_rows, 1))()
        n = self._c(lib().e3d_batch_history(self._h, i, rows, max_rows))
        return [rows[k].as_dict() for k in range(n)]

    def set_tile_order(self, order):
        """TILES_AUTO (default), TILES_LINEAR, TILES_XCD_CONTIGUOUS or TILES_XCD_CHUNKS; addresses only."""
        self._c(lib().e3d_set_tile_order(self._h, int(order)))

    def set_tile_cells(self, cells):
        """Cells per workgroup of the stage kernels: 64, 128 or 256 (default); addresses only."""
        self._c(lib().e3d_set_tile_cells(self._h, int(cells)))

    def set_slot_geometry(self, mode):
        """GEOMETRY_AUTO (default: Green-Gauss slot geometry rebuilt from node copies on tet
        meshes) or GEOMETRY_STREAMED (stored per-slot copies); same results either way."""
        self._c(lib().e3d_set_slot_geometry(self._h, int(mode)))

    def slot_geometry_nodes(self):
        """True when the Green-Gauss stage kernel rebuilds the slot geometry from node copies."""
        return bool(self._c(lib().e3d_slot_geometry_nodes(self._h)))

    def set_ordering(self, mode):
        """Memory order of the next upload: ORDER_REFERENCE, ORDER_RCM or ORDER_MORTON."""
        self._c(lib().e3d_set_ordering(self._h, mode))

    def init_freestream(self):
        self._c(lib().e3d_init_freestream(self._h))

    def get(self, block, which):
        nv, key = _ARR_SHAPE[which]
        n = self.sizes[block][key]
        out = np.empty(nv * n, dtype=np.float64)
        self._c(lib().e3d_get_array(self._h, block, which, out.ctypes.data_as(_dp)))
        return out.reshape(nv, n)

    def set(self, block, which, values):
        nv, key = _ARR_SHAPE[which]
        n = self.sizes[block][key]
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != nv * n:
            raise ValueError(f"expected {nv * n} values, got {a.size}")
        self._c(lib().e3d_set_array(self._h, block, which, a.ctypes.data_as(_dp)))

    def op(self, name, *args):
        return self._c(getattr(lib(), "e3d_" + name)(self._h, *args))

    def monitors(self):
        m = Monitor()
        self._c(lib().e3d_monitors(self._h, C.byref(m)))
        return m.as_dict()

    def iterate(self, n, history=True):
        hist = (Monitor * max(n, 1))() if history else None
        done = self._c(lib().e3d_iterate(self._h, n, hist))
        return [hist[i].as_dict() for i in range(done)] if history else done

    def kernel_time(self, name):
        ms = C.c_double()
        cnt = C.c_int()
        self._c(lib().e3d_kernel_time(self._h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def close(self):
        if self._h and self._h.value:
            lib().e3d_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MLP:
    """Coordinate MLP of the learned warm start (BF16 matrix-core kernels, FP32 master weights)."""

    def __init__(self, n_in, n_out, width=256, depth=5, seed=0, device=0):
        out = C.c_void_p()
        _check(lib().e3d_mlp_create(device, n_in, width, depth, n_out, seed, C.byref(out)))
        self._h = C.c_void_p(out.value)
        self.n_in, self.n_out, self.width, self.depth = n_in, n_out, width, depth

    def param_count(self):
        return int(lib().e3d_mlp_param_count(self._h))

    def get_params(self):
        out = np.empty(self.param_count(), np.float32)
        _check(lib().e3d_mlp_get_params(self._h, out.ctypes.data_as(_fp)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().e3d_mlp_set_params(self._h, p.ctypes.data_as(_fp)))

    def set_data(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        assert x.shape[1] == self.n_in and t.shape[1] == self.n_out and len(x) == len(t)
        self._x, self._t = x, t
        _check(lib().e3d_mlp_set_data(self._h, x.ctypes.data_as(_fp), t.ctypes.data_as(_fp), len(x)))

    def gradients(self, idx):
        idx = np.ascontiguousarray(idx, np.int32)
        g = np.empty(self.param_count(), np.float32)
        loss = C.c_float()
        _check(lib().e3d_mlp_gradients(self._h, idx.ctypes.data_as(_ip), len(idx), C.byref(loss), g.ctypes.data_as(_fp)))
        return loss.value, g

    def train(self, steps, batch=8192, lr=1e-3, cosine=True, seed=0):
        loss = np.empty(max(steps, 1), np.float32)
        _check(lib().e3d_mlp_train(self._h, steps, batch, lr, 1 if cosine else 0, seed, loss.ctypes.data_as(_fp)))
        return loss[:steps]

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((len(x), self.n_out), np.float32)
        _check(lib().e3d_mlp_predict(self._h, x.ctypes.data_as(_fp), len(x), y.ctypes.data_as(_fp)))
        return y

    def predict_noalloc(self, x_ptr, n, y_ptr):
        """Outputs for n feature rows already in device memory (x_ptr [n x n_in],
        y_ptr [n x n_out], float32, addresses as integers); allocates nothing."""
        _check(lib().e3d_mlp_predict_noalloc(self._h, C.c_void_p(x_ptr), n, C.c_void_p(y_ptr)))

    def time_step(self, batch, reps=10):
        ms, fl = C.c_double(), C.c_double()
        _check(lib().e3d_mlp_time_step(self._h, batch, reps, C.byref(ms), C.byref(fl)))
        return ms.value, fl.value

    def close(self):
        if self._h and self._h.value:
            lib().e3d_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
