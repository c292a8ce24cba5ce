"""ctypes binding of libjoemi (include/joemi.h).

The library must have been built in-tree (``python -m myjoe_amd.build``);
loading fails loudly otherwise -- there is no CPU fallback on this path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libjoemi.so")

JM_FP64, JM_MIXED = 0, 1
JM_RHS_IMPLICIT = 1
_STATUS = {0: "JM_OK", 1: "JM_ERR_ARG", 2: "JM_ERR_MESH", 3: "JM_ERR_PARAM", 4: "JM_ERR_HIP",
           5: "JM_ERR_UNSUPPORTED", 6: "JM_ERR_STATE"}

_lib = None


class JoemiError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class SolveReport(C.Structure):
    _fields_ = [("status", C.c_int), ("steps", C.c_int), ("resid", C.c_double * 5), ("resid0", C.c_double * 5),
                ("queued_s", C.c_double), ("run_s", C.c_double), ("fp64_retries", C.c_int),
                ("steps_abandoned", C.c_int)]


class StepInfo(C.Structure):
    _fields_ = [("step", C.c_int), ("lin_iters", C.c_int), ("lin_status", C.c_int), ("cfl", C.c_double),
                ("dt_min", C.c_double), ("resid", C.c_double * 5), ("resid_scal", C.c_double),
                ("scal_iters", C.c_int), ("resid_l2", C.c_double * 5), ("resid_scal2", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m myjoe_amd.build` (HIP path required)")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.jm_last_error.restype = C.c_char_p
        L.jm_version.restype = C.c_char_p
        L.jm_mesh_from_faces.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp]
        L.jm_mesh_read_msh.argtypes = [C.c_char_p, vp]
        L.jm_mesh_read_restart.argtypes = [C.c_char_p, vp]
        L.jm_mesh_restart_field.argtypes = [vp, C.c_char_p, vp, vp]
        L.jm_mesh_restart_step.argtypes = [vp, ip, dp]
        L.jm_write_restart.argtypes = [vp, C.c_char_p, C.c_int, vp, vp, vp, C.c_int, C.c_double]
        L.jm_mesh_from_ugp.argtypes = [vp, vp]
        L.jm_mesh_destroy.argtypes = [vp]
        L.jm_mesh_destroy.restype = None
        L.jm_mesh_counts.argtypes = [vp, vp]
        L.jm_mesh_array.argtypes = [vp, C.c_char_p, vp, vp]
        L.jm_mesh_zone.argtypes = [vp, C.c_int, C.c_char_p, C.c_int, ip, ip, ip]
        L.jm_ctx_create.argtypes = [vp, C.c_char_p, C.c_int, vp, vp]
        L.jm_ctx_destroy.argtypes = [vp]
        L.jm_ctx_destroy.restype = None
        L.jm_set_state.argtypes = [vp, vp, vp, vp]
        L.jm_set_scalar.argtypes = [vp, C.c_char_p, vp]
        L.jm_set_step.argtypes = [vp, C.c_int]
        L.jm_get_state.argtypes = [vp, vp, vp, vp]
        L.jm_init_uniform.argtypes = [vp]
        L.jm_calc_rhs.argtypes = [vp, C.c_int]
        L.jm_get.argtypes = [vp, C.c_char_p, vp, vp]
        L.jm_step.argtypes = [vp, vp]
        L.jm_step_graph.argtypes = [vp, vp]
        L.jm_run.argtypes = [vp, C.c_int, vp, vp]
        L.jm_set_boundary_state.argtypes = [vp, C.c_int, vp, vp, vp]
        L.jm_get_boundary_state.argtypes = [vp, C.c_int, vp, vp, vp]
        L.jm_block_bcgstab.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_double, C.c_double, C.c_int, vp, vp, vp]
        L.jm_block_bcgstab_noalloc.argtypes = L.jm_block_bcgstab.argtypes
        L.jm_block_bcgstab_noalloc_max_cells.argtypes = []
        L.jm_block_bcgstab_noalloc_max_cells.restype = C.c_int64
        L.jm_add_source_coupled.argtypes = [vp, C.c_int, vp, vp]
        L.jm_wall_values.argtypes = [vp, C.c_int, vp, vp]
        L.jm_time_step_kernels.argtypes = [vp, C.c_int, vp, C.c_char_p, C.c_int]
        L.jm_device_state.argtypes = [vp, vp]
        L.jm_device_rhs.argtypes = [vp, vp]
        L.jm_sync.argtypes = [vp]
        L.jm_time_residual_kernels.argtypes = [vp, C.c_int, vp, C.c_char_p, C.c_int]
        L.jm_mesh_quality.argtypes = [vp, vp, vp]
        L.jm_ctx_footprint.argtypes = [vp, C.c_int, vp]
        L.jm_server_create.argtypes = [C.c_int64, C.c_int, C.c_int, vp]
        L.jm_server_destroy.argtypes = [vp]
        L.jm_server_destroy.restype = None
        L.jm_server_submit.argtypes = [vp, vp, C.c_char_p, C.c_int, vp, vp, vp, vp]
        L.jm_server_status.argtypes = [vp, C.c_int, vp]
        L.jm_server_progress.argtypes = [vp, C.c_int, vp]
        L.jm_server_wait.argtypes = [vp, C.c_int]
        L.jm_server_result.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        L.jm_server_release.argtypes = [vp, C.c_int]
        L.jm_chemtable_load.argtypes = [C.c_char_p, vp]
        L.jm_chemtable_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp, vp]
        L.jm_chemtable_destroy.argtypes = [vp]
        L.jm_chemtable_destroy.restype = None
        L.jm_chemtable_info.argtypes = [vp, vp, vp]
        L.jm_chemtable_var.argtypes = [vp, C.c_char_p, ip]
        L.jm_chemtable_lookup.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_int, vp, vp, vp]
        L.jm_chemtable_time_lookup.argtypes = [vp, C.c_int64, vp, vp, vp, C.c_int, vp, vp, vp, C.c_int, dp]
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise JoemiError(st, lib().jm_last_error().decode())


def _ptr(a):
    return a.ctypes.data


class UgpView(C.Structure):
    """jm_ugp_view: the Ugp / UgpWithCv arrays of a host that already holds them."""
    _fields_ = [("ncv", C.c_int), ("nfa", C.c_int), ("nfa_b", C.c_int)] + \
        [(n, C.c_void_p) for n in ("cvofa", "faocv_i", "faocv_v", "nbocv_i", "nbocv_v", "x_fa", "fa_normal",
                                   "x_cv", "cv_volume", "nbocv_lsg", "bfa_lsg")] + \
        [("nzone", C.c_int)] + [(n, C.c_void_p) for n in ("zone_names", "zone_kind", "zone_ifa_f", "zone_ifa_l")]


_HIP = None


def hip():
    """The HIP runtime the library links (ROCm's libamdhip64). Device arrays
    that the library reads must come from this runtime: torch ships its own
    copy, and the two do not share a device in one process."""
    global _HIP
    if _HIP is None:
        lib()
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        _HIP = C.CDLL(os.path.join(rocm, "lib", "libamdhip64.so.7"))
        _HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _HIP.hipFree.argtypes = [C.c_void_p]
        _HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _HIP.hipDeviceSynchronize.argtypes = []
    return _HIP


class DeviceArray:
    """A device copy of a numpy array (hipMalloc of the library's runtime)."""

    def __init__(self, a):
        self.host = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        st = hip().hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 1))
        if st != 0:
            raise JoemiError(4, f"hipMalloc failed ({st})")
        self._copy(self.ptr, self.host.ctypes.data, 1)

    def _copy(self, dst, src, kind):
        st = hip().hipMemcpy(dst, src, self.host.nbytes, kind)  # 1 host->device, 2 device->host
        if st != 0:
            raise JoemiError(4, f"hipMemcpy failed ({st})")

    def numpy(self):
        out = np.empty_like(self.host)
        self._copy(out.ctypes.data, self.ptr, 2)
        return out

    def data_ptr(self):
        return self.ptr.value

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value and _HIP is not None:
            _HIP.hipFree(self.ptr)
            self.ptr = None


def block_bcgstab(nbocv_i, nbocv_v, A, rhs, phi, zero_abs=1e-8, zero_rel=1e-2, max_iter=30, stream=None):
    """jm_block_bcgstab on device arrays (DeviceArray: int32 nbocv_i[ncv+1],
    nbocv_v[nnz]; float64 A[nnz][5][5], rhs[ncv][5], phi[ncv][5], phi
    overwritten with the solution). Returns (iterations, status)."""
    ncv = rhs.host.size // 5
    it, st = C.c_int(), C.c_int()
    _check(lib().jm_block_bcgstab(ncv, nbocv_i.data_ptr(), nbocv_v.data_ptr(), A.data_ptr(), rhs.data_ptr(),
                                  phi.data_ptr(), zero_abs, zero_rel, max_iter, stream, C.byref(it), C.byref(st)))
    return it.value, st.value


def block_bcgstab_noalloc(nbocv_i, nbocv_v, A, rhs, phi, zero_abs=1e-8, zero_rel=1e-2, max_iter=30, stream=None):
    """jm_block_bcgstab_noalloc: block_bcgstab on the library's static work
    space (no device allocation; at most block_bcgstab_noalloc_max_cells()
    cells). Returns (iterations, status)."""
    ncv = rhs.host.size // 5
    it, st = C.c_int(), C.c_int()
    _check(lib().jm_block_bcgstab_noalloc(ncv, nbocv_i.data_ptr(), nbocv_v.data_ptr(), A.data_ptr(),
                                          rhs.data_ptr(), phi.data_ptr(), zero_abs, zero_rel, max_iter, stream,
                                          C.byref(it), C.byref(st)))
    return it.value, st.value


def block_bcgstab_noalloc_max_cells() -> int:
    return int(lib().jm_block_bcgstab_noalloc_max_cells())


class Mesh:
    """A mesh in the reference's numbering, built by libjoemi."""

    def __init__(self, handle):
        self.h = handle
        counts = np.zeros(8, np.int64)
        _check(lib().jm_mesh_counts(self.h, _ptr(counts)))
        self.nno, self.nfa, self.nfa_b, self.ncv, self.nnz, self.nzone, _, _ = (int(x) for x in counts)

    @classmethod
    def from_faces(cls, mm):
        """From a myjoe_amd.meshgen.MshMesh (face-file order, Joe orientation)."""
        h = C.c_void_p()
        x = np.ascontiguousarray(mm.x_no, np.float64)
        fp = np.ascontiguousarray(mm.face_ptr, np.int64)
        fn = np.ascontiguousarray(mm.face_nodes, np.int32)
        fc = np.ascontiguousarray(mm.face_cells, np.int32)
        fz = np.ascontiguousarray(mm.face_zone, np.int32)
        ids = np.array([z[0] for z in mm.zones], np.int32)
        names = (C.c_char_p * len(mm.zones))(*[z[2].encode() for z in mm.zones])
        _check(lib().jm_mesh_from_faces(mm.nno, _ptr(x), mm.nfa, _ptr(fp), _ptr(fn), _ptr(fc), _ptr(fz), mm.ncv,
                                        len(mm.zones), _ptr(ids), C.cast(names, C.c_void_p), C.byref(h)))
        return cls(h)

    @classmethod
    def from_ugp(cls, d: dict):
        """From the reference's own arrays (dump names as in as_dump: CVOFA,
        FAOCV_I/V, NBOCV_I/V, X_FA, FA_NORMAL, X_CV, CV_VOLUME, NBOCV_LSG,
        BFA_LSG, COUNTS, ZONES, ZONE_NAMES), taken over without renumbering."""
        cnt = np.asarray(d["COUNTS"]).reshape(-1)
        nfa, nfa_b, ncv = int(cnt[1]), int(cnt[2]), int(cnt[3])
        keep = {k: np.ascontiguousarray(np.asarray(d[k]).reshape(-1), np.int32)
                for k in ("CVOFA", "FAOCV_I", "FAOCV_V", "NBOCV_I", "NBOCV_V")}
        keep.update({k: np.ascontiguousarray(np.asarray(d[k]).reshape(-1)[:3 * ncv if k == "X_CV" else None],
                                             np.float64)
                     for k in ("X_FA", "FA_NORMAL", "X_CV", "NBOCV_LSG", "BFA_LSG")})
        keep["CV_VOLUME"] = np.ascontiguousarray(np.asarray(d["CV_VOLUME"]).reshape(-1)[:ncv], np.float64)
        zt = np.asarray(d["ZONES"], np.int32).reshape(-1, 3)
        names = bytes(np.asarray(d["ZONE_NAMES"], np.uint8)).decode().split("\n")[:len(zt)]
        zk, zf, zl = (np.ascontiguousarray(zt[:, i]) for i in range(3))
        cnames = (C.c_char_p * len(names))(*[n.encode() for n in names])
        v = UgpView(ncv, nfa, nfa_b, *[keep[k].ctypes.data for k in (
            "CVOFA", "FAOCV_I", "FAOCV_V", "NBOCV_I", "NBOCV_V", "X_FA", "FA_NORMAL", "X_CV", "CV_VOLUME",
            "NBOCV_LSG", "BFA_LSG")], len(names), C.cast(cnames, C.c_void_p), zk.ctypes.data, zf.ctypes.data,
            zl.ctypes.data)
        h = C.c_void_p()
        _check(lib().jm_mesh_from_ugp(C.byref(v), C.byref(h)))
        return cls(h)

    @classmethod
    def read_msh(cls, path):
        h = C.c_void_p()
        _check(lib().jm_mesh_read_msh(path.encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def read_restart(cls, path):
        """A Ugp restart file (restart.out): the mesh, and the file's step, time
        and cell fields in the mesh's numbering (restart_field)."""
        h = C.c_void_p()
        _check(lib().jm_mesh_read_restart(path.encode(), C.byref(h)))
        return cls(h)

    def restart_field(self, name):
        """A cell field read with the mesh from a restart file ([ncv] or [ncv][3])."""
        n = C.c_int64()
        _check(lib().jm_mesh_restart_field(self.h, name.encode(), None, C.byref(n)))
        out = np.empty(n.value, np.float64)
        _check(lib().jm_mesh_restart_field(self.h, name.encode(), _ptr(out), C.byref(n)))
        return out.reshape(-1, 3) if n.value == 3 * self.ncv else out

    def restart_step(self):
        st, t = C.c_int(), C.c_double()
        _check(lib().jm_mesh_restart_step(self.h, C.byref(st), C.byref(t)))
        return st.value, t.value

    def write_restart(self, path, fields: dict, step: int = 0, time: float = 0.0):
        """A Ugp restart file of this mesh and cell fields in its numbering
        (name -> [ncv] or [ncv][3] array), readable by the reference."""
        keep = [np.ascontiguousarray(v, np.float64) for v in fields.values()]
        dims = np.array([1 if a.size == self.ncv else 3 for a in keep], np.int32)
        for a, d in zip(keep, dims):
            if a.size != d * self.ncv:
                raise ValueError("a restart field needs ncv or 3 ncv values")
        names = (C.c_char_p * len(keep))(*[k.encode() for k in fields])
        ptrs = (C.c_void_p * len(keep))(*[a.ctypes.data for a in keep])
        _check(lib().jm_write_restart(self.h, path.encode(), len(keep), C.cast(names, C.c_void_p),
                                      C.cast(ptrs, C.c_void_p), _ptr(dims), int(step), float(time)))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.jm_mesh_destroy(self.h)
            self.h = None

    def array(self, name):
        n = C.c_int64()
        _check(lib().jm_mesh_array(self.h, name.encode(), None, C.byref(n)))
        dt = np.float64 if name in ("X_NO", "X_FA", "FA_NORMAL", "X_CV", "CV_VOLUME", "NBOCV_LSG", "BFA_LSG") \
            else np.int32
        out = np.empty(n.value, dt)
        _check(lib().jm_mesh_array(self.h, name.encode(), _ptr(out), C.byref(n)))
        return out

    def footprint(self, precision=JM_FP64) -> int:
        """Device bytes a solve on this mesh holds."""
        b = C.c_int64()
        _check(lib().jm_ctx_footprint(self.h, precision, C.byref(b)))
        return b.value

    def quality(self):
        q, n = C.c_double(), C.c_int()
        _check(lib().jm_mesh_quality(self.h, C.byref(q), C.byref(n)))
        return q.value, n.value

    def zones(self):
        out = []
        buf = C.create_string_buffer(256)
        k, f, l = C.c_int(), C.c_int(), C.c_int()
        for i in range(self.nzone):
            _check(lib().jm_mesh_zone(self.h, i, buf, 256, C.byref(k), C.byref(f), C.byref(l)))
            out.append((buf.value.decode(), k.value, f.value, l.value))
        return out

    def as_dump(self) -> dict:
        """Arrays under the reference dump names (for the oracle)."""
        d = {n: self.array(n) for n in ("X_NO", "NOOFA_I", "NOOFA_V", "CVOFA", "FAOCV_I", "FAOCV_V", "NBOCV_I",
                                         "NBOCV_V", "X_FA", "FA_NORMAL", "X_CV", "CV_VOLUME", "NBOCV_LSG",
                                         "BFA_LSG")}
        d["COUNTS"] = np.array([self.nno, self.nfa, self.nfa_b, self.ncv, self.ncv, self.nnz,
                                d["FAOCV_V"].size, d["NOOFA_V"].size], np.int32)
        zl = self.zones()
        d["ZONES"] = np.array([[z[1], z[2], z[3]] for z in zl], np.int32).reshape(-1)
        d["ZONE_NAMES"] = np.frombuffer("".join(z[0] + "\n" for z in zl).encode(), np.uint8)
        return d


class Solver:
    """One steady solve resident on the GPU (a JoeWithModels instance)."""

    def __init__(self, mesh: Mesh, joe_in: str, precision: int = JM_FP64, stream=None):
        self.mesh = mesh
        self.h = C.c_void_p()
        _check(lib().jm_ctx_create(mesh.h, joe_in.encode(), precision, stream, C.byref(self.h)))
        self.ncv = mesh.ncv

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.jm_ctx_destroy(self.h)
            self.h = None

    def set_state(self, rho, rhou, rhoE):
        rho = np.ascontiguousarray(rho, np.float64)
        rhou = np.ascontiguousarray(rhou, np.float64)
        rhoE = np.ascontiguousarray(rhoE, np.float64)
        assert rho.size == self.ncv and rhou.size == 3 * self.ncv and rhoE.size == self.ncv
        _check(lib().jm_set_state(self.h, _ptr(rho), _ptr(rhou), _ptr(rhoE)))

    def set_step(self, step: int):
        """After set_state: the step count a solve resumed from a restart file continues from."""
        _check(lib().jm_set_step(self.h, int(step)))

    def set_scalar(self, phi, name: str = "nuSA"):
        phi = np.ascontiguousarray(phi, np.float64)
        assert phi.size == self.ncv
        _check(lib().jm_set_scalar(self.h, name.encode(), _ptr(phi)))

    def state(self):
        rho, rhou, rhoE = np.empty(self.ncv), np.empty((self.ncv, 3)), np.empty(self.ncv)
        _check(lib().jm_get_state(self.h, _ptr(rho), _ptr(rhou), _ptr(rhoE)))
        return rho, rhou, rhoE

    def init_uniform(self):
        _check(lib().jm_init_uniform(self.h))

    def calc_rhs(self, implicit=False):
        _check(lib().jm_calc_rhs(self.h, JM_RHS_IMPLICIT if implicit else 0))

    def get(self, name):
        n = C.c_int64()
        _check(lib().jm_get(self.h, name.encode(), None, C.byref(n)))
        out = np.empty(n.value)
        _check(lib().jm_get(self.h, name.encode(), _ptr(out), C.byref(n)))
        return out

    def step(self) -> StepInfo:
        info = StepInfo()
        _check(lib().jm_step(self.h, C.byref(info)))
        return info

    def step_graph(self) -> StepInfo:
        info = StepInfo()
        _check(lib().jm_step_graph(self.h, C.byref(info)))
        return info

    def run(self, max_steps):
        hist = np.zeros((max_steps, 5))
        n = C.c_int()
        _check(lib().jm_run(self.h, max_steps, _ptr(hist), C.byref(n)))
        return n.value, hist[:n.value]

    def wall_values(self, zone):
        """jm_wall_values of a boundary zone (its index or name): [n][16]."""
        if isinstance(zone, str):
            zone = [z[0] for z in self.mesh.zones()].index(zone)
        n = C.c_int64()
        _check(lib().jm_wall_values(self.h, zone, None, C.byref(n)))
        out = np.empty((n.value, 16))
        _check(lib().jm_wall_values(self.h, zone, _ptr(out), C.byref(n)))
        return out

    def set_boundary_state(self, zone, T, vel, p):
        T = np.ascontiguousarray(T, np.float64)
        vel = np.ascontiguousarray(vel, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        _check(lib().jm_set_boundary_state(self.h, zone, _ptr(T), _ptr(vel), _ptr(p)))

    def boundary_state(self, zone, nface):
        """(T[n], vel[n][3], p[n]) that setBC last computed on a zone's faces."""
        T, vel, p = np.zeros(nface), np.zeros((nface, 3)), np.zeros(nface)
        _check(lib().jm_get_boundary_state(self.h, zone, _ptr(T), _ptr(vel), _ptr(p)))
        return T, vel, p

    def device_state_ptr(self) -> int:
        q = C.c_void_p()
        _check(lib().jm_device_state(self.h, C.byref(q)))
        return q.value

    def device_rhs_ptr(self) -> int:
        q = C.c_void_p()
        _check(lib().jm_device_rhs(self.h, C.byref(q)))
        return q.value

    def sync(self):
        _check(lib().jm_sync(self.h))

    def time_step_kernels(self, reps=10):
        """jm_time_step_kernels: mean ms of residual, Jacobian, Krylov set-up
        and one BiCGStab iteration."""
        ms = np.zeros(8)
        buf = C.create_string_buffer(256)
        _check(lib().jm_time_step_kernels(self.h, reps, _ptr(ms), buf, 256))
        names = buf.value.decode().split(",")
        return dict(zip(names, ms[:len(names)]))

    def time_kernels(self, reps=10):
        ms = np.zeros(8)
        buf = C.create_string_buffer(256)
        _check(lib().jm_time_residual_kernels(self.h, reps, _ptr(ms), buf, 256))
        names = buf.value.decode().split(",")
        return dict(zip(names, ms[:len(names)]))


class Server:
    """Many steady solves on one GPU (jm_server_*): admission by device
    footprint; batched=False: one stream and graph-captured steps per active
    solve; batched=True: the active solves step in lockstep, each kernel of
    the step and of the BiCGStab launched once for all of them. A mixed
    solve that goes non-finite is run again once in FP64 unless
    fp64_retry=False (the report's fp64_retries says so)."""

    QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, mem_budget_bytes: int, max_active: int = 8, batched: bool = False, fp64_retry: bool = True):
        self.h = C.c_void_p()
        flags = (1 if batched else 0) | (0 if fp64_retry else 2)
        _check(lib().jm_server_create(int(mem_budget_bytes), int(max_active), flags, C.byref(self.h)))
        self._meshes = {}

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.jm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def submit(self, mesh: Mesh, joe_in: str, precision: int = JM_FP64, state=None) -> int:
        t = C.c_int()
        if state is None:
            _check(lib().jm_server_submit(self.h, mesh.h, joe_in.encode(), precision, None, None, None, C.byref(t)))
        else:
            rho, rhou, rhoE = (np.ascontiguousarray(a, np.float64) for a in state)
            _check(lib().jm_server_submit(self.h, mesh.h, joe_in.encode(), precision, _ptr(rho), _ptr(rhou),
                                          _ptr(rhoE), C.byref(t)))
        self._meshes[t.value] = mesh  # keep the mesh alive until release
        return t.value

    def status(self, ticket: int) -> int:
        st = C.c_int()
        _check(lib().jm_server_status(self.h, ticket, C.byref(st)))
        return st.value

    def progress(self, ticket: int) -> int:
        """Pseudo-time steps the solve has taken so far."""
        n = C.c_int()
        _check(lib().jm_server_progress(self.h, ticket, C.byref(n)))
        return n.value

    def wait(self, ticket: int):
        _check(lib().jm_server_wait(self.h, ticket))

    def result(self, ticket: int):
        ncv = self._meshes[ticket].ncv
        rho, rhou, rhoE = np.empty(ncv), np.empty((ncv, 3)), np.empty(ncv)
        rep = SolveReport()
        _check(lib().jm_server_result(self.h, ticket, _ptr(rho), _ptr(rhou), _ptr(rhoE), C.byref(rep)))
        return (rho, rhou, rhoE), rep

    def release(self, ticket: int):
        _check(lib().jm_server_release(self.h, ticket))
        self._meshes.pop(ticket, None)


class Chemtable:
    """An FPVA chemistry table on the device (jm_chemtable; the reference's
    ChemtableCartesianLinear_single). lookup() takes host arrays and returns
    host arrays; lookup_device() works on DeviceArrays."""

    COEFF = tuple(range(10))     # LookupCoeff: RoM, T0, E0, Gam0, a_Gam, mu0, a_mu, lamOcp0, a_lamOcp, src_prog
    SELECTED = tuple(range(5))   # LookupSelectedCoeff: RoM, T0, E0, Gam0, a_Gam

    def __init__(self, handle):
        self.h = handle
        dims = np.zeros(4, np.int32)
        ref = np.zeros(3)
        _check(lib().jm_chemtable_info(self.h, _ptr(dims), _ptr(ref)))
        self.n1, self.n2, self.n3, self.nvar = (int(x) for x in dims)
        self.pref, self.toxi, self.tfuel = (float(x) for x in ref)

    @classmethod
    def load(cls, path):
        h = C.c_void_p()
        _check(lib().jm_chemtable_load(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def create(cls, x1, x2, x3, data, names=None):
        x1, x2, x3 = (np.ascontiguousarray(x, np.float64) for x in (x1, x2, x3))
        data = np.ascontiguousarray(data, np.float32)
        n1, n2, n3, nvar = data.shape
        nm = None
        if names is not None:
            keep = [s.encode() for s in names]
            nm = (C.c_char_p * len(keep))(*keep)
        h = C.c_void_p()
        _check(lib().jm_chemtable_create(n1, n2, n3, nvar, _ptr(x1), _ptr(x2), _ptr(x3), _ptr(data), nm,
                                         C.byref(h)))
        return cls(h)

    def var(self, name) -> int:
        i = C.c_int()
        _check(lib().jm_chemtable_var(self.h, name.encode(), C.byref(i)))
        return i.value

    def _vars(self, vars):
        v = np.ascontiguousarray([self.var(x) if isinstance(x, str) else int(x) for x in vars], np.int32)
        return v

    def lookup_device(self, zm, zv, pc, vars, out, stream=None):
        """jm_chemtable_lookup on DeviceArrays zm, zv, pc [n] into out [len(vars)][n]."""
        v = self._vars(vars)
        n = zm.host.size
        _check(lib().jm_chemtable_lookup(self.h, n, zm.data_ptr(), zv.data_ptr(), pc.data_ptr(), len(v), _ptr(v),
                                         out.data_ptr(), stream))

    def lookup(self, zm, zv, pc, vars):
        """Variables vars (indices or names) at the points: [len(vars)][n] float64."""
        zm, zv, pc = (np.ascontiguousarray(x, np.float64) for x in (zm, zv, pc))
        n = zm.size
        if zv.size != n or pc.size != n:
            raise ValueError("zm, zv, pc must have one value per point")
        if n == 0:
            return np.zeros((len(vars), 0))
        d = [DeviceArray(x) for x in (zm, zv, pc)]
        out = DeviceArray(np.zeros((len(vars), n)))
        self.lookup_device(*d, vars, out)
        return out.numpy()

    def time_lookup(self, zm, zv, pc, vars, out, reps=20, stream=None) -> float:
        """Mean ms of one lookup over DeviceArrays (HIP events on the stream)."""
        v = self._vars(vars)
        ms = C.c_double()
        _check(lib().jm_chemtable_time_lookup(self.h, zm.host.size, zm.data_ptr(), zv.data_ptr(), pc.data_ptr(),
                                              len(v), _ptr(v), out.data_ptr(), stream, reps, C.byref(ms)))
        return ms.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.jm_chemtable_destroy(self.h)
            self.h = None

