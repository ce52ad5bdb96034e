"""ctypes binding of the C ABI declared in include/meso_hip.h.

``Case`` wraps the host front-end (case file -> descriptors) and ``Device``
wraps the gfx950 step context. Both raise ``MesoError`` on a non-zero status.
The shared library is built in-tree (``meso-lts_amd/lib/libmeso_hip.so``);
loading fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libmeso_hip.so")

MESO_OK = 0
MESO_E_INVALID, MESO_E_IO, MESO_E_DEVICE, MESO_E_NUMERIC = -1, -2, -3, -4
MODEL_CDUGKS_BGK, MODEL_CDUGKS_SHAKHOV, MODEL_DUGKS_BGK = 1, 2, 3
STORAGE_FP64, STORAGE_FP32, STORAGE_FP32X, STORAGE_FP32E = 0, 1, 2, 3
LAYOUT_VELOCITY_MAJOR, LAYOUT_CELL_MAJOR = 0, 1
BC = dict(fluid_interior=0, farfield_inlet=1, pressure_inlet=2, freestream_inlet=3, farfield_outlet=4,
          pressure_outlet=5, wall=6, slip_wall=7, symmetry=8)

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)


class MarkDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("T_kind", C.c_int32), ("direction", C.c_int32), ("reserved", C.c_int32),
                ("velocity", C.c_double * 3), ("temperature", C.c_double), ("density", C.c_double),
                ("pressure", C.c_double), ("alpha_u", C.c_double), ("alpha_T", C.c_double)]


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int32), ("n_cells", C.c_int32), ("n_faces", C.c_int32), ("n_marks", C.c_int32),
                ("n_nbr", C.c_int32), ("n_cell_faces", C.c_int32), ("total_volume", C.c_double),
                ("min_cell_size", C.c_double), ("cell_pos", _d), ("cell_vol", _d), ("face_cells", _i),
                ("face_pos", _d), ("face_area", _d), ("face_normal", _d), ("cell_face_ptr", _i),
                ("cell_face_ids", _i), ("nbr_ptr", _i), ("nbr_ids", _i), ("lsq_dr", _d), ("lsq_w", _d),
                ("lsq_C", _d), ("face_mark", _i), ("marks", C.POINTER(MarkDesc))]


class DvsDesc(C.Structure):
    _fields_ = [("n_vel", C.c_int32), ("has_mirror", C.c_int32), ("max_magnitude", C.c_double), ("xi", _d),
                ("w", _d), ("mirror", _i)]


class StepParams(C.Structure):
    _fields_ = [("model", C.c_int32), ("dim", C.c_int32), ("gradient_switch", C.c_int32), ("K", C.c_int32),
                ("dt", C.c_double), ("Pr", C.c_double), ("R", C.c_double), ("T0", C.c_double), ("mu0", C.c_double),
                ("omega", C.c_double), ("RT", C.c_double), ("tau", C.c_double), ("limiter_switch", C.c_int32),
                ("reserved", C.c_int32), ("venkata_k", C.c_double)]


class TileConfig(C.Structure):
    _fields_ = [("patches", C.c_int32), ("update_cells", C.c_int32), ("moment_by_cells", C.c_int32),
                ("moment_size", C.c_int32), ("vchunk", C.c_int32), ("update_vt", C.c_int32),
                ("moment_vt", C.c_int32), ("update_owned", C.c_int32),
                ("moment_lds", C.c_int32), ("update_lds", C.c_int32)]


class StepReport(C.Structure):
    _fields_ = [("steps_done", C.c_int32), ("bad_cells", C.c_int32), ("first_bad_cell", C.c_int32),
                ("masked_cells", C.c_int32)]


class MesoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"meso status {code}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MesoError(MESO_E_INVALID, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "meso_case_open": ([C.c_char_p, C.c_double, C.POINTER(vp)], C.c_int),
            "meso_case_close": ([vp], None),
            "meso_case_last_error": ([], C.c_char_p),
            "meso_case_mesh": ([vp, C.POINTER(MeshDesc)], C.c_int),
            "meso_case_dvs": ([vp, C.POINTER(DvsDesc)], C.c_int),
            "meso_case_params": ([vp, C.POINTER(StepParams)], C.c_int),
            "meso_case_reload_settings": ([vp], C.c_int),
            "meso_case_initial_patch": ([vp, _d, _d, _d], C.c_int),
            "meso_case_cell_mask": ([vp, _i], C.c_int),
            "meso_case_setting_int": ([vp, C.c_char_p, C.c_int32, _i], C.c_int),
            "meso_case_setting_double": ([vp, C.c_char_p, C.c_double, _d], C.c_int),
            "meso_case_setting_str": ([vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32], C.c_int),
            "meso_case_write_output": ([vp, C.c_char_p, C.c_int32, C.c_double, _d, _d, _d, _d, _d, _d, _d],
                                       C.c_int),
            "meso_case_output_error": ([], C.c_char_p),
            "meso_mesh_renumber": ([C.c_char_p, C.c_char_p, C.c_int32, C.c_uint64, _i, C.c_int32, _i], C.c_int),
            "meso_mesh_renumber_error": ([], C.c_char_p),
            "meso_hip_create": ([C.POINTER(MeshDesc), C.POINTER(DvsDesc), C.POINTER(StepParams), C.c_int32,
                                 C.POINTER(vp)], C.c_int),
            "meso_hip_destroy": ([vp], None),
            "meso_hip_last_error": ([vp], C.c_char_p),
            "meso_hip_init_equilibrium": ([vp, _d, _d, _d], C.c_int),
            "meso_hip_init_from_macros": ([vp, _d, _d, _d, _d], C.c_int),
            "meso_hip_init_equilibrium_noalloc": ([vp, _d, _d, _d], C.c_int),
            "meso_hip_set_state": ([vp, _d, C.c_int32, _d, _d, _d, _d], C.c_int),
            "meso_hip_get_state": ([vp, _d, C.c_int32], C.c_int),
            "meso_hip_step": ([vp, C.c_int32, C.POINTER(StepReport)], C.c_int),
            "meso_hip_step_batch": ([C.POINTER(vp), C.c_int32, C.c_int32, C.POINTER(StepReport)], C.c_int),
            "meso_hip_batch_create": ([C.POINTER(vp), C.c_int32, C.POINTER(vp)], C.c_int),
            "meso_hip_batch_step": ([vp, C.c_int32, C.POINTER(StepReport)], C.c_int),
            "meso_hip_batch_set_graphs": ([vp, C.c_int32], C.c_int),
            "meso_hip_batch_set_rechunk": ([vp, C.c_int32], C.c_int),
            "meso_hip_batch_set_streams": ([vp, C.c_int32], C.c_int),
            "meso_hip_set_vchunk": ([vp, C.c_int32], C.c_int),
            "meso_hip_batch_info": ([vp, _i, _i, _i], C.c_int),
            "meso_hip_batch_destroy": ([vp], None),
            "meso_hip_autotune": ([vp, C.c_int32, _i], C.c_int),
            "meso_hip_set_graphs": ([vp, C.c_int32], C.c_int),
            "meso_hip_tiles": ([vp, C.POINTER(TileConfig)], C.c_int),
            "meso_hip_set_params": ([vp, C.POINTER(StepParams)], C.c_int),
            "meso_hip_set_mask": ([vp, _i], C.c_int),
            "meso_hip_get_macros": ([vp, _d, _d, _d, _d], C.c_int),
            "meso_hip_get_flux_moments": ([vp, _d, _d, _d], C.c_int),
            "meso_hip_residual": ([vp, _d], C.c_int),
            "meso_hip_profile": ([vp, C.c_int32], C.c_int),
            "meso_hip_kernel_times": ([vp, C.c_int32, C.POINTER(C.c_char_p), _d, _i, _i], C.c_int),
            "meso_hip_device_bytes": ([vp], C.c_int64),
            "meso_hip_sync": ([vp], C.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def dptr(a):
    return None if a is None else a.ctypes.data_as(_d)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_i)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Case:
    """A meso case (Config + gambit mesh + DVS) opened by the C++ front-end."""

    def __init__(self, path: str, mesh_scale: float = 0.0):
        L = lib()
        h = C.c_void_p()
        rc = L.meso_case_open(path.encode(), mesh_scale, C.byref(h))
        if rc != MESO_OK:
            raise MesoError(rc, L.meso_case_last_error().decode())
        self._h = h
        self.path = path
        self.mesh = MeshDesc()
        self.dvs = DvsDesc()
        self.params = StepParams()
        self._check(L.meso_case_mesh(h, C.byref(self.mesh)))
        self._check(L.meso_case_dvs(h, C.byref(self.dvs)))
        self._check(L.meso_case_params(h, C.byref(self.params)))

    def _check(self, rc):
        if rc != MESO_OK:
            raise MesoError(rc, lib().meso_case_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().meso_case_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_cells(self):
        return self.mesh.n_cells

    @property
    def n_faces(self):
        return self.mesh.n_faces

    @property
    def n_vel(self):
        return self.dvs.n_vel

    @property
    def model(self):
        return self.params.model

    @property
    def n_dist(self):
        return 2 if self.params.model == MODEL_CDUGKS_SHAKHOV else 1

    def reload_settings(self):
        self._check(lib().meso_case_reload_settings(self._h))
        self._check(lib().meso_case_params(self._h, C.byref(self.params)))

    def initial_patch(self):
        nc = self.n_cells
        rho, u, T = np.zeros(nc), np.zeros(nc * 3), np.zeros(nc)
        self._check(lib().meso_case_initial_patch(self._h, dptr(rho), dptr(u), dptr(T)))
        return rho, u.reshape(nc, 3), T

    def cell_mask(self):
        """1 for the cells of groups with `mask switch True` (meso_case_cell_mask)."""
        m = np.zeros(self.n_cells, dtype=np.int32)
        self._check(lib().meso_case_cell_mask(self._h, iptr(m)))
        return m

    def write_output(self, root, step, solution_time, rho, u, T=None, q=None, flux=None):
        """The files meso's output() writes for `step` (meso_case_write_output):
        rho (NC), u (NC x 3); Shakhov also T, q (NC x 3) and flux = (m0, m1, m2)."""
        def arr(x, n):
            return None if x is None else np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(n))
        nc = self.n_cells
        m0, m1, m2 = flux if flux is not None else (None, None, None)
        bufs = [arr(rho, nc), arr(u, 3 * nc), arr(T, nc), arr(q, 3 * nc), arr(m0, nc), arr(m1, 3 * nc), arr(m2, nc)]
        ptrs = [None if b is None else dptr(b) for b in bufs]
        rc = lib().meso_case_write_output(self._h, str(root).encode(), step, solution_time, *ptrs)
        if rc != MESO_OK:
            raise MesoError(rc, lib().meso_case_output_error().decode())

    def setting_int(self, key, dflt):
        out = C.c_int32()
        self._check(lib().meso_case_setting_int(self._h, key.encode(), dflt, C.byref(out)))
        return out.value

    def setting_double(self, key, dflt):
        out = C.c_double()
        self._check(lib().meso_case_setting_double(self._h, key.encode(), dflt, C.byref(out)))
        return out.value

    def setting_str(self, key, dflt=""):
        buf = C.create_string_buffer(4096)
        self._check(lib().meso_case_setting_str(self._h, key.encode(), dflt.encode(), buf, 4096))
        return buf.value.decode()

    # numpy views of the descriptor arrays (valid while the case is open)
    def arrays(self):
        m, d = self.mesh, self.dvs
        nc, nf = m.n_cells, m.n_faces
        a = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy()
        out = dict(
            cell_pos=a(m.cell_pos, 3 * nc).reshape(nc, 3), cell_vol=a(m.cell_vol, nc),
            face_cells=a(m.face_cells, 2 * nf).reshape(nf, 2), face_pos=a(m.face_pos, 3 * nf).reshape(nf, 3),
            face_area=a(m.face_area, nf), face_normal=a(m.face_normal, 3 * nf).reshape(nf, 3),
            cell_face_ptr=a(m.cell_face_ptr, nc + 1), cell_face_ids=a(m.cell_face_ids, m.n_cell_faces),
            nbr_ptr=a(m.nbr_ptr, nc + 1), nbr_ids=a(m.nbr_ids, m.n_nbr), lsq_dr=a(m.lsq_dr, 3 * m.n_nbr).reshape(-1, 3),
            lsq_w=a(m.lsq_w, m.n_nbr), lsq_C=a(m.lsq_C, 9 * nc).reshape(nc, 9), face_mark=a(m.face_mark, nf),
            xi=a(d.xi, 3 * d.n_vel).reshape(-1, 3), w=a(d.w, d.n_vel),
            total_volume=m.total_volume, min_cell_size=m.min_cell_size,
        )
        out["marks"] = [m.marks[i] for i in range(m.n_marks)]
        if d.has_mirror:
            out["mirror"] = a(d.mirror, 3 * d.n_vel).reshape(-1, 3)
        return out


class Device:
    """The gfx950 step context (one HIP stream, owns device memory)."""

    def __init__(self, case: Case, storage: int = STORAGE_FP64, params: StepParams | None = None):
        L = lib()
        h = C.c_void_p()
        self.case = case
        self.storage = storage
        rc = L.meso_hip_create(C.byref(case.mesh), C.byref(case.dvs), C.byref(params or case.params), storage,
                               C.byref(h))
        if rc != MESO_OK:
            raise MesoError(rc, L.meso_hip_last_error(None).decode())
        self._h = h
        self.nc, self.nv, self.nd = case.n_cells, case.n_vel, case.n_dist

    def _check(self, rc):
        if rc != MESO_OK:
            raise MesoError(rc, lib().meso_hip_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().meso_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_equilibrium(self, rho=None, u=None, T=None):
        if rho is None:
            rho, u, T = self.case.initial_patch()
        self._check(lib().meso_hip_init_equilibrium(self._h, dptr(as_f64(rho)), dptr(as_f64(u)), dptr(as_f64(T))))

    def init_equilibrium_noalloc(self, rho=None, u=None, T=None):
        """init_equilibrium through meso_hip_init_equilibrium_noalloc (no device allocation)."""
        if rho is None:
            rho, u, T = self.case.initial_patch()
        self._check(lib().meso_hip_init_equilibrium_noalloc(self._h, dptr(as_f64(rho)), dptr(as_f64(u)),
                                                            dptr(as_f64(T))))

    def init_from_macros(self, rho, u, T, q):
        self._check(lib().meso_hip_init_from_macros(self._h, dptr(as_f64(rho)), dptr(as_f64(u)), dptr(as_f64(T)),
                                                    dptr(as_f64(q))))

    def set_mask(self, mask=None):
        """Cells repaired instead of stopping the run (meso_hip_set_mask); None
        takes the case's `mask switch True` groups."""
        if mask is None:
            mask = self.case.cell_mask()
        m = np.ascontiguousarray(mask, dtype=np.int32)
        self._check(lib().meso_hip_set_mask(self._h, iptr(m)))

    def set_state(self, f, rho, u, T=None, q=None, layout=LAYOUT_VELOCITY_MAJOR):
        self._check(lib().meso_hip_set_state(self._h, dptr(as_f64(f)), layout, dptr(as_f64(rho)), dptr(as_f64(u)),
                                             None if T is None else dptr(as_f64(T)),
                                             None if q is None else dptr(as_f64(q))))

    def get_state(self, layout=LAYOUT_VELOCITY_MAJOR):
        f = np.zeros(self.nd * self.nv * self.nc)
        self._check(lib().meso_hip_get_state(self._h, dptr(f), layout))
        if layout == LAYOUT_VELOCITY_MAJOR:
            return f.reshape(self.nd, self.nv, self.nc)
        return f.reshape(self.nd, self.nc, self.nv)

    def step(self, n=1):
        rep = StepReport()
        rc = lib().meso_hip_step(self._h, n, C.byref(rep))
        if rc not in (MESO_OK, MESO_E_NUMERIC):
            self._check(rc)
        return rep

    def autotune(self, trial_steps=3):
        """Pick the velocity chunk of the patch kernels by timing trial steps
        (meso_hip_autotune); returns velocities per block (0: per-cell kernels)."""
        out = C.c_int32()
        self._check(lib().meso_hip_autotune(self._h, trial_steps, C.byref(out)))
        return out.value

    def tiles(self):
        """The tile shapes in use (meso_hip_tiles) as a dict."""
        t = TileConfig()
        self._check(lib().meso_hip_tiles(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in TileConfig._fields_ if k != "reserved"}

    def set_graphs(self, enable: bool):
        """Replay captured step pairs (meso_hip_set_graphs)."""
        self._check(lib().meso_hip_set_graphs(self._h, 1 if enable else 0))

    def set_vchunk(self, vchunk: int):
        """Velocities per block of the patch kernels (meso_hip_set_vchunk; 0: default)."""
        self._check(lib().meso_hip_set_vchunk(self._h, int(vchunk)))

    def set_params(self, p: StepParams):
        self._check(lib().meso_hip_set_params(self._h, C.byref(p)))

    def macros(self):
        nc = self.nc
        rho, u, T, q = np.zeros(nc), np.zeros(3 * nc), np.zeros(nc), np.zeros(3 * nc)
        self._check(lib().meso_hip_get_macros(self._h, dptr(rho), dptr(u), dptr(T), dptr(q)))
        return dict(rho=rho, u=u.reshape(nc, 3), T=T, q=q.reshape(nc, 3))

    def flux_moments(self):
        nc = self.nc
        m0, m1, m2 = np.zeros(nc), np.zeros(3 * nc), np.zeros(nc)
        self._check(lib().meso_hip_get_flux_moments(self._h, dptr(m0), dptr(m1), dptr(m2)))
        return m0, m1.reshape(nc, 3), m2

    def residual(self):
        out = np.zeros(8)
        self._check(lib().meso_hip_residual(self._h, dptr(out)))
        return out

    def profile(self, enable: bool):
        self._check(lib().meso_hip_profile(self._h, 1 if enable else 0))

    def kernel_times(self):
        n = 64
        names = (C.c_char_p * n)()
        tot = np.zeros(n)
        cnt = np.zeros(n, dtype=np.int32)
        got = C.c_int32()
        self._check(lib().meso_hip_kernel_times(self._h, n, names, dptr(tot), iptr(cnt), C.byref(got)))
        return {names[i].decode(): (float(tot[i]), int(cnt[i])) for i in range(got.value)}

    def device_bytes(self):
        return int(lib().meso_hip_device_bytes(self._h))

    def sync(self):
        self._check(lib().meso_hip_sync(self._h))


RENUMBER_RCB, RENUMBER_RANDOM, RENUMBER_KEEP = 0, 1, 2


def renumber_mesh(in_path: str, out_path: str, method: int = RENUMBER_RCB, seed: int = 0) -> np.ndarray:
    """Write the gambit file in_path to out_path with its cells renumbered
    (meso_mesh_renumber: RENUMBER_RCB locality order, RENUMBER_RANDOM seeded
    permutation, RENUMBER_KEEP file order); returns each old cell's new id."""
    L = lib()
    nelem = 0
    try:
        with open(in_path) as fp:  # NELEM from the header line after NUMNP
            for line in fp:
                if line.split()[:1] == ["NUMNP"]:
                    nelem = int(next(fp).split()[1])
                    break
    except (OSError, ValueError, IndexError, StopIteration):
        nelem = 0  # the library reports what is wrong with the file
    perm = np.zeros(nelem, dtype=np.int32)
    n = C.c_int32(0)
    rc = L.meso_mesh_renumber(in_path.encode(), out_path.encode(), method, seed, iptr(perm), nelem, C.byref(n))
    if rc != MESO_OK:
        raise MesoError(rc, L.meso_mesh_renumber_error().decode())
    return perm[:n.value]


def step_batch(devices, n=1):
    """Advance every Device of the list by n steps at once (meso_hip_step_batch);
    returns the per-device StepReports (a numerical failure shows in its report,
    as with Device.step). Raises on any other error."""
    k = len(devices)
    arr = (C.c_void_p * k)(*[d._h for d in devices])
    reps = (StepReport * k)()
    rc = lib().meso_hip_step_batch(arr, k, n, reps)
    if rc not in (MESO_OK, MESO_E_NUMERIC):
        raise MesoError(rc, lib().meso_hip_last_error(None).decode())
    return list(reps)


class Batch:
    """A fixed set of Devices advanced with one launch per stage
    (meso_hip_batch_*): the devices on the patch kernels share each stage's
    launch; the others keep their own launches in the same call."""

    def __init__(self, devices):
        self.devices = list(devices)
        k = len(self.devices)
        arr = (C.c_void_p * k)(*[d._h for d in self.devices])
        h = C.c_void_p()
        rc = lib().meso_hip_batch_create(arr, k, C.byref(h))
        if rc != MESO_OK:
            raise MesoError(rc, lib().meso_hip_last_error(None).decode())
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().meso_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        """Advance every device by n steps; returns the per-device StepReports."""
        k = len(self.devices)
        reps = (StepReport * k)()
        rc = lib().meso_hip_batch_step(self._h, n, reps)
        if rc not in (MESO_OK, MESO_E_NUMERIC):
            raise MesoError(rc, lib().meso_hip_last_error(None).decode())
        return list(reps)

    def set_graphs(self, enable: bool):
        rc = lib().meso_hip_batch_set_graphs(self._h, 1 if enable else 0)
        if rc != MESO_OK:
            raise MesoError(rc, lib().meso_hip_last_error(None).decode())

    def set_streams(self, n: int):
        """Sub-batches on their own streams (meso_hip_batch_set_streams, default 4)."""
        rc = lib().meso_hip_batch_set_streams(self._h, int(n))
        if rc != MESO_OK:
            raise MesoError(rc, lib().meso_hip_last_error(None).decode())

    def set_rechunk(self, enable: bool):
        """Let the plan set the fused devices' velocity chunks for the batch grid (default on)."""
        rc = lib().meso_hip_batch_set_rechunk(self._h, 1 if enable else 0)
        if rc != MESO_OK:
            raise MesoError(rc, lib().meso_hip_last_error(None).decode())

    def info(self):
        """(fused devices, kernel-instance groups, launches per fused step); builds the plan"""
        f, g, l = C.c_int32(), C.c_int32(), C.c_int32()
        rc = lib().meso_hip_batch_info(self._h, C.byref(f), C.byref(g), C.byref(l))
        if rc != MESO_OK:
            raise MesoError(rc, lib().meso_hip_last_error(None).decode())
        return f.value, g.value, l.value
