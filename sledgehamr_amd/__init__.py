"""MI355X (gfx950) implementation of sledgehamr's per-level time-step path.

Python binding over the C ABI in ``include/shm.h``.  Each project lives in
its own ``lib/libshm_<Project>.so`` built by ``sledgehamr_amd.build``; there is
no CPU fallback: a missing library or a device error raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = ["Simulation", "library_path", "hdf5_version", "PROJECTS", "ShmError",
           "PREC_FP64", "PREC_FP32C", "PREC_FP32C16", "PREC_FP32CS", "TE_COMPUTE",
           "TE_AVGDOWN", "TE_TAGS"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.environ.get("SHM_LIBDIR") or os.path.join(_HERE, "lib")
PROJECTS = ("RealScalar", "ComplexScalar", "PhaseTransitionGW")

PREC_INPUTS, PREC_FP64, PREC_FP32C, PREC_FP32C16, PREC_FP32CS = -1, 0, 1, 2, 3
TE_COMPUTE, TE_AVGDOWN, TE_TAGS = 1, 2, 4

# Exported symbols declared in include/shm.h (checked by the CPU tests).
EXPORTS = (
    "shm_create", "shm_destroy", "shm_last_error", "shm_project_name",
    "shm_get_info", "shm_set_box_table", "shm_get_box_table",
    "shm_level_info", "shm_set_level_time", "shm_upload", "shm_download",
    "shm_init_state", "shm_fill_rhs", "shm_download_rhs", "shm_advance_level",
    "shm_evolve", "shm_level_tag_counts", "shm_create_shadow", "shm_te_tag", "shm_gw_spectrum",
    "shm_gw_spectrum_mod", "shm_gw_modifiers", "shm_gw_modifier_name",
    "shm_te_tag_noalloc", "shm_event_log",
    "shm_output_functions", "shm_output_function_name", "shm_spectrum", "shm_projection",
    "shm_upload_array", "shm_fill_coarse_patch",
    "shm_regrid_select", "shm_add_boxes", "shm_remake_level", "shm_local_regrid",
    "shm_regrid_stats",
    "shm_sync", "shm_stream_probe", "shm_autotune", "shm_set_launch",
    "shm_get_launch", "shm_set_graphs", "shm_set_timing", "shm_get_timing",
    "shm_hdf5_version", "shm_read_initial_state", "shm_write_checkpoint",
    "shm_read_checkpoint", "shm_delete_checkpoint", "shm_write_gw_spectrum",
    "shm_write_spectra", "shm_write_projections",
)


def hdf5_version(project: str = "RealScalar"):
    """(major, minor, release) of the libhdf5 the library loads for its
    file entry points; raises ShmError when none can be loaded."""
    lib = load_library(project)
    v = (C.c_int * 3)()
    rc = lib.shm_hdf5_version(v)
    if rc:
        raise ShmError(rc, "libhdf5 could not be loaded (set SHM_HDF5_LIB)")
    return tuple(v)


class ShmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[shm status {code}] {msg}")
        self.code = code


class _Config(C.Structure):
    _fields_ = [("inputs", C.c_char_p), ("precision", C.c_int),
                ("device", C.c_int)]


class _Launch(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "variant", "tile_z", "xcd_split", "persistent", "zc_minor", "fused",
        "fused_variant")]


class _Info(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "ncomp", "nscalars", "nfields", "nghost", "stencil_radius",
        "precision", "max_level", "finest_level", "coarse_level_grid_size",
        "integrator_type", "with_gravitational_waves", "shadow_hierarchy")] + \
        [(n, C.c_double) for n in ("L", "cfl", "t_start", "t_end")] + \
        [("graph_replays", C.c_longlong), ("gw_groups", C.c_int)]


_LIBS: dict[str, C.CDLL] = {}


def library_path(project: str) -> str:
    return os.path.join(LIBDIR, f"libshm_{project}.so")


def load_library(project: str) -> C.CDLL:
    """Loads ``libshm_<project>.so``; raises if it was not built."""
    if project in _LIBS:
        return _LIBS[project]
    path = library_path(project)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} not found: build it with `python -m sledgehamr_amd.build`")
    lib = C.CDLL(path, mode=C.RTLD_LOCAL)
    vp, i, d, ll = C.c_void_p, C.c_int, C.c_double, C.c_longlong
    dp, ip, llp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_longlong)
    sig = {
        "shm_create": (i, [C.POINTER(_Config), C.POINTER(vp)]),
        "shm_destroy": (i, [vp]),
        "shm_last_error": (C.c_char_p, [vp]),
        "shm_project_name": (C.c_char_p, [vp]),
        "shm_get_info": (i, [vp, C.POINTER(_Info)]),
        "shm_set_box_table": (i, [vp, i, ip, i]),
        "shm_get_box_table": (i, [vp, i, ip, i, ip]),
        "shm_level_info": (i, [vp, i, llp, dp, ip]),
        "shm_set_level_time": (i, [vp, i, d, i]),
        "shm_upload": (i, [vp, i, i, dp]),
        "shm_download": (i, [vp, i, i, dp, i]),
        "shm_init_state": (i, [vp, i, i, C.c_ulonglong, dp, i]),
        "shm_fill_rhs": (i, [vp, i, d, d, i]),
        "shm_download_rhs": (i, [vp, i, i, dp]),
        "shm_advance_level": (i, [vp, i]),
        "shm_evolve": (i, [vp, i]),
        "shm_create_shadow": (i, [vp]),
        "shm_level_tag_counts": (i, [vp, i, llp]),
        "shm_te_tag": (i, [vp, i, i, C.POINTER(C.c_ubyte), llp, dp]),
        "shm_te_tag_noalloc": (i, [vp, i, i, C.POINTER(C.c_ubyte), llp]),
        "shm_event_log": (i, [vp, C.c_char_p, ip, dp, i, ip]),
        "shm_gw_spectrum": (i, [vp, i, i, i, i, dp, dp, ip]),
        "shm_gw_spectrum_mod": (i, [vp, i, i, i, i, i, dp, i, dp, dp, ip]),
        "shm_gw_modifiers": (i, [vp, ip]),
        "shm_gw_modifier_name": (i, [vp, i, C.c_char_p, i]),
        "shm_upload_array": (i, [vp, i, i, dp, i]),
        "shm_fill_coarse_patch": (i, [vp, i]),
        "shm_regrid_select": (i, [vp, i, C.POINTER(C.c_ubyte), ip, i, ip, dp]),
        "shm_add_boxes": (i, [vp, i, ip, i]),
        "shm_remake_level": (i, [vp, i, ip, i]),
        "shm_local_regrid": (i, [vp, i, ip]),
        "shm_regrid_stats": (i, [vp, llp, dp, i]),
        "shm_output_functions": (i, [vp, ip]),
        "shm_output_function_name": (i, [vp, i, C.c_char_p, i]),
        "shm_spectrum": (i, [vp, i, dp, llp, ip]),
        "shm_projection": (i, [vp, i, i, i, dp, ip, ip]),
        "shm_sync": (i, [vp]),
        "shm_stream_probe": (i, [vp, C.c_ulonglong, i, dp, dp, dp]),
        "shm_autotune": (i, [vp, i]),
        "shm_set_launch": (i, [vp, C.POINTER(_Launch)]),
        "shm_get_launch": (i, [vp, i, C.POINTER(_Launch)]),
        "shm_set_graphs": (i, [vp, i]),
        "shm_set_timing": (i, [vp, i]),
        "shm_get_timing": (i, [vp, dp, llp, llp, dp, llp, dp, llp]),
        "shm_hdf5_version": (i, [ip]),
        "shm_read_initial_state": (i, [vp, i, C.c_char_p]),
        "shm_write_checkpoint": (i, [vp, C.c_char_p, ip, dp, i, i]),
        "shm_read_checkpoint": (i, [vp, C.c_char_p, ip, dp, i, i]),
        "shm_delete_checkpoint": (i, [vp, C.c_char_p]),
        "shm_write_gw_spectrum": (i, [vp, C.c_char_p, i, dp, i]),
        "shm_write_spectra": (i, [vp, C.c_char_p, ip, i]),
        "shm_write_projections": (i, [vp, C.c_char_p, ip, ip, i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIBS[project] = lib
    return lib


def _inputs_text(inputs) -> str:
    if isinstance(inputs, str):
        return inputs
    lines = []
    for k, v in inputs.items():
        if isinstance(v, (list, tuple)):
            v = " ".join(str(x) for x in v)
        elif isinstance(v, bool):
            v = "true" if v else "false"
        lines.append(f"{k} = {v}")
    return "\n".join(lines) + "\n"


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Simulation:
    """One project instance on one GPU (a ``shm_ctx``)."""

    def __init__(self, project: str, inputs, precision: int = PREC_INPUTS,
                 device: int = 0):
        self.project = project
        self.lib = load_library(project)
        self._ctx = C.c_void_p()
        self._inputs = _inputs_text(inputs).encode()
        cfg = _Config(self._inputs, precision, device)
        rc = self.lib.shm_create(C.byref(cfg), C.byref(self._ctx))
        if rc != 0:
            msg = self.lib.shm_last_error(self._ctx).decode() if self._ctx else ""
            if self._ctx:
                self.lib.shm_destroy(self._ctx)
            self._ctx = C.c_void_p()
            raise ShmError(rc, msg)
        self.info = self._get_info()

    # -- plumbing ----------------------------------------------------------
    def _check(self, rc: int) -> None:
        if rc != 0:
            raise ShmError(rc, self.lib.shm_last_error(self._ctx).decode())

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self.lib.shm_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _get_info(self) -> dict:
        inf = _Info()
        self._check(self.lib.shm_get_info(self._ctx, C.byref(inf)))
        return {n: getattr(inf, n) for n, _ in _Info._fields_}

    # -- levels ------------------------------------------------------------
    def set_boxes(self, lev: int, boxes) -> None:
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        self._check(self.lib.shm_set_box_table(
            self._ctx, lev, b.ctypes.data_as(C.POINTER(C.c_int)), len(b)))
        self.info = self._get_info()

    def boxes(self, lev: int) -> np.ndarray:
        n = C.c_int()
        self._check(self.lib.shm_get_box_table(self._ctx, lev, None, 0, C.byref(n)))
        b = np.zeros((n.value, 6), dtype=np.int32)
        self._check(self.lib.shm_get_box_table(
            self._ctx, lev, b.ctypes.data_as(C.POINTER(C.c_int)), n.value,
            C.byref(n)))
        return b

    def level_info(self, lev: int) -> tuple[int, float, int]:
        n, t, s = C.c_longlong(), C.c_double(), C.c_int()
        self._check(self.lib.shm_level_info(self._ctx, lev, C.byref(n),
                                            C.byref(t), C.byref(s)))
        return n.value, t.value, s.value

    def set_time(self, lev: int, t: float, istep: int = 0) -> None:
        self._check(self.lib.shm_set_level_time(self._ctx, lev, t, istep))

    def ncells(self, lev: int) -> int:
        return self.level_info(lev)[0]

    # -- data --------------------------------------------------------------
    def upload(self, lev: int, comp: int, data) -> None:
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
        if a.size != self.ncells(lev):
            raise ValueError("data size does not match the level's valid cells")
        self._check(self.lib.shm_upload(self._ctx, lev, comp, _dptr(a)))

    def download(self, lev: int, comp: int, old: bool = False) -> np.ndarray:
        a = np.empty(self.ncells(lev), dtype=np.float64)
        self._check(self.lib.shm_download(self._ctx, lev, comp, _dptr(a), int(old)))
        return a

    def rhs(self, lev: int, comp: int) -> np.ndarray:
        a = np.empty(self.ncells(lev), dtype=np.float64)
        self._check(self.lib.shm_download_rhs(self._ctx, lev, comp, _dptr(a)))
        return a

    def init_state(self, lev: int, kind: int, seed: int, params=()) -> None:
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        self._check(self.lib.shm_init_state(
            self._ctx, lev, kind, seed, _dptr(p) if p.size else None, p.size))

    # -- operations --------------------------------------------------------
    def fill_rhs(self, lev: int, time: float, weight: float = 0.0,
                 add: bool = False) -> None:
        self._check(self.lib.shm_fill_rhs(self._ctx, lev, time, weight, int(add)))

    def advance(self, lev: int = 0, nsteps: int = 1) -> None:
        for _ in range(nsteps):
            self._check(self.lib.shm_advance_level(self._ctx, lev))

    def evolve(self, ncoarse: int) -> None:
        self._check(self.lib.shm_evolve(self._ctx, ncoarse))

    def level_tag_counts(self, lev: int) -> np.ndarray:
        """Counters [total, user, trunc...] of the last scheduled tagging."""
        c = np.zeros(2 + self.info["ncomp"], dtype=np.int64)
        self._check(self.lib.shm_level_tag_counts(
            self._ctx, lev, c.ctypes.data_as(C.POINTER(C.c_longlong))))
        return c

    def create_shadow(self) -> None:
        self._check(self.lib.shm_create_shadow(self._ctx))

    def te_tag(self, lev: int, flags: int = TE_COMPUTE | TE_AVGDOWN | TE_TAGS,
               want_te: bool = False):
        n = self.ncells(lev)
        mask = np.zeros(n, dtype=np.uint8)
        counts = np.zeros(2 + self.info["ncomp"], dtype=np.int64)
        te = None
        if want_te:
            te = np.zeros((self.info["ncomp"], n // 8), dtype=np.float64)
        self._check(self.lib.shm_te_tag(
            self._ctx, lev, flags, mask.ctypes.data_as(C.POINTER(C.c_ubyte)),
            counts.ctypes.data_as(C.POINTER(C.c_longlong)),
            _dptr(te) if te is not None else None))
        return mask, counts, te

    def events(self) -> list:
        """Level operations of the evolves so far: (kind, lev, t) with kind
        S, H, A, D, T or R (see shm_event_log)."""
        n = C.c_int(0)
        self._check(self.lib.shm_event_log(self._ctx, None, None, None, 0, C.byref(n)))
        m = max(n.value, 1)
        k = C.create_string_buffer(m)
        lv = np.zeros(m, dtype=np.int32)
        t = np.zeros(m)
        self._check(self.lib.shm_event_log(self._ctx, k, lv.ctypes.data_as(C.POINTER(C.c_int)),
                                           _dptr(t), n.value, C.byref(n)))
        return [(k.raw[i:i + 1].decode(), int(lv[i]), float(t[i])) for i in range(n.value)]

    def te_tag_noalloc(self, lev: int, flags: int = TE_COMPUTE | TE_TAGS):
        """shm_te_tag_noalloc: the tagging without device allocation (mask,
        counters)."""
        n = self.ncells(lev)
        mask = np.zeros(n, dtype=np.uint8)
        counts = np.zeros(2 + self.info["ncomp"], dtype=np.int64)
        self._check(self.lib.shm_te_tag_noalloc(
            self._ctx, lev, flags, mask.ctypes.data_as(C.POINTER(C.c_ubyte)),
            counts.ctypes.data_as(C.POINTER(C.c_longlong))))
        return mask, counts

    def gw_spectrum(self, mode: int = 0, nbins: int = 256, zero_padding: int = 1,
                    projection_type: int = 2, modifier: int = -1, params=()):
        """GW spectrum (bins, k values); modifier: id of a project spectrum
        modifier (-1: du_ij, unmodified) with its parameters."""
        n = C.c_int(0)
        if modifier < 0 and not len(params):
            self._check(self.lib.shm_gw_spectrum(self._ctx, mode, nbins, zero_padding,
                                                 projection_type, None, None,
                                                 C.byref(n)))
            bins = np.zeros(n.value)
            k = np.zeros(n.value)
            self._check(self.lib.shm_gw_spectrum(self._ctx, mode, nbins, zero_padding,
                                                 projection_type, _dptr(bins),
                                                 _dptr(k), C.byref(n)))
            return bins, k
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        pp = _dptr(p) if p.size else None
        self._check(self.lib.shm_gw_spectrum_mod(self._ctx, mode, nbins, zero_padding,
                                                 projection_type, modifier, pp, p.size,
                                                 None, None, C.byref(n)))
        bins = np.zeros(n.value)
        k = np.zeros(n.value)
        self._check(self.lib.shm_gw_spectrum_mod(self._ctx, mode, nbins, zero_padding,
                                                 projection_type, modifier, pp, p.size,
                                                 _dptr(bins), _dptr(k), C.byref(n)))
        return bins, k

    def gw_modifiers(self) -> list:
        """Names of the project's GW spectrum modifiers, by id."""
        n = C.c_int(0)
        self._check(self.lib.shm_gw_modifiers(self._ctx, C.byref(n)))
        out = []
        for i in range(n.value):
            buf = C.create_string_buffer(256)
            self._check(self.lib.shm_gw_modifier_name(self._ctx, i, buf, 256))
            out.append(buf.value.decode())
        return out

    def upload_array(self, lev: int, comp: int, data, layout: int) -> None:
        """Upload from the reference's array orders (1: whole level,
        i N^2 + j N + k; 2: per-box chunks, box-relative i ly lz + j lz + k)."""
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
        self._check(self.lib.shm_upload_array(self._ctx, lev, comp, _dptr(a), layout))

    def fill_coarse_patch(self, lev: int) -> None:
        """FillCoarsePatch: level lev from level lev-1 at the level's time."""
        self._check(self.lib.shm_fill_coarse_patch(self._ctx, lev))

    def regrid_select(self, lev: int, mask=None):
        """DetermineNewBoxArray of tagging level lev: (blocks of level lev+1
        as an (n, 3) array in (i, j, k) order, shortest distance or -1).
        mask: one byte per valid cell, or None for the last tagging's."""
        mp = None
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).ravel())
            if m.size != self.ncells(lev):
                raise ValueError("mask needs one byte per valid cell")
            mp = m.ctypes.data_as(C.POINTER(C.c_ubyte))
        n, dist = C.c_int(), C.c_double()
        self._check(self.lib.shm_regrid_select(self._ctx, lev, mp, None, 0, C.byref(n),
                                               C.byref(dist)))
        b = np.zeros((n.value, 3), dtype=np.int32)
        self._check(self.lib.shm_regrid_select(
            self._ctx, lev, mp, b.ctypes.data_as(C.POINTER(C.c_int)), n.value,
            C.byref(n), C.byref(dist)))
        return b, dist.value

    def add_boxes(self, lev: int, boxes) -> None:
        """AddBoxes: the level gains `boxes` (lo, hi rows) after its own."""
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        self._check(self.lib.shm_add_boxes(
            self._ctx, lev, b.ctypes.data_as(C.POINTER(C.c_int)), len(b)))

    def remake_level(self, lev: int, boxes) -> None:
        """RemakeLevel on a new box list."""
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        self._check(self.lib.shm_remake_level(
            self._ctx, lev, b.ctypes.data_as(C.POINTER(C.c_int)), len(b)))

    def local_regrid(self, lev: int) -> int:
        """DoAttemptRegrid from tagging level lev: 0 a global regrid is
        wanted, 1 boxes added, 2 delayed."""
        o = C.c_int()
        self._check(self.lib.shm_local_regrid(self._ctx, lev, C.byref(o)))
        return o.value

    def regrid_stats(self) -> dict:
        s = np.zeros(4, dtype=np.int64)
        nl = self.info["max_level"] + 1
        d = np.zeros(nl)
        self._check(self.lib.shm_regrid_stats(
            self._ctx, s.ctypes.data_as(C.POINTER(C.c_longlong)), _dptr(d), nl))
        return {"attempts": int(s[0]), "global_wanted": int(s[1]),
                "nregrids": int(s[2]), "veto_level": int(s[3]), "min_distance": d}

    def output_functions(self) -> list:
        """Names of the project's output functions, in id order."""
        n = C.c_int(0)
        self._check(self.lib.shm_output_functions(self._ctx, C.byref(n)))
        names = []
        for fid in range(n.value):
            buf = C.create_string_buffer(256)
            self._check(self.lib.shm_output_function_name(self._ctx, fid, buf, 256))
            names.append(buf.value.decode())
        return names

    def spectrum(self, fn: int):
        """Spectrum of output function fn on level 0: (bins, k_sq)."""
        n = C.c_int(0)
        self._check(self.lib.shm_spectrum(self._ctx, fn, None, None, C.byref(n)))
        bins = np.zeros(n.value)
        ks = np.zeros(n.value, dtype=np.int64)
        self._check(self.lib.shm_spectrum(
            self._ctx, fn, _dptr(bins), ks.ctypes.data_as(C.POINTER(C.c_longlong)),
            C.byref(n)))
        return bins, ks

    def projection(self, fn: int, mode: int = 0, max_level: int = -1):
        """Projection along z: (data, counts), both (dimN, dimN), x index first."""
        D = C.c_int(0)
        self._check(self.lib.shm_projection(self._ctx, fn, mode, max_level, None,
                                            None, C.byref(D)))
        data = np.zeros(D.value * D.value)
        cnt = np.zeros(D.value * D.value, dtype=np.int32)
        self._check(self.lib.shm_projection(
            self._ctx, fn, mode, max_level, _dptr(data),
            cnt.ctypes.data_as(C.POINTER(C.c_int)), C.byref(D)))
        return data.reshape(D.value, D.value), cnt.reshape(D.value, D.value)

    # HDF5 files (libhdf5 loaded at run time by the library).
    def read_initial_state(self, lev: int = 0, path: str | None = None) -> None:
        """FillLevel::FromInitialStateFile for level lev (path None:
        input.initial_state); a checkpoint folder restarts every level."""
        self._check(self.lib.shm_read_initial_state(
            self._ctx, lev, path.encode() if path else None))

    def write_checkpoint(self, folder: str, next_id=(), last_time_written=(),
                         idx_checkpoints: int = 0) -> None:
        nid = np.ascontiguousarray(np.asarray(next_id, dtype=np.int32))
        lt = np.ascontiguousarray(np.asarray(last_time_written, dtype=np.float64))
        if nid.size != lt.size:
            raise ValueError("next_id and last_time_written differ in length")
        self._check(self.lib.shm_write_checkpoint(
            self._ctx, folder.encode(),
            nid.ctypes.data_as(C.POINTER(C.c_int)) if nid.size else None,
            _dptr(lt) if lt.size else None, nid.size, idx_checkpoints))

    def read_checkpoint(self, folder: str, noutput: int | None = None,
                        idx_checkpoints: int = 0):
        """Restart from a checkpoint; with noutput, returns the stored
        (next_id, last_time_written) of the output types."""
        if noutput is None:
            self._check(self.lib.shm_read_checkpoint(self._ctx, folder.encode(), None,
                                                     None, 0, 0))
            return None
        nid = np.zeros(max(noutput, 1), dtype=np.int32)
        lt = np.zeros(max(noutput, 1))
        self._check(self.lib.shm_read_checkpoint(
            self._ctx, folder.encode(), nid.ctypes.data_as(C.POINTER(C.c_int)),
            _dptr(lt), noutput, idx_checkpoints))
        return nid[:noutput], lt[:noutput]

    def delete_checkpoint(self, folder: str) -> None:
        self._check(self.lib.shm_delete_checkpoint(self._ctx, folder.encode()))

    def write_gw_spectrum(self, prefix: str, modifier: int = -1, params=()) -> None:
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        self._check(self.lib.shm_write_gw_spectrum(
            self._ctx, prefix.encode(), modifier, _dptr(p) if p.size else None, p.size))

    def write_spectra(self, prefix: str, fns) -> None:
        f = np.ascontiguousarray(np.asarray(fns, dtype=np.int32))
        self._check(self.lib.shm_write_spectra(
            self._ctx, prefix.encode(), f.ctypes.data_as(C.POINTER(C.c_int)), f.size))

    def write_projections(self, prefix: str, fns, modes=None) -> None:
        f = np.ascontiguousarray(np.asarray(fns, dtype=np.int32))
        m = None if modes is None else np.ascontiguousarray(np.asarray(modes, dtype=np.int32))
        self._check(self.lib.shm_write_projections(
            self._ctx, prefix.encode(), f.ctypes.data_as(C.POINTER(C.c_int)),
            m.ctypes.data_as(C.POINTER(C.c_int)) if m is not None else None, f.size))

    def sync(self) -> None:
        self._check(self.lib.shm_sync(self._ctx))

    def stream_probe(self, nbytes: int = 4 << 30, reps: int = 10) -> dict:
        r, w, c = C.c_double(), C.c_double(), C.c_double()
        self._check(self.lib.shm_stream_probe(self._ctx, nbytes, reps, C.byref(r),
                                              C.byref(w), C.byref(c)))
        return {"read_gbs": r.value, "write_gbs": w.value, "copy_gbs": c.value}

    def autotune(self, lev: int = 0) -> None:
        self._check(self.lib.shm_autotune(self._ctx, lev))

    def set_launch(self, variant: int = -1, tile_z: int = 0, xcd_split: int = -1,
                   persistent: int = -1, zc_minor: int = -1, fused: int = -1,
                   fused_variant: int = -1) -> None:
        """Fixes fields of the stage-kernel launch shape (-1 / 0: not set).
        Each call replaces the whole shape: give every field to fix in one
        call."""
        cfg = _Launch(variant, tile_z, xcd_split, persistent, zc_minor, fused,
                      fused_variant)
        self._check(self.lib.shm_set_launch(self._ctx, C.byref(cfg)))

    def get_launch(self, lev: int = 0) -> dict:
        """Launch shape the stage kernels of level lev use now."""
        cfg = _Launch()
        self._check(self.lib.shm_get_launch(self._ctx, lev, C.byref(cfg)))
        return {n: getattr(cfg, n) for n, _ in _Launch._fields_}

    def set_graphs(self, on: bool) -> None:
        """Replay captured level steps (HIP graphs)."""
        self._check(self.lib.shm_set_graphs(self._ctx, int(on)))

    def set_timing(self, on: bool) -> None:
        self._check(self.lib.shm_set_timing(self._ctx, int(on)))

    def timing(self) -> dict:
        sm, hm, fm = C.c_double(), C.c_double(), C.c_double()
        sl, sc, hl, fl = (C.c_longlong(), C.c_longlong(), C.c_longlong(),
                          C.c_longlong())
        self._check(self.lib.shm_get_timing(
            self._ctx, C.byref(sm), C.byref(sl), C.byref(sc), C.byref(hm),
            C.byref(hl), C.byref(fm), C.byref(fl)))
        return {"stage_ms": sm.value, "stage_launches": sl.value,
                "stage_cells": sc.value, "halo_ms": hm.value,
                "halo_launches": hl.value, "fused_ms": fm.value,
                "fused_launches": fl.value}
