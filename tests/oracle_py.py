"""ctypes binding of the FP64 CPU oracle (oracle/_build/liboracle_<P>.so).

Test infrastructure only: tests use it as the checker; it is never the
product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_LIBS: dict[str, C.CDLL] = {}


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "-j4"], check=True)


def _load(project: str) -> C.CDLL:
    if project in _LIBS:
        return _LIBS[project]
    # ORACLE_LIB_OVERRIDE points at another build of the same library (for
    # example one compiled with a host sanitizer).
    path = os.environ.get("ORACLE_LIB_OVERRIDE") or os.path.join(
        ORACLE_DIR, "_build", f"liboracle_{project}.so")
    if not os.path.exists(path):
        build_oracle()
    lib = C.CDLL(path, mode=C.RTLD_LOCAL)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    dp, ip, llp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_longlong)
    sig = {
        "orc_last_error": (C.c_char_p, []),
        "orc_create": (vp, [C.c_char_p]),
        "orc_destroy": (None, [vp]),
        "orc_ncomp": (i, [vp]),
        "orc_ncells": (C.c_longlong, [vp, i]),
        "orc_set_boxes": (i, [vp, i, ip, i]),
        "orc_set_time": (i, [vp, i, d, i]),
        "orc_time": (d, [vp, i]),
        "orc_index_to_k": (d, [i, i, i]),
        "orc_upload": (i, [vp, i, i, dp]),
        "orc_download": (i, [vp, i, i, dp, i]),
        "orc_upload_old": (i, [vp, i, i, dp, d]),
        "orc_fill_patch": (i, [vp, i, d]),
        "orc_fill_coarse_patch": (i, [vp, i]),
        "orc_download_fab": (i, [vp, i, i, i, dp]),
        "orc_spectrum": (i, [vp, i, dp, llp, ip]),
        "orc_projection": (i, [vp, i, i, i, dp, ip, ip]),
        "orc_fill_rhs": (i, [vp, i, d, d, i]),
        "orc_download_rhs": (i, [vp, i, dp]),
        "orc_advance": (i, [vp, i]),
        "orc_create_shadow": (i, [vp]),
        "orc_shadow_time": (d, [vp]),
        "orc_te_tag": (i, [vp, i, i, C.POINTER(C.c_ubyte), llp, dp]),
        "orc_gw_spectrum": (i, [vp, i, i, i, i, dp, dp, ip]),
        "orc_events": (i, [vp, C.c_char_p, ip, dp, i, ip]),
        "orc_gw_spectrum_mod": (i, [vp, i, i, i, i, i, dp, i, dp, dp, ip]),
        "orc_init_state": (i, [vp, i, i, C.c_ulonglong, dp, i]),
        "orc_r2c": (None, [i, dp, dp]),
        "orc_evolve": (i, [vp, i]),
        "orc_last_tags": (i, [vp, i, llp, i]),
        "orc_gw_lambda": (d, [i, i, i, i, i, i, i, i, i]),
        "orc_unique_k2": (i, [i, llp, i]),
        "orc_boxes": (i, [vp, i, ip, i, ip]),
        "orc_regrid_select": (i, [vp, i, C.POINTER(C.c_ubyte), ip, i, ip, dp]),
        "orc_fix_nesting": (i, [vp, i, ip, i, ip, i, ip]),
        "orc_block_layout": (i, [ip, i, i, ip, i, ip]),
        "orc_add_boxes": (i, [vp, i, ip, i]),
        "orc_remake_level": (i, [vp, i, ip, i]),
        "orc_local_regrid": (i, [vp, i, ip]),
        "orc_regrid_stats": (i, [vp, llp, dp, i]),
        "orc_set_tags": (i, [vp, i, C.POINTER(C.c_ubyte)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIBS[project] = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def inputs_text(inputs) -> str:
    if isinstance(inputs, str):
        return inputs
    out = []
    for k, v in inputs.items():
        if isinstance(v, (list, tuple)):
            v = " ".join(str(x) for x in v)
        elif isinstance(v, bool):
            v = "true" if v else "false"
        out.append(f"{k} = {v}")
    return "\n".join(out) + "\n"


class Oracle:
    def __init__(self, project: str, inputs):
        self.lib = _load(project)
        self.h = self.lib.orc_create(inputs_text(inputs).encode())
        if not self.h:
            raise RuntimeError(self.lib.orc_last_error().decode())
        self.ncomp = self.lib.orc_ncomp(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def _ck(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.orc_last_error().decode())

    def ncells(self, lev):
        return self.lib.orc_ncells(self.h, lev)

    def set_boxes(self, lev, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        self._ck(self.lib.orc_set_boxes(self.h, lev,
                                        b.ctypes.data_as(C.POINTER(C.c_int)), len(b)))

    def set_time(self, lev, t, istep=0):
        self._ck(self.lib.orc_set_time(self.h, lev, t, istep))

    def time(self, lev):
        return self.lib.orc_time(self.h, lev)

    def upload(self, lev, comp, data):
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
        assert a.size == self.ncells(lev)
        self._ck(self.lib.orc_upload(self.h, lev, comp, _dp(a)))

    def download(self, lev, comp, old=False):
        a = np.empty(self.ncells(lev))
        self._ck(self.lib.orc_download(self.h, lev, comp, _dp(a), int(old)))
        return a

    def upload_old(self, lev, comp, data, t):
        """Valid cells of grid_old[lev] and its time."""
        a = np.ascontiguousarray(np.asarray(data, dtype=np.float64).ravel())
        assert a.size == self.ncells(lev)
        self._ck(self.lib.orc_upload_old(self.h, lev, comp, _dp(a), t))

    def fill_patch(self, lev, time):
        """FillPatch of grid_new[lev] at `time` (ghost cells included)."""
        self._ck(self.lib.orc_fill_patch(self.h, lev, time))

    def fill_coarse_patch(self, lev):
        """FillCoarsePatch: every cell of level lev from level lev-1."""
        self._ck(self.lib.orc_fill_coarse_patch(self.h, lev))

    def fab(self, lev, box, comp, shape):
        """Ghosted box `box` of grid_new[lev], indexed [k, j, i]; `shape` is
        its ghosted extent (nz, ny, nx)."""
        a = np.empty(int(np.prod(shape)))
        self._ck(self.lib.orc_download_fab(self.h, lev, box, comp, _dp(a)))
        return a.reshape(shape)

    def spectrum(self, fn):
        """Spectrum of output function fn on level 0: (bins, k_sq)."""
        n = C.c_int(0)
        self._ck(self.lib.orc_spectrum(self.h, fn, None, None, C.byref(n)))
        b = np.zeros(n.value)
        k = np.zeros(n.value, dtype=np.int64)
        self._ck(self.lib.orc_spectrum(self.h, fn, _dp(b),
                                       k.ctypes.data_as(C.POINTER(C.c_longlong)),
                                       C.byref(n)))
        return b, k

    def projection(self, fn, mode=0, max_level=-1):
        """Projection along z: (data, counts) as (dimN, dimN), x index first."""
        D = C.c_int(0)
        self._ck(self.lib.orc_projection(self.h, fn, mode, max_level, None, None,
                                         C.byref(D)))
        d = np.zeros(D.value ** 2)
        c = np.zeros(D.value ** 2, dtype=np.int32)
        self._ck(self.lib.orc_projection(self.h, fn, mode, max_level, _dp(d),
                                         c.ctypes.data_as(C.POINTER(C.c_int)),
                                         C.byref(D)))
        return d.reshape(D.value, D.value), c.reshape(D.value, D.value)

    def fill_rhs(self, lev, time, weight=0.0, add=False):
        self._ck(self.lib.orc_fill_rhs(self.h, lev, time, weight, int(add)))

    def rhs(self, lev, comp):
        a = np.empty(self.ncells(lev))
        self._ck(self.lib.orc_download_rhs(self.h, comp, _dp(a)))
        return a

    def advance(self, lev=0, nsteps=1):
        for _ in range(nsteps):
            self._ck(self.lib.orc_advance(self.h, lev))

    def evolve(self, ncoarse=1):
        """TimeStepper::Advance(0) repeated on the static layout."""
        self._ck(self.lib.orc_evolve(self.h, ncoarse))

    def last_tags(self, lev):
        """Tag counters [total, user, trunc...] of the last scheduled TE."""
        c = np.zeros(2 + self.ncomp, dtype=np.int64)
        self._ck(self.lib.orc_last_tags(self.h, lev,
                                        c.ctypes.data_as(C.POINTER(C.c_longlong)), c.size))
        return c

    def events(self):
        """Level operations of the evolves so far, in order: (kind, lev, t)
        with kind S (regrid scheduled for t), H (shadow level), A (advanced
        to t), D (lev+1 averaged onto lev), T (truncation errors and tags),
        R (regrid)."""
        n = C.c_int(0)
        self._ck(self.lib.orc_events(self.h, None, None, None, 0, C.byref(n)))
        k = C.create_string_buffer(max(n.value, 1))
        lv = np.zeros(max(n.value, 1), dtype=np.int32)
        t = np.zeros(max(n.value, 1))
        self._ck(self.lib.orc_events(self.h, k, lv.ctypes.data_as(C.POINTER(C.c_int)),
                                     _dp(t), n.value, C.byref(n)))
        return [(k.raw[i:i + 1].decode(), int(lv[i]), float(t[i])) for i in range(n.value)]

    def create_shadow(self):
        self._ck(self.lib.orc_create_shadow(self.h))

    def te_tag(self, lev, flags=7, want_te=False):
        n = self.ncells(lev)
        mask = np.zeros(n, dtype=np.uint8)
        counts = np.zeros(2 + self.ncomp, dtype=np.int64)
        te = np.zeros((self.ncomp, n // 8)) if want_te else None
        self._ck(self.lib.orc_te_tag(self.h, lev, flags,
                                     mask.ctypes.data_as(C.POINTER(C.c_ubyte)),
                                     counts.ctypes.data_as(C.POINTER(C.c_longlong)),
                                     _dp(te) if te is not None else None))
        return mask, counts, te

    def gw_spectrum(self, mode=0, nbins=256, zero_padding=1, projection_type=2,
                    modifier=-1, params=()):
        """GW spectrum with an optional project spectrum modifier (by id) and
        its parameters."""
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        pp = _dp(p) if p.size else None
        n = C.c_int(0)
        self._ck(self.lib.orc_gw_spectrum_mod(self.h, mode, nbins, zero_padding,
                                              projection_type, modifier, pp, p.size,
                                              None, None, C.byref(n)))
        b, k = np.zeros(n.value), np.zeros(n.value)
        self._ck(self.lib.orc_gw_spectrum_mod(self.h, mode, nbins, zero_padding,
                                              projection_type, modifier, pp, p.size,
                                              _dp(b), _dp(k), C.byref(n)))
        return b, k

    def init_state(self, lev, kind, seed, params=()):
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        self._ck(self.lib.orc_init_state(self.h, lev, kind, seed,
                                         _dp(p) if p.size else None, p.size))

    def index_to_k(self, a, N, ptype):
        return self.lib.orc_index_to_k(a, N, ptype)


def spectral_lib(project: str = "PhaseTransitionGW") -> C.CDLL:
    """The oracle library of a project, for its project-independent spectral
    helpers (r2c, Lambda, unique |k|^2)."""
    return _load(project)


def oracle_r2c(a: np.ndarray) -> np.ndarray:
    """The oracle's R2C DFT of a real N^3 array indexed [z, y, x]; returns
    complex [kz, ky, kx] with kx = 0..N/2."""
    lib = spectral_lib()
    N = a.shape[0]
    src = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros(N * N * (N // 2 + 1) * 2)
    lib.orc_r2c(N, _dp(src), _dp(out))
    return out.view(np.complex128).reshape(N, N, N // 2 + 1)


def oracle_lambda(i, j, l, m, abc, N, ptype) -> float:
    return spectral_lib().orc_gw_lambda(i, j, l, m, abc[0], abc[1], abc[2], N, ptype)


def oracle_unique_k2(N: int) -> np.ndarray:
    lib = spectral_lib()
    n = lib.orc_unique_k2(N, None, 0)
    out = np.zeros(n, dtype=np.int64)
    lib.orc_unique_k2(N, out.ctypes.data_as(C.POINTER(C.c_longlong)), n)
    return out


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def block_layout(project, blocks, bf):
    """Boxes of a block set: z-runs joined along y then x (oracle)."""
    lib = _load(project)
    b = np.ascontiguousarray(np.asarray(blocks, dtype=np.int32).reshape(-1, 3))
    n = C.c_int()
    if lib.orc_block_layout(_ip(b), len(b), bf, None, 0, C.byref(n)) != 0:
        raise RuntimeError(lib.orc_last_error().decode())
    out = np.zeros((n.value, 6), dtype=np.int32)
    if lib.orc_block_layout(_ip(b), len(b), bf, _ip(out), n.value, C.byref(n)) != 0:
        raise RuntimeError(lib.orc_last_error().decode())
    return out


def _oracle_regrid_methods():
    def boxes(self, lev):
        n = C.c_int()
        self._ck(self.lib.orc_boxes(self.h, lev, None, 0, C.byref(n)))
        b = np.zeros((n.value, 6), dtype=np.int32)
        self._ck(self.lib.orc_boxes(self.h, lev, _ip(b), n.value, C.byref(n)))
        return b

    def regrid_select(self, lev, mask=None):
        mp = None
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).ravel())
            assert m.size == self.ncells(lev)
            mp = m.ctypes.data_as(C.POINTER(C.c_ubyte))
        n, d = C.c_int(), C.c_double()
        self._ck(self.lib.orc_regrid_select(self.h, lev, mp, None, 0, C.byref(n), C.byref(d)))
        b = np.zeros((n.value, 3), dtype=np.int32)
        self._ck(self.lib.orc_regrid_select(self.h, lev, mp, _ip(b), n.value, C.byref(n),
                                            C.byref(d)))
        return b, d.value

    def fix_nesting(self, l, blocks):
        b = np.ascontiguousarray(np.asarray(blocks, dtype=np.int32).reshape(-1, 3))
        n = C.c_int()
        self._ck(self.lib.orc_fix_nesting(self.h, l, _ip(b), len(b), None, 0, C.byref(n)))
        out = np.zeros((n.value, 3), dtype=np.int32)
        self._ck(self.lib.orc_fix_nesting(self.h, l, _ip(b), len(b), _ip(out), n.value,
                                          C.byref(n)))
        return out

    def add_boxes(self, lev, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        self._ck(self.lib.orc_add_boxes(self.h, lev, _ip(b), len(b)))

    def remake_level(self, lev, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.int32).reshape(-1, 6))
        self._ck(self.lib.orc_remake_level(self.h, lev, _ip(b), len(b)))

    def local_regrid(self, lev):
        o = C.c_int()
        self._ck(self.lib.orc_local_regrid(self.h, lev, C.byref(o)))
        return o.value

    def regrid_stats(self, nlev):
        s = np.zeros(4, dtype=np.int64)
        d = np.zeros(nlev)
        self._ck(self.lib.orc_regrid_stats(self.h, s.ctypes.data_as(C.POINTER(C.c_longlong)),
                                           _dp(d), nlev))
        return {"attempts": int(s[0]), "global_wanted": int(s[1]),
                "nregrids": int(s[2]), "veto_level": int(s[3]), "min_distance": d}

    def set_tags(self, lev, mask):
        """Tag mask of level lev at its current time, as a tagging leaves it."""
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).ravel())
        assert m.size == self.ncells(lev)
        self._ck(self.lib.orc_set_tags(self.h, lev, m.ctypes.data_as(C.POINTER(C.c_ubyte))))

    for f in (set_tags, boxes, regrid_select, fix_nesting, add_boxes, remake_level, local_regrid,
              regrid_stats):
        setattr(Oracle, f.__name__, f)


_oracle_regrid_methods()
