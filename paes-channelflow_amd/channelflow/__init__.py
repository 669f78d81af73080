"""Python host binding of the MI355X channel-flow RK3 timestep.

This is a thin ctypes layer over the C-ABI declared in ``include/channelflow.h``
(the same entry points the Fortran driver binds with ISO_C_BINDING).  It
mirrors the reference driver's vocabulary: a run is created from the
GRIDSIZE/PARAMS/TIMING namelist values (les.F90:30-34), fields are loaded like
a restart file (read_field, read_write.f90:1-11), advanced stage by stage
(lower/upper/comp1/comprhs/press/comp2, les.F90:254-285) and read back like
sav_field (read_write.f90:31-40).

The shared library must have been built (``__graft_entry__.build()``); if it
is missing, importing this package raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libchannelflow.so")

CF_OK = 0
CF_E_ARG = -1
CF_E_HIP = -2
CF_E_UTAU = -3
CF_E_NONFINITE = -4
CF_E_ECLAMP = -5
CF_E_NODEV = -6
CF_PREC_FP64 = 0
CF_PREC_FP32 = 1
CF_NKERN = 11

EXPORTS = (
    "cf_create", "cf_destroy", "cf_set_state", "cf_seeded_init", "cf_get_state", "cf_get_tendency",
    "cf_get_comp1_tendency", "cf_stage", "cf_step", "cf_set_dt", "cf_sync", "cf_get_divz",
    "cf_get_scalars", "cf_compmn", "cf_time_steps", "cf_kernel_name", "cf_device_bytes",
    "cf_last_error",
    # learned SGS closure
    "cf_mlp_create", "cf_mlp_destroy", "cf_mlp_set_params", "cf_mlp_get_params", "cf_mlp_set_norm",
    "cf_mlp_get_norm", "cf_mlp_save", "cf_mlp_load", "cf_mlp_forward", "cf_mlp_reserve",
    "cf_mlp_put_samples", "cf_mlp_loss_grad", "cf_mlp_train", "cf_mlp_reset_optimizer",
    "cf_mlp_last_error", "cf_set_closure", "cf_get_closure_counts", "cf_filter_snapshot",
    "cf_calc_spectra", "cf_set_closure_grid",
    # allocation-free variant (no reference counterpart)
    "cf_calc_spectra_noalloc",
)
CF_CLOSURE_GRID_DEALIASED = 0
CF_CLOSURE_GRID_SMALL = 1
CF_MLP_NIN = 11
CF_MLP_NOUT = 6
CF_MLP_NPARAM = 9478


class CfParams(ctypes.Structure):
    """struct cf_params of include/channelflow.h."""
    _fields_ = [
        ("nnx", ctypes.c_int32), ("nny", ctypes.c_int32), ("nnz", ctypes.c_int32),
        ("xl", ctypes.c_double), ("yl", ctypes.c_double), ("zl", ctypes.c_double),
        ("z0", ctypes.c_double), ("dt", ctypes.c_double),
        ("imodel", ctypes.c_int32), ("isurface", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("device", ctypes.c_int32),
        ("alk", ctypes.c_double), ("ugal", ctypes.c_double), ("vgal", ctypes.c_double),
    ]


class ChannelFlowError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"channelflow error {code}: {msg}")
        self.code = code


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libchannelflow.so and declare the C signatures.  Raises when the
    library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    vp = ctypes.c_void_p
    lib.cf_create.argtypes = [ctypes.POINTER(CfParams), ctypes.POINTER(vp)]
    lib.cf_destroy.argtypes = [vp]
    lib.cf_set_state.argtypes = [vp, dp, dp, dp, dp]
    lib.cf_seeded_init.argtypes = [vp, ctypes.c_uint64, ctypes.c_double]
    lib.cf_get_state.argtypes = [vp, dp, dp, dp, dp, dp]
    lib.cf_get_tendency.argtypes = [vp, dp, dp, dp, dp]
    lib.cf_get_comp1_tendency.argtypes = [vp, dp, dp, dp]
    lib.cf_stage.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.cf_step.argtypes = [vp, ctypes.c_int]
    lib.cf_set_dt.argtypes = [vp, ctypes.c_double]
    lib.cf_sync.argtypes = [vp]
    lib.cf_get_divz.argtypes = [vp, dp]
    lib.cf_get_scalars.argtypes = [vp, dp]
    lib.cf_compmn.argtypes = [vp, dp]
    lib.cf_time_steps.argtypes = [vp, ctypes.c_int, dp, dp]
    lib.cf_kernel_name.argtypes = [ctypes.c_int]
    lib.cf_kernel_name.restype = ctypes.c_char_p
    lib.cf_device_bytes.argtypes = [vp]
    lib.cf_device_bytes.restype = ctypes.c_int64
    lib.cf_last_error.argtypes = [vp]
    lib.cf_last_error.restype = ctypes.c_char_p
    fp = ctypes.POINTER(ctypes.c_float)
    i64 = ctypes.c_int64
    lib.cf_mlp_create.argtypes = [ctypes.c_int32, ctypes.POINTER(vp)]
    lib.cf_mlp_destroy.argtypes = [vp]
    lib.cf_mlp_set_params.argtypes = [vp, fp]
    lib.cf_mlp_get_params.argtypes = [vp, fp]
    lib.cf_mlp_set_norm.argtypes = [vp, fp, fp, fp, fp]
    lib.cf_mlp_get_norm.argtypes = [vp, fp, fp, fp, fp]
    lib.cf_mlp_save.argtypes = [vp, ctypes.c_char_p]
    lib.cf_mlp_load.argtypes = [vp, ctypes.c_char_p]
    lib.cf_mlp_forward.argtypes = [vp, i64, fp, fp]
    lib.cf_mlp_reserve.argtypes = [vp, i64]
    lib.cf_mlp_put_samples.argtypes = [vp, i64, i64, fp, fp]
    lib.cf_mlp_loss_grad.argtypes = [vp, i64, i64, fp, fp]
    lib.cf_mlp_train.argtypes = [vp, ctypes.c_int32, i64, ctypes.c_float, ctypes.c_uint64, fp]
    lib.cf_mlp_reset_optimizer.argtypes = [vp]
    lib.cf_mlp_last_error.argtypes = [vp]
    lib.cf_mlp_last_error.restype = ctypes.c_char_p
    lib.cf_set_closure.argtypes = [vp, vp]
    lib.cf_get_closure_counts.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.cf_filter_snapshot.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, dp]
    lib.cf_calc_spectra.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), dp, dp, dp]
    lib.cf_calc_spectra_noalloc.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), dp, dp, dp]
    lib.cf_set_closure_grid.argtypes = [vp, ctypes.c_int32]
    for name in EXPORTS:
        if name not in ("cf_kernel_name", "cf_device_bytes", "cf_last_error", "cf_mlp_last_error"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def z0_of_retau(re_tau: float, h: float = 1.0, kappa: float = 0.4, B: float = 5.0) -> float:
    """Roughness length standing in for Re_tau: the reference has no molecular
    viscosity, so Re_tau enters only through z0 = h exp(-kappa B) / Re_tau."""
    return h * math.exp(-kappa * B) / re_tau


@dataclass
class RunParams:
    nnx: int
    nny: int
    nnz: int
    xl: float = 2.0 * math.pi
    yl: float = math.pi
    zl: float = 2.0
    z0: float = 7.519e-4
    dt: float = 0.01
    imodel: int = 3
    isurface: int = 1
    precision: int = CF_PREC_FP64
    device: int = 0
    alk: float = 0.0
    ugal: float = 0.0
    vgal: float = 0.0

    def to_c(self) -> CfParams:
        return CfParams(self.nnx, self.nny, self.nnz, self.xl, self.yl, self.zl, self.z0, self.dt,
                        self.imodel, self.isurface, self.precision, self.device, self.alk,
                        self.ugal, self.vgal)


class ChannelFlow:
    """One run on one GPU (a cf_ctx)."""

    def __init__(self, params: RunParams):
        self.lib = load_library()
        self.params = params
        self._h = ctypes.c_void_p()
        cp = params.to_c()
        rc = self.lib.cf_create(ctypes.byref(cp), ctypes.byref(self._h))
        if rc != CF_OK:
            raise ChannelFlowError(rc, self.lib.cf_last_error(None).decode())
        self.shape = (params.nnz, params.nny, params.nnx)

    def _check(self, rc):
        if rc != CF_OK:
            raise ChannelFlowError(rc, self.lib.cf_last_error(self._h).decode())

    def close(self):
        if self._h:
            self.lib.cf_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"expected shape {self.shape}, got {a.shape}")
        return a

    def set_state(self, u, v, w, e):
        u, v, w, e = (self._arr(x) for x in (u, v, w, e))
        self._check(self.lib.cf_set_state(self._h, _ptr(u), _ptr(v), _ptr(w), _ptr(e)))

    def seeded_init(self, seed: int = 10972, amp: float = 0.1):
        self._check(self.lib.cf_seeded_init(self._h, seed, amp))

    def get_state(self):
        out = [np.empty(self.shape) for _ in range(5)]
        self._check(self.lib.cf_get_state(self._h, *(_ptr(a) for a in out)))
        return dict(zip("uvwep", out))

    def get_tendency(self):
        out = [np.empty(self.shape) for _ in range(4)]
        self._check(self.lib.cf_get_tendency(self._h, *(_ptr(a) for a in out)))
        return dict(zip(("ru", "rv", "rw", "re"), out))

    def get_comp1_tendency(self):
        out = [np.empty(self.shape) for _ in range(3)]
        self._check(self.lib.cf_get_comp1_tendency(self._h, *(_ptr(a) for a in out)))
        return dict(zip(("ru", "rv", "rw"), out))

    def stage(self, istage: int, mnout: bool = False):
        self._check(self.lib.cf_stage(self._h, istage, int(bool(mnout))))

    def step(self, nsteps: int = 1):
        self._check(self.lib.cf_step(self._h, nsteps))

    def set_dt(self, dt: float):
        self._check(self.lib.cf_set_dt(self._h, dt))

    def sync(self):
        self._check(self.lib.cf_sync(self._h))

    def divz(self):
        out = np.empty(self.params.nnz)
        self._check(self.lib.cf_get_divz(self._h, _ptr(out)))
        return out

    def scalars(self):
        out = np.empty(4)
        self._check(self.lib.cf_get_scalars(self._h, _ptr(out)))
        return dict(zip(("utau", "utau_s", "dudt", "dvdt"), out))

    def compmn(self):
        """meanbuf(19, nnz) in the allocate.f90:89-111 row order."""
        out = np.empty((self.params.nnz, 19))
        self._check(self.lib.cf_compmn(self._h, _ptr(out)))
        return out.T.copy()

    def time_steps(self, nsteps: int, per_kernel: bool = False):
        tot = np.zeros(1)
        ker = np.zeros(CF_NKERN) if per_kernel else None
        self._check(self.lib.cf_time_steps(self._h, nsteps, _ptr(tot), _ptr(ker)))
        if per_kernel:
            names = [self.lib.cf_kernel_name(k).decode() for k in range(CF_NKERN)]
            return float(tot[0]), dict(zip(names, ker))
        return float(tot[0])

    def device_bytes(self) -> int:
        return int(self.lib.cf_device_bytes(self._h))

    def calc_spectra(self, zlevels):
        """calc_spectra.f90 on the device for 1-based levels ``zlevels``:
        (hats, var, lscales) in the layout of the reference's uhat(r, iz):
        hats[f] is (nnx/2+1, nlev) for f = u, v, w; var[f] (nlev,);
        lscales (l111, l112, l221, l222, l331, l332), each (nlev,)."""
        return self._spectra(self.lib.cf_calc_spectra, zlevels)

    def calc_spectra_noalloc(self, zlevels):
        """calc_spectra through cf_calc_spectra_noalloc (no device allocation;
        at most 64 levels); same results."""
        return self._spectra(self.lib.cf_calc_spectra_noalloc, zlevels)

    def _spectra(self, fn, zlevels):
        zl = np.ascontiguousarray(zlevels, dtype=np.int32)
        nl = zl.size
        nb = self.params.nnx // 2 + 1
        h = np.empty((3, nl, nb))
        var = np.empty((3, nl))
        ls = np.empty((6, nl))
        self._check(fn(self._h, nl, zl.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), _ptr(h), _ptr(var), _ptr(ls)))
        return [h[f].T.copy() for f in range(3)], [var[f] for f in range(3)], [ls[q] for q in range(6)]

    def set_closure(self, net: "SgsNet | None"):
        """Use the learned closure ``net`` from the next stage on (None: TKE)."""
        self._check(self.lib.cf_set_closure(self._h, net._h if net is not None else None))

    def set_closure_grid(self, grid: int):
        """CF_CLOSURE_GRID_DEALIASED (3/2 grid) or CF_CLOSURE_GRID_SMALL (nnx x nny)."""
        self._check(self.lib.cf_set_closure_grid(self._h, grid))

    def closure_counts(self):
        """(cells that used the network's stress, cells evaluated) since the last call."""
        out = (ctypes.c_int64 * 2)()
        self._check(self.lib.cf_get_closure_counts(self._h, out))
        return int(out[0]), int(out[1])

    def filter_snapshot(self, fx: int, fy: int, fz: int):
        """Box filter of the current state onto the (nnx/fx, nny/fy, nnz/fz)
        grid: dict of u, v, wc, wf, e, uu, vv, ww, uv, uw, vw and the fine-grid
        SGS stresses s11 s22 s33 s12 s13 s23, arrays (ncz, ncy, ncx)."""
        nz, ny, nx = self.shape
        shp = (nz // fz, ny // fy, nx // fx)
        out = np.empty((len(FILTER_NAMES),) + shp)
        self._check(self.lib.cf_filter_snapshot(self._h, fx, fy, fz, _ptr(out)))
        return dict(zip(FILTER_NAMES, out))


FILTER_NAMES = ("u", "v", "wc", "wf", "e", "uu", "vv", "ww", "uv", "uw", "vw",
                "s11", "s22", "s33", "s12", "s13", "s23")


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class SgsNet:
    """The learned SGS closure network on a device (a cf_mlp): parameters,
    standardisation, Adam state and training samples."""

    def __init__(self, device: int = 0, path: str | None = None):
        self.lib = load_library()
        self._h = ctypes.c_void_p()
        rc = self.lib.cf_mlp_create(device, ctypes.byref(self._h))
        if rc != CF_OK:
            raise ChannelFlowError(rc, self.lib.cf_mlp_last_error(None).decode())
        if path is not None:
            self.load(path)

    def _check(self, rc):
        if rc != CF_OK:
            raise ChannelFlowError(rc, self.lib.cf_mlp_last_error(self._h).decode())

    def close(self):
        if self._h:
            self.lib.cf_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _f32(a, shape=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        if shape is not None and a.shape != shape:
            raise ValueError(f"expected shape {shape}, got {a.shape}")
        return a

    def set_params(self, flat):
        p = self._f32(flat, (CF_MLP_NPARAM,))
        self._check(self.lib.cf_mlp_set_params(self._h, _fptr(p)))

    def get_params(self):
        p = np.empty(CF_MLP_NPARAM, dtype=np.float32)
        self._check(self.lib.cf_mlp_get_params(self._h, _fptr(p)))
        return p

    def set_norm(self, in_mu, in_sd, out_mu, out_sd):
        a = [self._f32(in_mu, (11,)), self._f32(in_sd, (11,)), self._f32(out_mu, (6,)), self._f32(out_sd, (6,))]
        self._check(self.lib.cf_mlp_set_norm(self._h, *(_fptr(x) for x in a)))

    def get_norm(self):
        a = [np.empty(11, np.float32), np.empty(11, np.float32), np.empty(6, np.float32), np.empty(6, np.float32)]
        self._check(self.lib.cf_mlp_get_norm(self._h, *(_fptr(x) for x in a)))
        return tuple(a)

    def save(self, path: str):
        self._check(self.lib.cf_mlp_save(self._h, os.fsencode(path)))

    def load(self, path: str):
        self._check(self.lib.cf_mlp_load(self._h, os.fsencode(path)))

    def forward(self, feats):
        f = self._f32(feats)
        n = f.shape[0]
        out = np.empty((n, CF_MLP_NOUT), dtype=np.float32)
        self._check(self.lib.cf_mlp_forward(self._h, n, _fptr(f), _fptr(out)))
        return out

    def reserve(self, capacity: int):
        self._check(self.lib.cf_mlp_reserve(self._h, int(capacity)))

    def put_samples(self, offset: int, feats, targets):
        f = self._f32(feats)
        t = self._f32(targets)
        assert f.shape[0] == t.shape[0] and f.shape[1] == CF_MLP_NIN and t.shape[1] == CF_MLP_NOUT
        self._check(self.lib.cf_mlp_put_samples(self._h, int(offset), f.shape[0], _fptr(f), _fptr(t)))

    def loss_grad(self, offset: int, n: int):
        loss = np.zeros(1, np.float32)
        g = np.empty(CF_MLP_NPARAM, np.float32)
        self._check(self.lib.cf_mlp_loss_grad(self._h, int(offset), int(n), _fptr(loss), _fptr(g)))
        return float(loss[0]), g

    def train(self, nsteps: int, batch: int, lr: float, seed: int = 0):
        losses = np.empty(nsteps, np.float32)
        self._check(self.lib.cf_mlp_train(self._h, nsteps, int(batch), lr, seed, _fptr(losses)))
        return losses

    def reset_optimizer(self):
        self._check(self.lib.cf_mlp_reset_optimizer(self._h))
