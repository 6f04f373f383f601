"""gample_amd -- Python binding of libgample_amd.so (the MI355X PIC step).

The library is the product: a C-ABI over hand-written HIP kernels for gfx950
(include/gample_amd.h).  This module is a thin ctypes layer used by the tests
and the benchmark.  There is no fallback: if the shared library is missing,
importing :func:`lib` raises, and every call that fails on the device raises
:class:`GampleError` with the library's own message.

Names follow the reference's modules (Particle / EMField / Moment /
Collision, reference src/*.h); see :mod:`gample_amd.host` for the JSON
condition-file front end.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GAMPLE_AMD_LIB selects another build of the library (e.g. a tuning variant)
LIB_PATH = os.environ.get("GAMPLE_AMD_LIB") or os.path.join(_HERE, "libgample_amd.so")

GP_OK = 0
GP_NOT_CONVERGED = 1
PBC = {"none": -1, "periodic": 0, "Delete": 1, "Reflect": 2}
FBC = {"periodic": 0, "Dirichlet": 1, "Neumann": 2}

(F_RHO, F_EX, F_EY, F_EZ, F_BX, F_BY, F_BZ, F_PHI, F_JX, F_JY, F_DEL_EX, F_DEL_EY, F_DEL_BZ, F_SOL, F_CHI, F_AX,
 F_AY, F_RHS) = range(18)
# elliptic solver choices of solve_fields / solve_darwin / step
SOLVER_MG, SOLVER_NN, SOLVER_NN_WARM_MG, SOLVER_NN_FCG = 0, 1, 2, 3
G_RHO_ACC, G_JX_ACC, G_JY_ACC = range(3)


class GampleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gample_amd error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [
        ("ngx", C.c_int), ("ngy", C.c_int),
        ("dx", C.c_double), ("dy", C.c_double),
        ("weight_order", C.c_int), ("boris_order", C.c_int),
        ("pbc", C.c_int * 4), ("fbc", C.c_int * 4),
        ("fbc_value", C.c_double * 4),
        ("e0", C.c_double * 3), ("b0", C.c_double * 3),
        ("cur_dens", C.c_int), ("em_model", C.c_int), ("device", C.c_int),
    ]


class SpeciesDesc(C.Structure):
    _fields_ = [
        ("q", C.c_double), ("m", C.c_double), ("w", C.c_double),
        ("capacity", C.c_int64), ("vnorm", C.c_double), ("vth_hint", C.c_double),
    ]


class Gen(C.Structure):
    _fields_ = [
        ("gen_type", C.c_int),
        ("xmin", C.c_float), ("xmax", C.c_float), ("ymin", C.c_float), ("ymax", C.c_float),
        ("ele_u", C.c_float * 3), ("ele_vth", C.c_float),
        ("ion_u", C.c_float * 3), ("ion_vth", C.c_float),
        ("pert_delta", C.c_float), ("pert_mode", C.c_int),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("refinements", C.c_int), ("relres", C.c_double),
        ("seconds", C.c_double), ("used_nn", C.c_int), ("fell_back", C.c_int),
    ]


class Diag(C.Structure):
    _fields_ = [
        ("kinetic", C.c_double * 8), ("we_ref", C.c_double), ("we_cgs", C.c_double),
        ("joule", C.c_double), ("count", C.c_int64 * 8), ("absorbed", (C.c_int64 * 4) * 8),
        ("mean_cathode", C.c_double),
    ]


EXPORTS = [
    "gp_ctx_create", "gp_ctx_destroy", "gp_last_error", "gp_sync", "gp_dims",
    "gp_species_create", "gp_species_upload", "gp_species_download", "gp_species_count",
    "gp_species_grid_get", "gp_species_read", "gp_load_pair", "gp_generate", "gp_set_cathode", "gp_inject_wall", "gp_reset_grids", "gp_deposit",
    "gp_solve_fields", "gp_solve_darwin", "gp_darwin_stats", "gp_push", "gp_reduce", "gp_reduce_noalloc", "gp_sort", "gp_set_sort_drift", "gp_sort_stats", "gp_moments",
    "gp_step", "gp_field_get", "gp_field_set", "gp_field_energy", "gp_timing", "gp_kernel_time", "gp_stream_probe", "gp_nn_add_increment",
    "gp_nn_set_system", "gp_nn_add_darwin_increments",
    "gp_nn_init", "gp_nn_train", "gp_nn_infer", "gp_nn_save", "gp_nn_load", "gp_nn_add_pair", "gp_nn_params",
    "gp_nn_dims", "gp_nn_forward", "gp_nn_grad", "gp_solve_rhs", "gp_nn_config",
]

_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libgample_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C gample_amd)")
        L = C.CDLL(LIB_PATH)
        P, I64, D, I = C.c_void_p, C.c_int64, C.c_double, C.c_int
        fp = C.POINTER(C.c_float)
        L.gp_last_error.restype = C.c_char_p
        L.gp_ctx_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
        L.gp_ctx_destroy.argtypes = [P]
        L.gp_sync.argtypes = [P]
        L.gp_dims.argtypes = [P, C.POINTER(C.c_int)]
        L.gp_species_create.argtypes = [P, C.POINTER(SpeciesDesc), C.POINTER(C.c_int)]
        L.gp_species_upload.argtypes = [P, I, I64, fp, fp, fp, fp, fp]
        L.gp_species_download.argtypes = [P, I, I64, C.POINTER(C.c_int64), fp, fp, fp, fp, fp]
        L.gp_species_count.argtypes = [P, I, C.POINTER(C.c_int64)]
        L.gp_species_read.argtypes = [P, I, I64, I64, C.POINTER(C.c_int64), fp, fp, fp, fp, fp]
        L.gp_species_grid_get.argtypes = [P, I, I, C.POINTER(C.c_int64), I64]
        L.gp_load_pair.argtypes = [P, I, I, C.POINTER(Gen), C.c_uint64, I64, I64]
        L.gp_generate.argtypes = [P, I, I, C.POINTER(Gen), I64, I]
        L.gp_set_cathode.argtypes = [P, D]
        L.gp_inject_wall.argtypes = [P, I, I64, C.c_float, fp, D, C.c_uint64, C.c_uint64, I]
        L.gp_reset_grids.argtypes = [P]
        L.gp_deposit.argtypes = [P, I]
        L.gp_solve_fields.argtypes = [P, D, D, I, I, C.POINTER(SolveStats), C.POINTER(Diag)]
        L.gp_solve_darwin.argtypes = [P, D, D, I, I, C.POINTER(SolveStats)]
        L.gp_darwin_stats.argtypes = [P, C.POINTER(SolveStats)]
        L.gp_push.argtypes = [P, I, D, I]
        L.gp_reduce.argtypes = [P, I, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.gp_reduce_noalloc.argtypes = [P, I, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        L.gp_sort.argtypes = [P, I]
        L.gp_set_sort_drift.argtypes = [P, D]
        L.gp_sort_stats.argtypes = [P, I, C.POINTER(C.c_int64)]
        L.gp_moments.argtypes = [P, I, C.POINTER(C.c_double)]
        L.gp_step.argtypes = [P, D, D, I, I, I, I, C.POINTER(SolveStats), C.POINTER(Diag)]
        L.gp_field_get.argtypes = [P, I, P, I64]
        L.gp_field_set.argtypes = [P, I, P, I64]
        L.gp_field_energy.argtypes = [P, D, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gp_timing.argtypes = [P, I]
        L.gp_kernel_time.argtypes = [P, I, C.POINTER(C.c_double), C.POINTER(C.c_int64)]
        L.gp_stream_probe.argtypes = [P, I, I, C.POINTER(C.c_double)]
        L.gp_nn_add_increment.argtypes = [P, C.POINTER(I)]
        L.gp_nn_set_system.argtypes = [P, I]
        L.gp_nn_add_darwin_increments.argtypes = [P, C.POINTER(I)]
        L.gp_nn_init.argtypes = [P, I, C.c_uint64]
        L.gp_nn_train.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double), I, I, I, C.c_float,
                                  C.POINTER(C.c_float)]
        L.gp_nn_infer.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gp_nn_save.argtypes = [P, C.c_char_p]
        L.gp_nn_load.argtypes = [P, C.c_char_p]
        L.gp_nn_add_pair.argtypes = [P]
        L.gp_nn_config.argtypes = [P, I, I, I]
        L.gp_nn_params.argtypes = [P, fp, fp, C.POINTER(C.c_int64)]
        L.gp_nn_dims.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.gp_nn_forward.argtypes = [P, fp, I, fp]
        L.gp_nn_grad.argtypes = [P, fp, fp, I, I, fp, fp]
        L.gp_solve_rhs.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double), D, I, I, C.POINTER(SolveStats)]
        _lib = L
    return _lib


def _check(rc: int, allow_nc: bool = False) -> int:
    if rc < 0 or (rc == GP_NOT_CONVERGED and not allow_nc):
        raise GampleError(rc, lib().gp_last_error().decode())
    return rc


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(C.POINTER(C.c_float))


def make_config(ngx, ngy, dx, dy, order=3, boris=2, pbc=("periodic",) * 4, fbc=("periodic",) * 4,
                fbc_value=(0.0,) * 4, e0=(0.0, 0.0, 0.0), b0=(0.0, 0.0, 0.0), cur_dens=0, em_model=1,
                device=0) -> Config:
    c = Config()
    c.ngx, c.ngy, c.dx, c.dy = ngx, ngy, dx, dy
    c.weight_order, c.boris_order = order, boris
    for k in range(4):
        c.pbc[k] = PBC[pbc[k]] if isinstance(pbc[k], str) else int(pbc[k])
        c.fbc[k] = FBC[fbc[k]] if isinstance(fbc[k], str) else int(fbc[k])
        c.fbc_value[k] = fbc_value[k]
    for k in range(3):
        c.e0[k], c.b0[k] = e0[k], b0[k]
    c.cur_dens, c.em_model, c.device = cur_dens, em_model, device
    return c


class Context:
    """One device context: grids, solver hierarchy and species stores."""

    def __init__(self, cfg: Config):
        self.cfg = cfg
        self.h = C.c_void_p()
        _check(lib().gp_ctx_create(C.byref(cfg), C.byref(self.h)))
        d = (C.c_int * 6)()
        _check(lib().gp_dims(self.h, d))
        self.nx, self.ny, self.nkx, self.nky, self.ikmin, self.jkmin = list(d)

    def close(self):
        if self.h:
            lib().gp_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # species -----------------------------------------------------------
    def add_species(self, q, m, w, capacity, vth_hint=0.0, vnorm=0.0) -> int:
        d = SpeciesDesc(q, m, w, int(capacity), vnorm, vth_hint)
        sid = C.c_int()
        _check(lib().gp_species_create(self.h, C.byref(d), C.byref(sid)))
        return sid.value

    def upload(self, sid, x, y, vx, vy, vz):
        arrs = [_f32(a) for a in (x, y, vx, vy, vz)]
        n = len(arrs[0][0])
        _check(lib().gp_species_upload(self.h, sid, n, *[p for _, p in arrs]))

    def count(self, sid) -> int:
        n = C.c_int64()
        _check(lib().gp_species_count(self.h, sid, C.byref(n)))
        return n.value

    def download(self, sid):
        n = self.count(sid)
        out = [np.empty(n, np.float32) for _ in range(5)]
        nn = C.c_int64()
        _check(lib().gp_species_download(self.h, sid, n, C.byref(nn),
                                         *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in out]))
        return out

    def read(self, sid, first, count):
        """particles first .. first+count-1 in store order (clipped to the count)"""
        out = [np.empty(max(count, 0), np.float32) for _ in range(5)]
        nn = C.c_int64()
        _check(lib().gp_species_read(self.h, sid, first, count, C.byref(nn),
                                     *[a.ctypes.data_as(C.POINTER(C.c_float)) for a in out]))
        return [a[:nn.value] for a in out]

    def acc(self, sid, which=G_RHO_ACC):
        shape = {G_RHO_ACC: (self.ny + 1, self.nx + 1), G_JX_ACC: (self.ny + 1, self.nx + 2),
                 G_JY_ACC: (self.ny + 2, self.nx + 1)}[which]
        a = np.empty(shape, np.int64)
        _check(lib().gp_species_grid_get(self.h, sid, which, a.ctypes.data_as(C.POINTER(C.c_int64)), a.size))
        return a

    def load_pair(self, se, si, gen: Gen, seed: int, first_id: int, n: int):
        _check(lib().gp_load_pair(self.h, se, si, C.byref(gen), seed, first_id, n))

    def generate(self, se, si, gen: Gen, addn: int, deposit=1):
        """append addn generated pairs (electrons only for si < 0 or
        gen_type 3) from the context's continuing generation streams"""
        _check(lib().gp_generate(self.h, se, si, C.byref(gen), int(addn), deposit))

    def set_cathode(self, x_cm):
        _check(lib().gp_set_cathode(self.h, float(x_cm)))

    def inject_wall(self, sid, n, vth, u, dt, seed, stream, deposit=1):
        uu = (C.c_float * 3)(*u)
        _check(lib().gp_inject_wall(self.h, sid, n, vth, uu, dt, seed, stream, deposit))

    # step pieces -------------------------------------------------------
    def reset_grids(self):
        _check(lib().gp_reset_grids(self.h))

    def deposit(self, sid):
        _check(lib().gp_deposit(self.h, sid))

    def solve_fields(self, dt=0.0, tol=1e-8, maxiter=200, solver=0, diag=False):
        st = SolveStats()
        dg = Diag() if diag else None
        rc = _check(lib().gp_solve_fields(self.h, dt, tol, maxiter, solver, C.byref(st),
                                          C.byref(dg) if dg is not None else None), allow_nc=True)
        return (st, dg, rc) if diag else (st, rc)

    def solve_rhs(self, b, x0=None, tol=1e-8, maxiter=400, solver=0):
        """S x = b on the Poisson grid for a given b; returns (x, stats, rc)"""
        b = np.ascontiguousarray(b, np.float64).ravel()
        x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64).ravel()
        st = SolveStats()
        rc = _check(lib().gp_solve_rhs(self.h, b.ctypes.data_as(C.POINTER(C.c_double)),
                                       x.ctypes.data_as(C.POINTER(C.c_double)), tol, maxiter, solver, C.byref(st)),
                    allow_nc=True)
        return x, st, rc

    def solve_darwin(self, dt, tol=1e-8, maxiter=400, solver=0):
        st = (SolveStats * 3)()
        rc = _check(lib().gp_solve_darwin(self.h, dt, tol, maxiter, solver, st), allow_nc=True)
        return list(st), rc

    def darwin_stats(self):
        """statistics of the chi, Ax, Ay solves of the last Darwin solve"""
        st = (SolveStats * 3)()
        _check(lib().gp_darwin_stats(self.h, st))
        return list(st)

    def push(self, sid, dt, deposit=0):
        _check(lib().gp_push(self.h, sid, dt, deposit))

    def reduce(self, sid):
        cnt = (C.c_int64 * 4)()
        ke = C.c_double()
        _check(lib().gp_reduce(self.h, sid, cnt, C.byref(ke)))
        return list(cnt), ke.value

    def reduce_noalloc(self, sid):
        """gp_reduce with its scratch in the library's static device array"""
        cnt = (C.c_int64 * 4)()
        ke = C.c_double()
        _check(lib().gp_reduce_noalloc(self.h, sid, cnt, C.byref(ke)))
        return list(cnt), ke.value

    def sort(self, sid):
        _check(lib().gp_sort(self.h, sid))

    def set_sort_drift(self, fraction):
        """gp_step sorts a species when its last push gathered more than
        fraction * n particles outside their tile windows (0: off)"""
        _check(lib().gp_set_sort_drift(self.h, fraction))

    def sort_stats(self, sid):
        """(sorts so far, drifted particles of the last push read back)"""
        out = (C.c_int64 * 2)()
        _check(lib().gp_sort_stats(self.h, sid, out))
        return out[0], out[1]

    def moments(self, sid):
        a = np.empty((10, self.ny + 1, self.nx + 1), np.float64)
        _check(lib().gp_moments(self.h, sid, a.ctypes.data_as(C.POINTER(C.c_double))))
        return a

    def step(self, dt, tol=1e-8, maxiter=200, solver=0, sort_interval=0, cycle=1, diag=False):
        st = SolveStats()
        dg = Diag() if diag else None
        rc = _check(lib().gp_step(self.h, dt, tol, maxiter, solver, sort_interval, cycle, C.byref(st),
                                  C.byref(dg) if dg is not None else None), allow_nc=True)
        return st, dg

    def sync(self):
        _check(lib().gp_sync(self.h))

    # fields ------------------------------------------------------------
    def field_shape(self, which):
        nx, ny = self.nx, self.ny
        return {F_RHO: ((ny + 1, nx + 1), np.float64), F_EX: ((ny + 1, nx + 2), np.float32),
                F_EY: ((ny + 2, nx + 1), np.float32), F_EZ: ((ny + 1, nx + 1), np.float32),
                F_BX: ((ny + 2, nx + 1), np.float32), F_BY: ((ny + 1, nx + 2), np.float32),
                F_BZ: ((ny + 2, nx + 2), np.float32), F_PHI: ((ny + 3, nx + 3), np.float64),
                F_JX: ((ny + 1, nx + 2), np.float64), F_JY: ((ny + 2, nx + 1), np.float64),
                F_DEL_EX: ((ny + 1, nx + 2), np.float32), F_DEL_EY: ((ny + 2, nx + 1), np.float32),
                F_DEL_BZ: ((ny + 2, nx + 2), np.float32),
                F_SOL: ((self.nky + 1, self.nkx + 1), np.float64),
                F_RHS: ((self.nky + 1, self.nkx + 1), np.float64)}.get(which) or self._darwin_shape(which)

    def _darwin_shape(self, which):
        perx = self.cfg.fbc[0] == 0
        pery = self.cfg.fbc[2] == 0
        mx, my = self.cfg.ngx - (1 if perx else 0) + 1, self.cfg.ngy - (1 if pery else 0) + 1
        if which in (F_CHI, F_AX, F_AY):
            return ((my, mx), np.float64)
        raise KeyError(which)

    def field(self, which):
        shape, dt = self.field_shape(which)
        a = np.empty(shape, dt)
        _check(lib().gp_field_get(self.h, which, a.ctypes.data, a.nbytes))
        return a

    def set_field(self, which, a):
        shape, dt = self.field_shape(which)
        a = np.ascontiguousarray(a, dtype=dt).reshape(shape)
        _check(lib().gp_field_set(self.h, which, a.ctypes.data, a.nbytes))

    def energy(self, dt=0.0):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(lib().gp_field_energy(self.h, dt, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    # learned preconditioner -------------------------------------------
    def nn_init(self, base_channels=32, seed=1):
        _check(lib().gp_nn_init(self.h, base_channels, seed))

    def nn_config(self, loss_kind=1, smooth=2, fcg_cap=0):
        _check(lib().gp_nn_config(self.h, loss_kind, smooth, fcg_cap))

    def nn_add_pair(self):
        _check(lib().gp_nn_add_pair(self.h))


    def nn_add_increment(self):
        """append (b_n - b_{n-1}, x_n - x_{n-1}) of the last two Poisson solves
        (device resident); False on the first call, which only records them"""
        a = C.c_int()
        _check(lib().gp_nn_add_increment(self.h, C.byref(a)))
        return bool(a.value)

    def nn_set_system(self, sys):
        """training pairs for 0 the Poisson solve, 1 the Darwin solves (a
        change drops the pairs held for the other, keeps the weights)"""
        _check(lib().gp_nn_set_system(self.h, sys))

    def nn_add_darwin_increments(self):
        """append the last Darwin solve's chi, Ax, Ay increments; the pairs added"""
        a = C.c_int()
        _check(lib().gp_nn_add_darwin_increments(self.h, C.byref(a)))
        return a.value

    def nn_train(self, epochs, batch=4, lr=1e-3, rho=None, phi=None):
        loss = C.c_float()
        if rho is None:
            _check(lib().gp_nn_train(self.h, None, None, 0, epochs, batch, lr, C.byref(loss)))
        else:
            rho = np.ascontiguousarray(rho, np.float64)
            phi = np.ascontiguousarray(phi, np.float64)
            n = rho.shape[0]
            _check(lib().gp_nn_train(self.h, rho.ctypes.data_as(C.POINTER(C.c_double)),
                                     phi.ctypes.data_as(C.POINTER(C.c_double)), n, epochs, batch, lr, C.byref(loss)))
        return loss.value

    def nn_params(self, set_to=None):
        cnt = C.c_int64()
        _check(lib().gp_nn_params(self.h, None, None, C.byref(cnt)))
        out = np.empty(cnt.value, np.float32)
        if set_to is not None:
            v, vp = _f32(set_to)
            assert v.size == cnt.value
            _check(lib().gp_nn_params(self.h, None, vp, None))
        _check(lib().gp_nn_params(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), None, None))
        return out

    def nn_dims(self):
        H, W = C.c_int(), C.c_int()
        _check(lib().gp_nn_dims(self.h, C.byref(H), C.byref(W)))
        return H.value, W.value

    def nn_forward(self, x):
        x, xp = _f32(x)
        out = np.empty_like(x)
        _check(lib().gp_nn_forward(self.h, xp, x.shape[0], out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def nn_grad(self, x, t, kind=0):
        (x, xp), (t, tp) = _f32(x), _f32(t)
        cnt = C.c_int64()
        _check(lib().gp_nn_params(self.h, None, None, C.byref(cnt)))
        g = np.empty(cnt.value, np.float32)
        loss = C.c_float()
        _check(lib().gp_nn_grad(self.h, xp, tp, x.shape[0], kind, C.byref(loss), g.ctypes.data_as(C.POINTER(C.c_float))))
        return loss.value, g

    def nn_infer(self, rhs):
        rhs = np.ascontiguousarray(rhs, np.float64).ravel()
        out = np.empty_like(rhs)
        _check(lib().gp_nn_infer(self.h, rhs.ctypes.data_as(C.POINTER(C.c_double)),
                                 out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def nn_save(self, path):
        _check(lib().gp_nn_save(self.h, path.encode()))

    def nn_load(self, path):
        _check(lib().gp_nn_load(self.h, path.encode()))

    # timing ------------------------------------------------------------
    def timing(self, enable=True):
        _check(lib().gp_timing(self.h, 1 if enable else 0))

    def stream_probe(self, sid, reps=10):
        """ms per launch of a kernel streaming the species' arrays in the push's
        pattern (read 4 floats, write 4 per particle) into its scratch store"""
        ms = C.c_double()
        _check(lib().gp_stream_probe(self.h, sid, reps, C.byref(ms)))
        return ms.value

    def kernel_time(self, which=0):
        t, n = C.c_double(), C.c_int64()
        _check(lib().gp_kernel_time(self.h, which, C.byref(t), C.byref(n)))
        return t.value, n.value
