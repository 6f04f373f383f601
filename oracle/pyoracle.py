"""ctypes binding of the CPU oracle (oracle/libgp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package gample_amd/.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libgp_oracle.so")

PBC = {"none": -1, "periodic": 0, "Delete": 1, "Reflect": 2}
FBC = {"periodic": 0, "Dirichlet": 1, "Neumann": 2}


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


class Grid(C.Structure):
    _fields_ = [("ngx", C.c_int), ("ngy", C.c_int), ("ngb", C.c_int), ("order", C.c_int),
                ("boris_order", C.c_int), ("pbc", C.c_int * 4), ("dx", C.c_float), ("dy", C.c_float)]


class Gen(C.Structure):
    _fields_ = [("gen_type", C.c_int), ("xmin", C.c_float), ("xmax", C.c_float), ("ymin", C.c_float),
                ("ymax", C.c_float), ("ele_u", C.c_float * 3), ("ele_vth", C.c_float),
                ("ion_u", C.c_float * 3), ("ion_vth", C.c_float), ("dx", C.c_float), ("dy", C.c_float),
                ("ngb", C.c_float), ("pert_delta", C.c_float), ("pert_mode", C.c_int),
                ("xdom", C.c_float), ("ydom", C.c_float)]


class Field(C.Structure):
    _fields_ = [("ngx", C.c_int), ("ngy", C.c_int), ("ngb", C.c_int), ("dx", C.c_float), ("dy", C.c_float),
                ("fbc", C.c_int * 4), ("fval", C.c_double * 4)]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.gpo_weights_fast.argtypes = [C.c_int, C.c_float, P]
        L.gpo_weights_ref.argtypes = [C.c_int, C.c_float, P]
        L.gpo_weights_f64.argtypes = [C.c_int, C.c_double, P]
        L.gpo_push_fast.argtypes = [P, P, P, P, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_int64, P, P, P, P, P, C.c_int]
        L.gpo_push_ref.argtypes = [P, P, P, P, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_int64, P, P, P, P, P, P, C.c_int]
        L.gpo_deposit_fx.argtypes = [P, C.c_int64, P, P, P, P, C.c_float, P, P, P, C.c_int]
        L.gpo_deposit_ref.argtypes = [P, C.c_int64, P, P, P, P, C.c_float, P, P, P]
        L.gpo_deposit_f64.argtypes = [P, C.c_int64, P, P, P, P, C.c_double, P, P, P]
        L.gpo_fx_from_float.argtypes = [C.c_float, C.c_int]
        L.gpo_fx_from_float.restype = C.c_int64
        L.gpo_moments_f64.argtypes = [P, C.c_int64, P, P, P, P, P, C.c_double, P]
        L.gpo_reduce_literal.argtypes = [P, P, P, P, P, P, P, P]
        L.gpo_kinetic_energy.argtypes = [C.c_int64, P, P, P, P, C.c_float, C.c_float]
        L.gpo_kinetic_energy.restype = C.c_double
        L.gpo_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        L.gpo_splitmix64.restype = C.c_uint64
        L.gpo_pcg32.argtypes = [C.POINTER(C.c_uint64), C.c_uint64]
        L.gpo_pcg32.restype = C.c_uint32
        L.gpo_rand01.argtypes = [C.c_uint32]
        L.gpo_rand01.restype = C.c_float
        L.gpo_load_pair.argtypes = [P, C.c_uint64, C.c_int64, C.c_int64] + [P] * 10
        L.gpo_stream_states.argtypes = [C.c_uint64, C.c_int64, C.c_int64, P, P]
        L.gpo_generate.argtypes = [P, P, P, C.c_int64] + [P] * 10
        L.gpo_cathode_ramp.argtypes = [P, P, C.c_double, P]
        L.gpo_cathode_ramp.restype = C.c_double
        L.gpo_inject_wall.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_float, P, C.c_float, C.c_float,
                                      C.c_float, C.c_float] + [P] * 5
        L.gpo_poisson_dims.argtypes = [P] + [C.POINTER(C.c_int)] * 4
        L.gpo_fold_rho.argtypes = [P, P]
        L.gpo_poisson_csr.argtypes = [P, P, P, P, P]
        L.gpo_poisson_csr.restype = C.c_int64
        L.gpo_poisson_rhs.argtypes = [P, P, P, P]
        L.gpo_cg.argtypes = [C.c_int64, P, P, P, P, P, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.gpo_relres.argtypes = [C.c_int64, P, P, P, P, P]
        L.gpo_relres.restype = C.c_double
        L.gpo_reconstruct_phi.argtypes = [P, P, P]
        L.gpo_gradient.argtypes = [P, P, P, P]
        L.gpo_field_energy.argtypes = [P, P, P, P, P, C.c_double] + [C.POINTER(C.c_double)] * 3
        L.gpo_darwin_dims.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.gpo_darwin_csr.argtypes = [P, P, P, P]
        L.gpo_darwin_csr.restype = C.c_int64
        L.gpo_darwin_rhs_chi.argtypes = [P, P, P, P]
        L.gpo_darwin_rhs_a.argtypes = [P, P, P, P, P, P]
        _L = L
    return _L


def ptr(a):
    return None if a is None else a.ctypes.data


def grid(ngx, ngy, order, pbc=("periodic",) * 4, boris=2, dx=1.0, dy=1.0) -> Grid:
    g = Grid()
    g.ngx, g.ngy, g.order, g.ngb, g.boris_order = ngx, ngy, order, order // 2, boris
    for k in range(4):
        g.pbc[k] = PBC[pbc[k]] if isinstance(pbc[k], str) else int(pbc[k])
    g.dx, g.dy = dx, dy
    return g


def field(ngx, ngy, ngb, dx, dy, fbc=("periodic",) * 4, fval=(0.0,) * 4) -> Field:
    f = Field()
    f.ngx, f.ngy, f.ngb, f.dx, f.dy = ngx, ngy, ngb, dx, dy
    for k in range(4):
        f.fbc[k] = FBC[fbc[k]] if isinstance(fbc[k], str) else int(fbc[k])
        f.fval[k] = fval[k]
    return f


def weights(order, h, kind="fast"):
    out = np.zeros(6, np.float64 if kind == "f64" else np.float32)
    fn = {"fast": lib().gpo_weights_fast, "ref": lib().gpo_weights_ref, "f64": lib().gpo_weights_f64}[kind]
    fn(order, h, out.ctypes.data)
    return out[: order + 1]


def push_fast(g, Ex, Ey, Bz, consts, parts, uniform_bz=False, bz0=0.0, threads=8):
    x, y, vx, vy, vz = [np.ascontiguousarray(a, np.float32).copy() for a in parts]
    cE, cB, cX, cY = consts
    lib().gpo_push_fast(C.byref(g), ptr(Ex), ptr(Ey), ptr(Bz), int(uniform_bz), bz0, cE, cB, cX, cY,
                        len(x), ptr(x), ptr(y), ptr(vx), ptr(vy), ptr(vz), threads)
    return x, y, vx, vy, vz


def push_ref(g, Ex, Ey, Bz, consts, parts, threads=8):
    x, y, vx, vy, vz = [np.ascontiguousarray(a, np.float32).copy() for a in parts]
    flag = np.zeros(len(x), np.int32)
    cE, cB, cX, cY = consts
    lib().gpo_push_ref(C.byref(g), ptr(Ex), ptr(Ey), ptr(Bz), cE, cB, cX, cY, len(x), ptr(x), ptr(y), ptr(vx),
                       ptr(vy), ptr(vz), ptr(flag), threads)
    return x, y, vx, vy, vz, flag


def deposit_fx(g, parts, inv_vnorm=1.0, current=False, threads=8):
    x, y, vx, vy = [np.ascontiguousarray(a, np.float32) for a in parts[:4]]
    nx, ny = g.ngx + 2 * g.ngb, g.ngy + 2 * g.ngb
    rho = np.zeros((ny + 1, nx + 1), np.int64)
    jx = np.zeros((ny + 1, nx + 2), np.int64) if current else None
    jy = np.zeros((ny + 2, nx + 1), np.int64) if current else None
    rc = lib().gpo_deposit_fx(C.byref(g), len(x), ptr(x), ptr(y), ptr(vx), ptr(vy), inv_vnorm, ptr(rho), ptr(jx),
                              ptr(jy), threads)
    assert rc == 0
    return rho, jx, jy


def deposit_ref(g, parts, scale, current=False):
    x, y, vx, vy = [np.ascontiguousarray(a, np.float32) for a in parts[:4]]
    nx, ny = g.ngx + 2 * g.ngb, g.ngy + 2 * g.ngb
    rho = np.zeros((ny + 1, nx + 1), np.float32)
    jx = np.zeros((ny + 1, nx + 2), np.float32) if current else None
    jy = np.zeros((ny + 2, nx + 1), np.float32) if current else None
    lib().gpo_deposit_ref(C.byref(g), len(x), ptr(x), ptr(y), ptr(vx), ptr(vy), scale, ptr(rho), ptr(jx), ptr(jy))
    return rho, jx, jy


def deposit_f64(g, parts, scale, current=False):
    x, y, vx, vy = [np.ascontiguousarray(a, np.float32) for a in parts[:4]]
    nx, ny = g.ngx + 2 * g.ngb, g.ngy + 2 * g.ngb
    rho = np.zeros((ny + 1, nx + 1), np.float64)
    jx = np.zeros((ny + 1, nx + 2), np.float64) if current else None
    jy = np.zeros((ny + 2, nx + 1), np.float64) if current else None
    lib().gpo_deposit_f64(C.byref(g), len(x), ptr(x), ptr(y), ptr(vx), ptr(vy), scale, ptr(rho), ptr(jx), ptr(jy))
    return rho, jx, jy


def moments_f64(g, parts, scale):
    x, y, vx, vy, vz = [np.ascontiguousarray(a, np.float32) for a in parts]
    nx, ny = g.ngx + 2 * g.ngb, g.ngy + 2 * g.ngb
    m = np.zeros((10, ny + 1, nx + 1), np.float64)
    lib().gpo_moments_f64(C.byref(g), len(x), ptr(x), ptr(y), ptr(vx), ptr(vy), ptr(vz), scale, ptr(m))
    return m


def reduce_literal(parts, flags):
    x, y, vx, vy, vz = [np.ascontiguousarray(a, np.float32).copy() for a in parts]
    fl = np.ascontiguousarray(flags, np.int32).copy()
    n = C.c_int64(len(x))
    cnt = (C.c_int * 4)()
    lib().gpo_reduce_literal(C.byref(n), ptr(x), ptr(y), ptr(vx), ptr(vy), ptr(vz), ptr(fl), cnt)
    k = n.value
    return (x[:k], y[:k], vx[:k], vy[:k], vz[:k]), list(cnt), k


def kinetic(parts, m, w):
    x, _, vx, vy, vz = [np.ascontiguousarray(a, np.float32) for a in parts]
    return lib().gpo_kinetic_energy(len(x), ptr(x), ptr(vx), ptr(vy), ptr(vz), m, w)


def load_pair(gen: Gen, seed, first_id, n, ion=True):
    e = [np.zeros(n, np.float32) for _ in range(5)]
    i = [np.zeros(n, np.float32) for _ in range(5)] if ion else [None] * 5
    lib().gpo_load_pair(C.byref(gen), seed, first_id, n, *[ptr(a) for a in e], *[ptr(a) for a in i])
    return e, (i if ion else None)


def stream_states(seed, first_id, n):
    """(state, inc) of generation threads first_id .. first_id + n - 1"""
    st = np.zeros(n, np.uint64)
    inc = np.zeros(n, np.uint64)
    lib().gpo_stream_states(seed, first_id, n, ptr(st), ptr(inc))
    return st, inc


def generate(gen: Gen, state, inc, addn, ion=True):
    """one generation call on the persistent states (advanced in place)"""
    assert state.dtype == np.uint64 and len(state) >= addn and len(inc) >= addn
    e = [np.zeros(addn, np.float32) for _ in range(5)]
    i = [np.zeros(addn, np.float32) for _ in range(5)] if ion else [None] * 5
    lib().gpo_generate(C.byref(gen), ptr(state), ptr(inc), addn, *[ptr(a) for a in e], *[ptr(a) for a in i])
    return e, (i if ion else None)


def cathode_ramp(f, sol, pos, phi):
    """apply the hollow-cathode ramp to phi in place; returns the mean"""
    assert phi.dtype == np.float64 and phi.flags.c_contiguous
    return lib().gpo_cathode_ramp(C.byref(f), ptr(np.ascontiguousarray(sol, np.float64)), pos, ptr(phi))


def inject_wall(seed, stream, n, vth, u, dtdx, lo, ngx, ngy):
    out = [np.zeros(n, np.float32) for _ in range(5)]
    uu = np.asarray(u, np.float32)
    lib().gpo_inject_wall(seed, stream, n, vth, ptr(uu), dtdx, lo, ngx, ngy, *[ptr(a) for a in out])
    return out


def poisson_dims(f):
    a = [C.c_int() for _ in range(4)]
    lib().gpo_poisson_dims(C.byref(f), *[C.byref(v) for v in a])
    return [v.value for v in a]


def poisson_csr(f):
    nkx, nky, _, _ = poisson_dims(f)
    n = (nkx + 1) * (nky + 1)
    ptr_ = np.zeros(n + 1, np.int32)
    col = np.zeros(5 * n, np.int32)
    val = np.zeros(5 * n, np.float64)
    rbc = np.zeros(n, np.float64)
    nnz = lib().gpo_poisson_csr(C.byref(f), ptr(ptr_), ptr(col), ptr(val), ptr(rbc))
    return ptr_, col[:nnz], val[:nnz], rbc


def darwin_csr(f):
    a, b = C.c_int(), C.c_int()
    lib().gpo_darwin_dims(C.byref(f), C.byref(a), C.byref(b))
    n = (a.value + 1) * (b.value + 1)
    ptr_ = np.zeros(n + 1, np.int32)
    col = np.zeros(5 * n, np.int32)
    val = np.zeros(5 * n, np.float64)
    nnz = lib().gpo_darwin_csr(C.byref(f), ptr(ptr_), ptr(col), ptr(val))
    return ptr_, col[:nnz], val[:nnz], (a.value, b.value)


def fold(f, rho):
    r = np.ascontiguousarray(rho, np.float64).copy()
    lib().gpo_fold_rho(C.byref(f), ptr(r))
    return r


def fold_acc(acc, ngb, ngx, ngy, perx=True, pery=True):
    """The periodic branch of the ghost-charge fold (EMField.mm:749-850,
    x then y) on an int64 node accumulator of shape (ny+1, nx+1): each ghost
    node's charge is added to its interior image one period away, and the
    ghost is cleared.  Integer sums are exact, so two accumulators that hold
    the same charge at different periodic images fold to the same array."""
    a = np.array(acc, np.int64, copy=True)
    nx, ny = ngx + 2 * ngb, ngy + 2 * ngb
    assert a.shape == (ny + 1, nx + 1)
    if perx:
        for g in range(ngb):
            a[:, g + ngx] += a[:, g]
            a[:, g] = 0
        for g in range(ngb + ngx, nx + 1):
            a[:, g - ngx] += a[:, g]
            a[:, g] = 0
    if pery:
        for g in range(ngb):
            a[g + ngy, :] += a[g, :]
            a[g, :] = 0
        for g in range(ngb + ngy, ny + 1):
            a[g - ngy, :] += a[g, :]
            a[g, :] = 0
    return a


def poisson_rhs(f, rho, rbc):
    nkx, nky, _, _ = poisson_dims(f)
    out = np.zeros((nkx + 1) * (nky + 1), np.float64)
    lib().gpo_poisson_rhs(C.byref(f), ptr(np.ascontiguousarray(rho, np.float64)), ptr(rbc), ptr(out))
    return out


def cg(ptr_, col, val, rhs, x0=None, tol=1e-12, maxiter=100000, project=False):
    n = len(rhs)
    x = np.zeros(n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
    rr = C.c_double()
    it = lib().gpo_cg(n, ptr(ptr_), ptr(col), ptr(val), ptr(np.ascontiguousarray(rhs, np.float64)), ptr(x), tol,
                      maxiter, int(project), C.byref(rr))
    return x, it, rr.value


def relres(ptr_, col, val, b, x):
    return lib().gpo_relres(len(b), ptr(ptr_), ptr(col), ptr(val), ptr(np.ascontiguousarray(b, np.float64)),
                            ptr(np.ascontiguousarray(x, np.float64)))


def reconstruct(f, sol):
    nx, ny = f.ngx + 2 * f.ngb, f.ngy + 2 * f.ngb
    phi = np.zeros((ny + 3, nx + 3), np.float64)
    lib().gpo_reconstruct_phi(C.byref(f), ptr(np.ascontiguousarray(sol, np.float64)), ptr(phi))
    return phi


def gradient(f, phi):
    nx, ny = f.ngx + 2 * f.ngb, f.ngy + 2 * f.ngb
    Ex = np.zeros((ny + 1, nx + 2), np.float32)
    Ey = np.zeros((ny + 2, nx + 1), np.float32)
    lib().gpo_gradient(C.byref(f), ptr(np.ascontiguousarray(phi, np.float64)), ptr(Ex), ptr(Ey))
    return Ex, Ey


def field_energy(f, Ex, Ey, jx=None, jy=None, dt=0.0):
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().gpo_field_energy(C.byref(f), ptr(Ex), ptr(Ey), ptr(jx), ptr(jy), dt, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def darwin_rhs_chi(f, jx, jy):
    a, b = darwin_csr(f)[3]
    out = np.zeros((a + 1) * (b + 1))
    lib().gpo_darwin_rhs_chi(C.byref(f), ptr(np.ascontiguousarray(jx, np.float64)),
                             ptr(np.ascontiguousarray(jy, np.float64)), ptr(out))
    return out


def darwin_rhs_a(f, jx, jy, chi):
    a, b = darwin_csr(f)[3]
    ox = np.zeros((a + 1) * (b + 1))
    oy = np.zeros((a + 1) * (b + 1))
    lib().gpo_darwin_rhs_a(C.byref(f), ptr(np.ascontiguousarray(jx, np.float64)),
                           ptr(np.ascontiguousarray(jy, np.float64)), ptr(np.ascontiguousarray(chi, np.float64)),
                           ptr(ox), ptr(oy))
    return ox, oy
