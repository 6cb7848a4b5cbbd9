"""ctypes binding of the FP64 CPU oracle (oracle/lib/liboracle.so).

The oracle is the checker: tests compare the HIP path against it, never the other way round.
Vector fields use the (nx, ny, nz, 3) numpy indexing of the reference's af.array and are
converted to the library's flat [3][nz][ny][nx] layout here.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "lib", "liboracle.so")

MU0 = 4e-7 * np.pi

DEMAG, EXCHANGE, DMI, UNIAXIAL, ZEEMAN, RKKY = 1, 2, 3, 4, 5, 6
ATOM_DIPOLE, ATOM_EXCHANGE, ATOM_DMI, ATOM_UNIAXIAL, ATOM_ZEEMAN = 11, 12, 13, 14, 15


class Mesh(C.Structure):
    _fields_ = [("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


# int (*)(void* user, double t, const double* m, double* h_out, void* stream); the oracle passes
# host pointers and a null stream
FIELD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p)
HOMOGENEOUS_CALLBACK = 8


class TermDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("flags", C.c_int32), ("value", C.c_double),
                ("axis", C.c_double * 3), ("field", C.c_double * 3),
                ("array0", C.POINTER(C.c_double)), ("array1", C.POINTER(C.c_double)),
                ("index0", C.POINTER(C.c_uint32)), ("callback", FIELD_FN), ("callback_user", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(Mesh), C.c_double, dp, C.POINTER(TermDesc), C.c_int, dp, C.c_int]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_add_term.argtypes = [C.c_void_p, C.POINTER(TermDesc), C.c_int]
        L.oracle_set_m.argtypes = [C.c_void_p, dp, C.c_int]
        L.oracle_get_m.argtypes = [C.c_void_p, dp]
        L.oracle_set_t.argtypes = [C.c_void_p, C.c_double]
        L.oracle_get_t.argtypes = [C.c_void_p]
        L.oracle_get_t.restype = C.c_double
        L.oracle_get_steps.argtypes = [C.c_void_p]
        L.oracle_get_steps.restype = C.c_int64
        L.oracle_has_ms_field.argtypes = [C.c_void_p]
        L.oracle_heff.argtypes = [C.c_void_p, C.c_uint32, dp]
        L.oracle_energy.argtypes = [C.c_void_p, C.c_uint32, dp, dp]
        L.oracle_mean_m.argtypes = [C.c_void_p, dp]
        L.oracle_llg_init.argtypes = [C.c_void_p, C.c_double, C.c_char_p, C.c_double, C.c_double, C.c_double,
                                      C.c_double, C.c_int, C.c_int, C.c_int]
        L.oracle_llg_set_alpha.argtypes = [C.c_void_p, C.c_double]
        L.oracle_llg_set_alpha_field.argtypes = [C.c_void_p, dp]
        i64p = C.POINTER(C.c_int64)
        L.oracle_llg_step.argtypes = [C.c_void_p, dp, dp, dp, i64p, i64p, i64p]
        L.oracle_llg_integrate_dense.argtypes = [C.c_void_p, C.c_double, C.c_double, dp, C.c_int64, i64p]
        L.oracle_llg_attempt.argtypes = [C.c_void_p, C.c_double, dp, dp]
        L.oracle_lbfgs.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.c_double, dp, i64p, i64p,
                                   C.POINTER(C.c_int)]
        L.oracle_lbfgs_trace.argtypes = [C.c_void_p, C.c_double, C.c_int64, dp, dp, C.c_int64, i64p]
        L.oracle_llg_counters.argtypes = [C.c_void_p, i64p]
        L.oracle_demag_tensor.argtypes = [C.POINTER(Mesh), dp, C.c_int]
        L.oracle_newell_nxx.argtypes = [C.c_int] * 3 + [C.c_double] * 3
        L.oracle_newell_nxx.restype = C.c_double
        L.oracle_newell_nxy.argtypes = [C.c_int] * 3 + [C.c_double] * 3
        L.oracle_newell_nxy.restype = C.c_double
        L.oracle_fft.argtypes = [dp, C.c_int64, C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def to_flat(v):
    """(nx, ny, nz, 3) array -> flat [3][nz][ny][nx] float64."""
    v = np.asarray(v, dtype=np.float64)
    return np.ascontiguousarray(v.transpose(3, 2, 1, 0)).ravel()


def from_flat(f, nx, ny, nz):
    return np.ascontiguousarray(np.asarray(f).reshape(3, nz, ny, nx).transpose(3, 2, 1, 0))


def scalar_flat(s):
    """(nx, ny, nz) scalar field -> flat [nz][ny][nx]."""
    s = np.asarray(s, dtype=np.float64)
    return np.ascontiguousarray(s.transpose(2, 1, 0)).ravel()


def term(kind, value=0.0, axis=(0.0, 0.0, 1.0), field=(0.0, 0.0, 0.0), array0=None, array1=None, index0=None,
         callback=None, homogeneous_callback=False):
    """Build a TermDesc; arrays are kept alive on the returned object.  callback(t, m_flat) returns
    the external field at time t: a 3-vector with homogeneous_callback, else the flat 3N field."""
    d = TermDesc()
    d.kind = kind
    d.value = value
    d.axis[:] = list(axis)
    d.field[:] = list(field)
    keep = []
    if callback is not None:
        d.flags = HOMOGENEOUS_CALLBACK if homogeneous_callback else 0

        def fn(user, t, m_ptr, h_ptr, stream, _cb=callback, _hom=homogeneous_callback):
            try:
                if _hom:
                    v = np.asarray(_cb(t, None), dtype=np.float64)
                    C.memmove(h_ptr, v.ctypes.data, 24)
                else:
                    n3 = d._n3
                    m = np.ctypeslib.as_array(C.cast(m_ptr, C.POINTER(C.c_double)), shape=(n3,)).copy()
                    v = np.ascontiguousarray(_cb(t, m), dtype=np.float64)
                    C.memmove(h_ptr, v.ctypes.data, 8 * n3)
                return 0
            except Exception:  # noqa: BLE001 - reported as a failed evaluation
                return 1
        f = FIELD_FN(fn)
        d.callback = f
        keep.append(f)
    if array0 is not None:
        a = np.ascontiguousarray(array0, dtype=np.float64)
        keep.append(a)
        d.array0 = _dp(a)
    if array1 is not None:
        a = np.ascontiguousarray(array1, dtype=np.float64)
        keep.append(a)
        d.array1 = _dp(a)
    if index0 is not None:
        a = np.ascontiguousarray(index0, dtype=np.uint32)
        keep.append(a)
        d.index0 = a.ctypes.data_as(C.POINTER(C.c_uint32))
    d._keep = keep
    return d


class Sim:
    """One oracle simulation: mesh + State + terms (+ optional LLG integrator)."""

    def __init__(self, mesh, Ms, m, terms, Ms_field=None, nthreads=0):
        self.mesh = Mesh(*mesh) if not isinstance(mesh, Mesh) else mesh
        self.nx, self.ny, self.nz = self.mesh.nx, self.mesh.ny, self.mesh.nz
        self.N = self.nx * self.ny * self.nz
        self.terms = list(terms)
        for t in self.terms:
            t._n3 = 3 * self.N  # full-field callbacks read and write 3N values
        arr = (TermDesc * max(1, len(self.terms)))(*self.terms)
        m_flat = to_flat(m) if np.ndim(m) == 4 else np.ascontiguousarray(m, dtype=np.float64)
        msf = None
        if Ms_field is not None:
            msf = scalar_flat(Ms_field) if np.ndim(Ms_field) == 3 else np.ascontiguousarray(Ms_field, np.float64)
        self.h = lib().oracle_create(C.byref(self.mesh), float(Ms), _dp(msf) if msf is not None else None, arr,
                                     len(self.terms), _dp(m_flat), nthreads)
        if not self.h:
            raise RuntimeError("oracle_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def add_term(self, d, nthreads=0):
        d._n3 = 3 * self.N
        self.terms.append(d)
        assert lib().oracle_add_term(self.h, C.byref(d), nthreads) == 0

    @property
    def has_ms_field(self):
        return bool(lib().oracle_has_ms_field(self.h))

    def m_flat(self):
        out = np.empty(3 * self.N)
        lib().oracle_get_m(self.h, _dp(out))
        return out

    def m(self):
        return from_flat(self.m_flat(), self.nx, self.ny, self.nz)

    def set_m(self, m, normalize=True):
        f = to_flat(m) if np.ndim(m) == 4 else np.ascontiguousarray(m, dtype=np.float64)
        lib().oracle_set_m(self.h, _dp(f), int(normalize))

    @property
    def t(self):
        return lib().oracle_get_t(self.h)

    @t.setter
    def t(self, v):
        lib().oracle_set_t(self.h, v)

    @property
    def steps(self):
        return lib().oracle_get_steps(self.h)

    def heff_flat(self, mask=0xFFFFFFFF):
        out = np.empty(3 * self.N)
        assert lib().oracle_heff(self.h, mask, _dp(out)) == 0
        return out

    def heff(self, mask=0xFFFFFFFF):
        return from_flat(self.heff_flat(mask), self.nx, self.ny, self.nz)

    def energy(self, mask=0xFFFFFFFF):
        e = np.zeros(max(1, len(self.terms)))
        tot = C.c_double()
        assert lib().oracle_energy(self.h, mask, _dp(e), C.byref(tot)) == 0
        return tot.value, e[: len(self.terms)]

    def mean_m(self):
        out = np.zeros(3)
        lib().oracle_mean_m(self.h, _dp(out))
        return out

    def llg_init(self, alpha, scheme="RKF45", hmin=1e-15, hmax=3.5e-10, atol=1e-6, rtol=1e-6,
                 dissipation_term_only=False, normalize=True, analytic_rhs=False):
        lib().oracle_llg_init(self.h, alpha, scheme.encode(), hmin, hmax, atol, rtol, int(dissipation_term_only),
                              int(normalize), int(analytic_rhs))

    def set_alpha(self, a):
        lib().oracle_llg_set_alpha(self.h, a)

    def set_alpha_field(self, alpha):
        """Cell-wise damping (nx, ny, nz) or flat [N]; None returns to the scalar alpha."""
        if alpha is None:
            lib().oracle_llg_set_alpha_field(self.h, None)
            return
        a = scalar_flat(alpha) if np.ndim(alpha) == 3 else np.ascontiguousarray(alpha, dtype=np.float64)
        self._alpha_keep = a
        lib().oracle_llg_set_alpha_field(self.h, _dp(a))

    def step(self):
        h, e, hn = C.c_double(), C.c_double(), C.c_double()
        acc, rej, ev = C.c_int64(), C.c_int64(), C.c_int64()
        assert lib().oracle_llg_step(self.h, C.byref(h), C.byref(e), C.byref(hn), C.byref(acc), C.byref(rej),
                                     C.byref(ev)) == 0
        return dict(h=h.value, err=e.value, h_next=hn.value, accepted=acc.value, rejected=rej.value,
                    heff_evals=ev.value)

    def attempt(self, dt):
        dm = np.empty(3 * self.N)
        e = C.c_double()
        assert lib().oracle_llg_attempt(self.h, dt, _dp(dm), C.byref(e)) == 0
        return dm, e.value

    def integrate_dense(self, time, dt_write, max_rows=100000):
        out = np.zeros(4 * max_rows)
        rows = C.c_int64()
        assert lib().oracle_llg_integrate_dense(self.h, time, dt_write, _dp(out), max_rows, C.byref(rows)) == 0
        return out[: 4 * rows.value].reshape(rows.value, 4)

    def counters(self):
        """Controller counters: accepted, rejected, hmin, hmax, maxscale, minscale."""
        out = (C.c_int64 * 6)()
        lib().oracle_llg_counters(self.h, out)
        return dict(zip(["accepted", "rejected", "hmin", "hmax", "maxscale", "minscale"], list(out)))

    def lbfgs_trace(self, tol=1e-6, maxiter=230, max_rows=100000):
        """LBFGS without the torque stop; returns (E, rows) with one row per line search:
        f(0), g(0).s, step, f(step), g(step).s, cvsrch info, evaluations."""
        E = C.c_double()
        out = np.zeros(7 * max_rows)
        rows = C.c_int64()
        assert lib().oracle_lbfgs_trace(self.h, tol, maxiter, C.byref(E), _dp(out), max_rows, C.byref(rows)) == 0
        return E.value, out[: 7 * rows.value].reshape(rows.value, 7)

    def lbfgs(self, tol=1e-6, maxiter=230, torque_tol=0.0):
        E, it, ev, rs = C.c_double(), C.c_int64(), C.c_int64(), C.c_int()
        assert lib().oracle_lbfgs(self.h, tol, maxiter, torque_tol, C.byref(E), C.byref(it), C.byref(ev),
                                  C.byref(rs)) == 0
        return dict(E=E.value, iterations=it.value, evals=ev.value, reason=rs.value)


def demag_tensor(mesh, nthreads=0):
    m = Mesh(*mesh) if not isinstance(mesh, Mesh) else mesh
    n0, n1 = 2 * m.nx, 2 * m.ny
    n2 = 1 if m.nz == 1 else 2 * m.nz
    out = np.empty(6 * n0 * n1 * n2)
    lib().oracle_demag_tensor(C.byref(m), _dp(out), nthreads)
    return out.reshape(6, n2, n1, n0)


def fft(x, sign=-1):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    lib().oracle_fft(a.ctypes.data_as(C.POINTER(C.c_double)), a.size, sign)
    return a
