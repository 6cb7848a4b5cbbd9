"""Python host layer over libmafx mirroring the reference's `magnumaf` module for the hot path
(python/magnumaf.pyx:522-1836): Mesh, State, the field terms, LLGIntegrator and LBFGS_Minimizer,
with the reference's class names, argument names, defaults and units.

Arrays in and out use the reference's indexing m[x, y, z, c] (numpy, shape (nx, ny, nz, 3)).
The magnetization lives on the GPU (a torch float64 tensor, torch serving only as the device
allocator); every field evaluation, integrator step and minimizer iteration runs in libmafx.
"""
import ctypes as C
import math
import os
import sys

import numpy as np

from . import _abi
from ._abi import check, lib

__all__ = ["Mesh", "State", "Constants", "DemagField", "ExchangeField", "DmiField", "UniaxialAnisotropyField",
           "ExternalField", "RKKYExchangeField", "AtomisticDipoleDipoleField", "AtomisticExchangeField",
           "AtomisticDmiField", "AtomisticUniaxialAnisotropyField", "AtomisticExternalField", "LLGIntegrator",
           "LBFGS_Minimizer", "Context", "to_flat", "from_flat"]


class Constants:
    """magnumaf/include/magnumaf/constants.hpp:9-22"""
    mu0 = 4e-7 * math.pi
    gamma = 1.760859644e11 * mu0
    mu_b = 9.274009994e-24
    e = -1.6021766208e-19
    kb = 1.38064852e-23
    hbar = 1.0545718e-34


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("magnumaf (MI355X build) needs a GPU: torch.cuda.is_available() is False")
    return torch


def to_flat(v):
    """(nx, ny, nz, 3) -> flat [3][nz][ny][nx] float64 (the af::array column-major layout)."""
    v = np.asarray(v, dtype=np.float64)
    return np.ascontiguousarray(v.transpose(3, 2, 1, 0)).ravel()


def from_flat(f, nx, ny, nz):
    return np.ascontiguousarray(np.asarray(f).reshape(3, nz, ny, nx).transpose(3, 2, 1, 0))


def _scalar_flat(s, mesh):
    s = np.asarray(s, dtype=np.float64)
    if s.ndim == 4:
        s = s[..., 0]
    if s.ndim == 3:
        return np.ascontiguousarray(s.transpose(2, 1, 0)).ravel()
    return np.ascontiguousarray(s.ravel())


def from_scalar_flat(f, mesh):
    """flat [nz][ny][nx] -> (nx, ny, nz)."""
    return np.ascontiguousarray(np.asarray(f).reshape(mesh.nz, mesh.ny, mesh.nx).transpose(2, 1, 0))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Mesh:
    """Mesh(nx, ny, nz, dx, dy, dz) (mesh.hpp:11-20, magnumaf.pyx:522-616)."""

    def __init__(self, nx, ny, nz, dx, dy, dz):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)

    def _c(self):
        return _abi.Mesh(self.nx, self.ny, self.nz, self.dx, self.dy, self.dz)

    @property
    def n(self):
        return self.nx * self.ny * self.nz

    def key(self):
        return (self.nx, self.ny, self.nz, self.dx, self.dy, self.dz)


class Context:
    """One libmafx evaluation context: mesh, material, an ordered term list, precision.

    The context remembers the version of every term and of the state's material it was built
    from; refresh() uploads what changed since (mafx_ctx_update_term, mafx_ctx_set_material), so
    a term or material changed between evaluations is seen by the next one, as in the reference,
    whose terms and State are read at every evaluation."""

    def __init__(self, mesh, Ms, Ms_field, terms, precision="f64", material_version=0):
        self.mesh = mesh
        self.terms = list(terms)
        descs, self._keep_terms = [], []
        for t in self.terms:
            d, keep = t._desc(mesh)
            descs.append(d)
            self._keep_terms.append((d, keep))
        self._versions = [t._version for t in self.terms]
        arr = (_abi.TermDesc * len(descs))(*descs)
        msf = None
        if Ms_field is not None:
            msf = np.ascontiguousarray(Ms_field, dtype=np.float64)
        prec = _abi.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        self.precision = prec
        h = C.c_void_p()
        torch = _torch()
        torch.cuda.init()
        check(lib().mafx_ctx_create(C.byref(mesh._c()), float(Ms) if Ms is not None else 0.0,
                                    _dp(msf) if msf is not None else None, arr, len(descs), prec, C.byref(h)))
        self.h = h
        self.material_version = material_version
        check(lib().mafx_ctx_set_stream(self.h, C.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def refresh(self, state):
        """Uploads the parameters of terms changed since the last call, the state's material when it
        changed, and the state's time (read by time-dependent external fields)."""
        for i, t in enumerate(self.terms):
            if t._version != self._versions[i]:
                d, keep = t._desc(self.mesh)
                check(lib().mafx_ctx_update_term(self.h, i, C.byref(d)))
                self._keep_terms[i] = (d, keep)
                self._versions[i] = t._version
        if state is not None:
            if state._material_version != self.material_version:
                msf = state.Ms_field
                if msf is not None:
                    msf = np.ascontiguousarray(msf, dtype=np.float64)
                Ms = state._Ms if msf is None else 0.0
                check(lib().mafx_ctx_set_material(self.h, float(Ms), _dp(msf) if msf is not None else None))
                self.material_version = state._material_version
            check(lib().mafx_ctx_set_time(self.h, float(state.t)))

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the process releases the device memory itself
        if getattr(self, "h", None) and not _finalizing():
            lib().mafx_ctx_destroy(self.h)
            self.h = None

    def heff(self, m_dev, mask=0xFFFFFFFF, out=None, f32=False):
        torch = _torch()
        n3 = 3 * self.mesh.n
        if out is None:
            out = torch.empty(n3, dtype=torch.float32 if f32 else torch.float64, device="cuda")
        check(lib().mafx_heff(self.h, C.c_void_p(m_dev.data_ptr()), mask, C.c_void_p(out.data_ptr()), int(f32)))
        return out

    def energy(self, m_dev, mask=0xFFFFFFFF):
        e = np.zeros(len(self.terms))
        tot = C.c_double()
        check(lib().mafx_energy(self.h, C.c_void_p(m_dev.data_ptr()), mask, _dp(e), C.byref(tot)))
        return tot.value, e

    def tensor_source(self):
        """0 no demag tensor, 1 computed, 2 read from the on-disk cache."""
        v = C.c_int()
        check(lib().mafx_ctx_tensor_source(self.h, C.byref(v)))
        return v.value

    def device_bytes(self):
        b = C.c_int64()
        check(lib().mafx_ctx_device_bytes(self.h, C.byref(b)))
        return b.value

    def prof(self, on=True):
        check(lib().mafx_prof_enable(self.h, int(on)))

    def prof_reset(self):
        check(lib().mafx_prof_reset(self.h))

    def prof_report(self):
        buf = C.create_string_buffer(1 << 16)
        check(lib().mafx_prof_report(self.h, buf, len(buf)))
        rows = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms, by = line.split()
            rows[name] = dict(count=int(cnt), ms=float(ms), bytes=float(by))
        return rows


class State:
    """State(mesh, Ms, m, verbose=True, mute_warning=False) (state.cpp:91-111, magnumaf.pyx:694-878).

    m is normalised with zero vectors kept at zero; when Ms is a float and some |m| == 0, an
    Ms_field = Ms * (|m| != 0) is built as the reference does (state.cpp:71-89).  Ms may be a
    float or a cell-wise field (nx, ny, nz) / (nx, ny, nz, 1); assigning state.Ms or state.Ms_field
    later changes the material of every following evaluation (magnumaf.pyx:777-784,839-842)."""

    def __init__(self, mesh, Ms, m, verbose=True, mute_warning=False, precision="f64"):
        torch = _torch()
        self.mesh = mesh
        self.t = 0.0
        self.steps = 0
        self.verbose = verbose
        mf = to_flat(m) if np.ndim(m) == 4 else np.array(m, dtype=np.float64).ravel()
        N = mesh.n
        a, b, c = mf[:N], mf[N:2 * N], mf[2 * N:]
        nrm = np.sqrt(a * a + b * b + c * c)
        with np.errstate(invalid="ignore", divide="ignore"):
            mf = np.where(np.concatenate([nrm, nrm, nrm]) != 0, mf / np.concatenate([nrm, nrm, nrm]), 0.0)
        self._material_version = 0
        if np.ndim(Ms) == 0:
            self._Ms = float(Ms)
            self._Ms_field = None
            if np.min(nrm) == 0:
                if verbose:
                    print("Info: in state.cpp: initial m has values with zero norm, building Ms_field array")
                self._Ms_field = self._Ms * (nrm != 0).astype(np.float64)
        else:
            self._Ms = float("nan")
            self._Ms_field = _scalar_flat(Ms, mesh)
        self._m = torch.from_numpy(mf).to("cuda")
        self._ctx = {}

    # -- material: State::Ms, State::Ms_field (magnumaf.pyx:777-784,839-842)
    @property
    def Ms(self):
        """Ms as the reference's Python returns it: the float, or the field when one is set."""
        return self._Ms if self._Ms_field is None else from_scalar_flat(self._Ms_field, self.mesh)

    @Ms.setter
    def Ms(self, value):
        if np.ndim(value) == 0:
            self._Ms = float(value)
        else:
            self._Ms_field = _scalar_flat(value, self.mesh)
        self._material_version += 1

    @property
    def Ms_field(self):
        """The cell-wise Ms (flat [N], x fastest) or None."""
        return self._Ms_field

    @Ms_field.setter
    def Ms_field(self, value):
        self._Ms_field = None if value is None else _scalar_flat(value, self.mesh)
        self._material_version += 1

    # -- material key for context reuse
    def _context(self, terms, precision="f64"):
        key = (tuple(id(t) for t in terms), precision)
        ctx = self._ctx.get(key)
        if ctx is None:
            ctx = Context(self.mesh, self._Ms if self._Ms_field is None else 0.0, self._Ms_field, terms, precision,
                          self._material_version)
            self._ctx[key] = ctx
        ctx.refresh(self)
        return ctx

    @property
    def m(self):
        return from_flat(self._m.cpu().numpy(), self.mesh.nx, self.mesh.ny, self.mesh.nz)

    @m.setter
    def m(self, value):
        mf = to_flat(value)
        N = self.mesh.n
        a, b, c = mf[:N], mf[N:2 * N], mf[2 * N:]
        nrm = np.sqrt(a * a + b * b + c * c)
        nrm3 = np.concatenate([nrm, nrm, nrm])
        with np.errstate(invalid="ignore", divide="ignore"):
            mf = np.where(nrm3 != 0, mf / nrm3, 0.0)
        self._m.copy_(_torch().from_numpy(mf))

    def set_m_raw(self, value):
        """Sets m without renormalisation (for energy sweeps over prescribed states)."""
        self._m.copy_(_torch().from_numpy(to_flat(value)))

    @property
    def m_dev(self):
        return self._m

    def _util_ctx(self):
        # a context with one zero external field serves the vector helpers
        if "_util" not in self._ctx:
            self._ctx["_util"] = Context(self.mesh, self._Ms if self._Ms_field is None else 0.0, self._Ms_field,
                                         [ExternalField(homogeneous=(0.0, 0.0, 0.0))], "f64", self._material_version)
        return self._ctx["_util"]

    def normalize(self):
        check(lib().mafx_normalize(self._util_ctx().h, C.c_void_p(self._m.data_ptr()), 1))

    def mean_m(self, i=None):
        """<m> over cells with |m| != 0 (state.cpp:32-45)."""
        out = np.zeros(3)
        check(lib().mafx_mean_m(self._util_ctx().h, C.c_void_p(self._m.data_ptr()), _dp(out)))
        mean = (float(out[0]), float(out[1]), float(out[2]))
        return mean if i is None else mean[i]

    def mean_mx(self):
        return self.mean_m(0)

    def mean_my(self):
        return self.mean_m(1)

    def mean_mz(self):
        return self.mean_m(2)

    def __str__(self):
        """'t <mx> <my> <mz>' as the reference's Python State prints it (magnumaf.pyx:754-757)."""
        mx, my, mz = self.mean_m()
        return str(self.t) + " " + str(mx) + " " + str(my) + " " + str(mz)

    def stream_line(self):
        """'t<TAB><mx><TAB><my><TAB><mz>' as the reference's C++ operator<< prints it (state.cpp:26-30)."""
        from . import io
        return io.stream_line(self.t, self.mean_m())

    def write_vti(self, outputname):
        """State::write_vti (state.cpp:139): m as cell data of a .vti image grid."""
        from . import io
        return io.write_vti(outputname, self.m, (self.mesh.dx, self.mesh.dy, self.mesh.dz))

    def read_vti(self, inputname):
        """State::_vti_reader (state.cpp:141): m (not renormalised) from a .vti file of this mesh."""
        from . import io
        m, _ = io.read_vti(inputname)
        if m.shape[:3] != (self.mesh.nx, self.mesh.ny, self.mesh.nz):
            raise ValueError(f"read_vti: file holds {m.shape[:3]} cells, the mesh {self.mesh.nx, self.mesh.ny, self.mesh.nz}")
        self.set_m_raw(m)

    _vti_reader = read_vti


# ---------------------------------------------------------------- field terms
class FieldTerm:
    """Base of the field terms (field_term.hpp:12-51): H_in_Apm(state), Energy_in_J(state).

    Assigning any public attribute of a term bumps its version; contexts holding the term upload
    the new parameters before their next evaluation (Context.refresh)."""

    _version = 0

    def __setattr__(self, name, value):
        object.__setattr__(self, name, value)
        if not name.startswith("_"):
            object.__setattr__(self, "_version", self._version + 1)

    def _desc(self, mesh):
        raise NotImplementedError

    def H_in_Apm(self, state, precision="f64"):
        ctx = state._context([self], precision)
        h = ctx.heff(state.m_dev)
        return from_flat(h.cpu().numpy(), state.mesh.nx, state.mesh.ny, state.mesh.nz)

    def H_in_T(self, state):
        return self.H_in_Apm(state) * Constants.mu0

    def Energy_in_J(self, state, precision="f64"):
        ctx = state._context([self], precision)
        return ctx.energy(state.m_dev)[0]

    def Energy_in_eV(self, state):
        return self.Energy_in_J(state) / 1.6021766208e-19


def _td(kind, value=0.0, axis=(0.0, 0.0, 1.0), field=(0.0, 0.0, 0.0)):
    d = _abi.TermDesc()
    d.kind = kind
    d.value = float(value)
    d.axis[:] = [float(v) for v in axis]
    d.field[:] = [float(v) for v in field]
    return d


def _is_array(v):
    return not isinstance(v, (int, float)) and np.ndim(v) > 0 and np.size(v) > 3


class DemagField(FieldTerm):
    """DemagField(mesh, verbose=False, caching=False, nthreads=4) (demag_field.hpp:14-30,
    magnumaf.pyx:1086-1121).  caching reads the transformed tensor from the on-disk cache
    (MAFX_CACHE_DIR, else <tmp>/magnumaf.cache) and writes it there when absent; verbose reports
    the cache activity.  nthreads is accepted for the reference's signature (the host tensor uses
    every hardware thread).  Keyword only, not in the reference, the real-space tensor the transform
    starts from: by default the FP64 path uses the reference's whole tensor (host, reference
    operation order) and the FP32 paths its non-negative octant extended by parity; host_tensor
    selects the whole host tensor, octant_tensor the extended octant, device_tensor the formulas
    evaluated on the GPU, on either path.  full_tensor (FP32 paths) applies the reference's whole
    tensor spectrum, far-field noise included, as the symmetric quarter plus an FP16 residual:
    FP32 arithmetic on the reference's tensor, for 0.40 GB more tensor bytes at 1024x1024x4."""

    def __init__(self, mesh, verbose=False, caching=False, nthreads=4, *, host_tensor=False, device_tensor=False,
                 octant_tensor=False, full_tensor=False):
        self.mesh = mesh
        self.verbose, self.caching, self.nthreads = bool(verbose), bool(caching), nthreads
        self.host_tensor = bool(host_tensor)
        self.device_tensor = bool(device_tensor)
        self.octant_tensor = bool(octant_tensor)
        self.full_tensor = bool(full_tensor)

    def _desc(self, mesh):
        d = _td(_abi.DEMAG)
        d.flags = ((_abi.TERM_CACHING if self.caching else 0) | (_abi.TERM_VERBOSE if self.verbose else 0)
                   | (_abi.TERM_HOST_TENSOR if self.host_tensor else 0)
                   | (_abi.TERM_DEVICE_TENSOR if self.device_tensor else 0)
                   | (_abi.TERM_OCTANT_TENSOR if self.octant_tensor else 0)
                   | (_abi.TERM_FULL_TENSOR if self.full_tensor else 0))
        return d, []


class ExchangeField(FieldTerm):
    """ExchangeField(A) in J/m (exchange_field.hpp:8-16)."""

    def __init__(self, A, mesh=None, verbose=False):
        self.A = float(A)

    def _desc(self, mesh):
        return _td(_abi.EXCHANGE, self.A), []


class DmiField(FieldTerm):
    """DmiField(D, D_axis=[0, 0, 1]) interface DMI (dmi_field.hpp:8-23, magnumaf.pyx:1324-1344): D in
    J/m^2, a float or a cell-wise field (nx, ny, nz) / (nx, ny, nz, 1) (dmi_field.cpp:17-19,97-109)."""

    def __init__(self, D, D_axis=(0.0, 0.0, 1.0)):
        self.D = np.asarray(D, dtype=np.float64) if _is_array(D) else float(D)
        self.D_axis = list(D_axis)

    def _desc(self, mesh):
        if isinstance(self.D, float):
            return _td(_abi.DMI, self.D, self.D_axis), []
        f = _scalar_flat(self.D, mesh)
        d = _td(_abi.DMI, 0.0, self.D_axis)
        d.array0 = _dp(f)
        return d, [f]


class UniaxialAnisotropyField(FieldTerm):
    """UniaxialAnisotropyField(Ku1, Ku1_axis=[0, 0, 1]) (uniaxial_anisotropy_field.hpp:8-35,
    magnumaf.pyx:1493-1503): Ku1 in J/m^3, a float or a cell-wise field (nx, ny, nz) / (nx, ny, nz, 1);
    Ku1_axis a 3-vector or a cell-wise axis field (nx, ny, nz, 3), normalised per cell
    (uniaxial_anisotropy_field.cpp:11-25,58-89)."""

    def __init__(self, Ku1, Ku1_axis=(0.0, 0.0, 1.0)):
        self.Ku1 = np.asarray(Ku1, dtype=np.float64) if _is_array(Ku1) else float(Ku1)
        if _is_array(Ku1_axis):
            self.Ku1_axis = np.asarray(Ku1_axis, dtype=np.float64)
        else:
            n = math.sqrt(sum(float(v) ** 2 for v in Ku1_axis))
            self.Ku1_axis = [float(v) / n for v in Ku1_axis]

    @property
    def Ku1_field(self):
        return None if isinstance(self.Ku1, float) else self.Ku1

    def _desc(self, mesh):
        keep = []
        scalar_axis = not isinstance(self.Ku1_axis, np.ndarray)
        d = _td(_abi.UNIAXIAL, self.Ku1 if isinstance(self.Ku1, float) else 0.0,
                self.Ku1_axis if scalar_axis else (0.0, 0.0, 1.0))
        if not isinstance(self.Ku1, float):
            k = self.Ku1
            if k.ndim == 4 and k.shape[3] == 3:
                if not (np.array_equal(k[..., 0], k[..., 1]) and np.array_equal(k[..., 0], k[..., 2])):
                    raise ValueError("UniaxialAnisotropyField: a Ku1 field of shape (nx, ny, nz, 3) must hold equal "
                                     "components (the legacy tiled form); pass (nx, ny, nz)")
                k = k[..., 0]
            f = _scalar_flat(k, mesh)
            d.array0 = _dp(f)
            keep.append(f)
        if not scalar_axis:
            a = self.Ku1_axis
            if a.shape[:3] != (mesh.nx, mesh.ny, mesh.nz):
                a = np.broadcast_to(a, (mesh.nx, mesh.ny, mesh.nz, 3))
            f = to_flat(a)
            d.array1 = _dp(f)
            keep.append(f)
        return d, keep


class _CallbackState:
    """What a field callback sees of the State: t, mesh and (for full-field callbacks) m."""

    def __init__(self, t, mesh, m_flat):
        self.t = t
        self.mesh = mesh
        self._m = m_flat

    @property
    def m(self):
        if self._m is None:
            raise AttributeError("homogeneous field callbacks do not receive m")
        return from_flat(self._m, self.mesh.nx, self.mesh.ny, self.mesh.nz)


class ExternalField(FieldTerm):
    """ExternalField(array (nx, ny, nz, 3) in A/m), ExternalField(homogeneous=(hx, hy, hz)), or a
    time-dependent field ExternalField(func) with func(state) -> (nx, ny, nz, 3) array
    (external_field.hpp:11-13, external_field.cpp:28-36: evaluated at every field evaluation, at
    the stage times of the integrator).  With homogeneous_callback=True func(state) returns one
    (hx, hy, hz) and the state it receives carries t and mesh only.  set_homogeneous_field changes
    the field between evaluations (external_field.cpp:21-26)."""

    kind = _abi.ZEEMAN

    def __init__(self, array_in=None, homogeneous=None, homogeneous_callback=False):
        self.callback = None
        self.homogeneous_callback = bool(homogeneous_callback)
        if callable(array_in):
            self.callback = array_in
            array_in = None
        self.array = None if array_in is None else np.asarray(array_in, dtype=np.float64)
        self.homogeneous = None if homogeneous is None else [float(v) for v in homogeneous]

    def set_homogeneous_field(self, x, y, z):
        self.homogeneous = [float(x), float(y), float(z)]
        self.array = None
        self.callback = None

    def _desc(self, mesh):
        if self.callback is not None:
            return self._callback_desc(mesh)
        if self.array is not None:
            a = self.array
            if a.shape[:3] != (mesh.nx, mesh.ny, mesh.nz):
                a = np.broadcast_to(a, (mesh.nx, mesh.ny, mesh.nz, 3))
            flat = to_flat(a)
            d = _td(self.kind)
            d.array0 = _dp(flat)
            return d, [flat]
        return _td(self.kind, field=self.homogeneous or (0, 0, 0)), []

    def _callback_desc(self, mesh):
        func, hom, n3 = self.callback, self.homogeneous_callback, 3 * mesh.n

        def fn(user, t, m_ptr, h_ptr, stream):
            try:
                if hom:
                    v = np.asarray(func(_CallbackState(t, mesh, None)), dtype=np.float64).ravel()
                    if v.size != 3:
                        return 2
                    C.memmove(h_ptr, v.ctypes.data, 24)
                else:
                    m = np.ctypeslib.as_array(C.cast(m_ptr, C.POINTER(C.c_double)), shape=(n3,)).copy()
                    v = np.asarray(func(_CallbackState(t, mesh, m)), dtype=np.float64)
                    if v.ndim == 4 or v.shape[:1] != (n3,):
                        v = to_flat(np.broadcast_to(v, (mesh.nx, mesh.ny, mesh.nz, 3)))
                    C.memmove(h_ptr, np.ascontiguousarray(v).ctypes.data, 8 * n3)
                return 0
            except Exception as e:  # noqa: BLE001 - a failed evaluation, reported through the status
                sys.stderr.write(f"ExternalField callback raised: {e!r}\n")
                return 1

        f = _abi.FIELD_FN(fn)
        d = _td(self.kind)
        d.flags = _abi.TERM_HOMOGENEOUS_CALLBACK if hom else _abi.TERM_HOST_CALLBACK
        d.callback = f
        return d, [f]


class RKKYExchangeField(FieldTerm):
    """RKKYExchangeField(rkky_values, exchange_values, mesh, rkky_indices=None)
    (rkky_exchange_field.hpp:37-55): scalar fields (nx, ny, nz) or (nx, ny, nz, 3)."""

    def __init__(self, rkky_values, exchange_values, mesh, rkky_indices=None, verbose=True):
        self.rk = _scalar_flat(rkky_values, mesh)
        self.A = _scalar_flat(exchange_values, mesh)
        self.idx = None
        if rkky_indices is not None:
            r = np.asarray(rkky_indices)
            if r.ndim == 4:
                r = r[..., 0]
            self.idx = np.ascontiguousarray(r.transpose(2, 1, 0)).ravel().astype(np.uint32)

    def _desc(self, mesh):
        d = _td(_abi.RKKY)
        d.array0 = _dp(self.rk)
        d.array1 = _dp(self.A)
        if self.idx is not None:
            d.index0 = self.idx.ctypes.data_as(C.POINTER(C.c_uint32))
        return d, [self.rk, self.A, self.idx]


class AtomisticDipoleDipoleField(FieldTerm):
    def __init__(self, mesh):
        self.mesh = mesh

    def _desc(self, mesh):
        return _td(_abi.ATOM_DIPOLE), []


class AtomisticExchangeField(FieldTerm):
    def __init__(self, J_atom):
        self.J_atom = float(J_atom)

    def _desc(self, mesh):
        return _td(_abi.ATOM_EXCHANGE, self.J_atom), []


class AtomisticDmiField(FieldTerm):
    """Python default axis [0, 0, -1] (magnumaf.pyx:1637)."""

    def __init__(self, D_atom, D_atom_axis=(0.0, 0.0, -1.0)):
        self.D_atom = float(D_atom)
        self.axis = list(D_atom_axis)

    def _desc(self, mesh):
        return _td(_abi.ATOM_DMI, self.D_atom, self.axis), []


class AtomisticUniaxialAnisotropyField(FieldTerm):
    def __init__(self, K_atom, K_atom_axis=(0.0, 0.0, 1.0)):
        self.K_atom = float(K_atom)
        self.axis = list(K_atom_axis)

    def _desc(self, mesh):
        return _td(_abi.ATOM_UNIAXIAL, self.K_atom, self.axis), []


class AtomisticExternalField(ExternalField):
    kind = _abi.ATOM_ZEEMAN


# ---------------------------------------------------------------- integrator
class LLGIntegrator:
    """LLGIntegrator(alpha, terms, mode="RKF45", hmin=1e-15, hmax=3.5e-10, atol=1e-6, rtol=1e-6,
    dissipation_term_only=False) (magnumaf.pyx:912-1033, llg_integrator.hpp:26-82).  mode is one of
    "RKF45", "DP45", "BS45", "DP78", "BS23".

    alpha is a float or a cell-wise damping field (nx, ny, nz) / (nx, ny, nz, 1)
    (LLGIntegrator(af::array alpha_field, ...), llg_integrator.hpp:32-40, magnumaf.pyx:962-967).

    precision="f32" selects the FP32-field path (demag FFT and stencils in FP32, state and RK in FP64);
    "f32_demag" keeps the stencils in FP64 (FP32 demag only), "f32_stencils" the demag."""

    def __init__(self, alpha, terms, mode="RKF45", hmin=1e-15, hmax=3.5e-10, atol=1e-6, rtol=1e-6,
                 dissipation_term_only=False, precision="f64"):
        self._alpha_field = None
        if _is_array(alpha):
            self._alpha_field = np.asarray(alpha, dtype=np.float64)
            self._alpha = 0.0
        else:
            self._alpha = float(alpha)
        self.terms = list(terms)
        self.mode = mode
        self.hmin, self.hmax, self.atol, self.rtol = hmin, hmax, atol, rtol
        self.dissipation_term_only = dissipation_term_only
        self.precision = precision
        self._llg = None
        self._ctx = None
        self._ctrl = None  # (h, errold, reject) carried across rebuilds
        self.accumulated_steps = 0

    def _bind(self, state):
        ctx = state._context(self.terms, self.precision)
        if self._llg is not None and self._ctx is ctx:
            return
        if self._llg is not None:
            self._ctrl = self.controller_state()
            lib().mafx_llg_destroy(self._llg)
            self._llg = None
        h = C.c_void_p()
        check(lib().mafx_llg_create(ctx.h, self._alpha, self.mode.encode(), self.hmin, self.hmax, self.atol,
                                    self.rtol, int(self.dissipation_term_only), C.byref(h)))
        self._llg, self._ctx = h, ctx
        if self._alpha_field is not None:
            self._upload_alpha_field()
        if self._ctrl is not None:
            check(lib().mafx_llg_set_controller_state(self._llg, *self._ctrl))

    def _upload_alpha_field(self):
        f = _scalar_flat(self._alpha_field, self._ctx.mesh)
        check(lib().mafx_llg_set_alpha_field(self._llg, _dp(f)))

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter exit the process releases the device memory itself
        if getattr(self, "_llg", None) and not _finalizing():
            lib().mafx_llg_destroy(self._llg)
            self._llg = None

    @property
    def alpha(self):
        """The scalar alpha, or the alpha field when the integrator was given one."""
        return self._alpha if self._alpha_field is None else self._alpha_field

    @alpha.setter
    def alpha(self, v):
        if _is_array(v):
            self._alpha_field = np.asarray(v, dtype=np.float64)
            if self._llg is not None:
                self._upload_alpha_field()
            return
        self._alpha = float(v)
        self._alpha_field = None
        if self._llg is not None:
            check(lib().mafx_llg_set_alpha_field(self._llg, None))
            check(lib().mafx_llg_set_alpha(self._llg, self._alpha))

    def controller_state(self):
        h, e, r = C.c_double(), C.c_double(), C.c_int()
        check(lib().mafx_llg_get_controller_state(self._llg, C.byref(h), C.byref(e), C.byref(r)))
        return h.value, e.value, r.value

    def add_terms(self, *args):
        if self._llg is not None:
            self._ctrl = self.controller_state()
            lib().mafx_llg_destroy(self._llg)
            self._llg, self._ctx = None, None
        self.terms.extend(args)

    def step(self, state):
        self._bind(state)
        t = C.c_double(state.t)
        check(lib().mafx_llg_step(self._llg, C.c_void_p(state.m_dev.data_ptr()), C.byref(t)))
        state.t = t.value
        state.steps += 1
        self.accumulated_steps += 1

    def run_until(self, state, t_end, max_steps=1 << 62):
        self._bind(state)
        t = C.c_double(state.t)
        n = C.c_int64()
        check(lib().mafx_llg_run_until(self._llg, C.c_void_p(state.m_dev.data_ptr()), C.byref(t), t_end, max_steps,
                                       C.byref(n)))
        state.t = t.value
        state.steps += n.value
        self.accumulated_steps += n.value
        return n.value

    def relax(self, state, precision=1e-10, ncalcE=100, nprint=1000, verbose=True):
        """relax(state, precision=1e-10, ncalcE=100, nprint=1000) (magnumaf.pyx:993-998,
        llg_integrator.cpp:43-66): steps until the relative energy change over ncalcE steps is
        below precision.  Returns the number of steps taken."""
        self._bind(state)
        t = C.c_double(state.t)
        n = C.c_int64()
        check(lib().mafx_llg_relax(self._llg, C.c_void_p(state.m_dev.data_ptr()), C.byref(t), float(precision),
                                   int(ncalcE), int(nprint), int(bool(verbose)), int(state.steps), C.byref(n)))
        state.t = t.value
        state.steps += n.value
        self.accumulated_steps += n.value
        return n.value

    def stats(self):
        s = _abi.LlgStats()
        check(lib().mafx_llg_get_stats(self._llg, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _abi.LlgStats._fields_}

    def integrate_dense(self, state, time_in_s, write_every_dt_in_s, filename=None, verbose=False, append=False):
        """Returns rows (t, <mx>, <my>, <mz>); also written to `filename` in the reference's
        't<TAB>mx<TAB>my<TAB>mz' line format when given."""
        self._bind(state)
        rows_max = int(time_in_s / write_every_dt_in_s) + 8
        out = np.zeros(4 * rows_max)
        t = C.c_double(state.t)
        nr = C.c_int64()
        check(lib().mafx_llg_integrate_dense(self._llg, C.c_void_p(state.m_dev.data_ptr()), C.byref(t), time_in_s,
                                             write_every_dt_in_s, _dp(out), rows_max, C.byref(nr)))
        state.t = t.value
        rows = out[: 4 * nr.value].reshape(nr.value, 4)
        if filename:
            from . import io
            with open(filename, "a" if append else "w") as f:
                for r in rows:
                    f.write(io.stream_line(r[0], r[1:]) + "\n")
        return rows

    def save_state(self, path, state):
        """Writes a resume file (.npz): m bit for bit, t, steps, this integrator's scheme and
        parameters and its controller state, so that load_state + step continues exactly as the
        uninterrupted run would.  The step counters of stats() are not part of it."""
        self._bind(state)
        h, errold, reject = self.controller_state()
        g = state.mesh
        np.savez(path, m=state.m_dev.cpu().numpy(), t=state.t, steps=state.steps,
                 mesh=np.array([g.nx, g.ny, g.nz], dtype=np.int64), spacing=np.array([g.dx, g.dy, g.dz]),
                 scheme=np.array(self.mode), alpha=self._alpha,
                 params=np.array([self.hmin, self.hmax, self.atol, self.rtol]),
                 dissipation_term_only=int(self.dissipation_term_only), h=h, errold=errold, reject=reject,
                 accumulated_steps=self.accumulated_steps)

    def load_state(self, path, state):
        """Restores state (m, t, steps) and this integrator's controller state from save_state's
        file.  The mesh, scheme and controller parameters must match."""
        with np.load(path, allow_pickle=False) as f:
            g = state.mesh
            if tuple(f["mesh"]) != (g.nx, g.ny, g.nz) or not np.array_equal(f["spacing"], [g.dx, g.dy, g.dz]):
                raise ValueError("load_state: the file was written for another mesh")
            if str(f["scheme"]) != self.mode or not np.array_equal(f["params"], [self.hmin, self.hmax, self.atol,
                                                                                 self.rtol]):
                raise ValueError("load_state: the file was written by an integrator with another scheme or controller")
            state.m_dev.copy_(_torch().from_numpy(np.ascontiguousarray(f["m"])))
            state.t = float(f["t"])
            state.steps = int(f["steps"])
            self._bind(state)
            check(lib().mafx_llg_set_controller_state(self._llg, float(f["h"]), float(f["errold"]), int(f["reject"])))
            self.accumulated_steps = int(f["accumulated_steps"])

    def H_in_Apm(self, state):
        ctx = state._context(self.terms, self.precision)
        h = ctx.heff(state.m_dev)
        return from_flat(h.cpu().numpy(), state.mesh.nx, state.mesh.ny, state.mesh.nz)

    def Eeff_in_J(self, state):
        return state._context(self.terms, self.precision).energy(state.m_dev)[0]

    E = Eeff_in_J


class LBFGS_Minimizer:
    """LBFGS_Minimizer(terms=[], tol=1e-6, maxiter=230, verbose=0) (magnumaf.pyx:1786-1836,
    lbfgs_minimizer.hpp:9-24).  torque_tol > 0 adds the stop max|m x H| / max|H| <= torque_tol.

    precision="f32" (or "f32_demag") minimises on the FP32-field path.  Its energy sums carry the
    FP32 field's rounding, which near the minimum exceeds the energy change of an iteration; with
    torque_tol > 0 a run that stalls there (five iterations without a decrease) goes on from the same
    iterate on the FP64 path, keeping its curvature pairs (fp64_finish, default True;
    mafx_lbfgs_set_finish_context).  last_stats then splits the iterations into "fp32_iterations"
    and "fp64_iterations".
    convergence_file, when set, receives the reference's of_convergence_ rows
    (lbfgs_minimizer.cpp:148-153) of every later minimisation."""

    def __init__(self, terms=(), tol=1e-6, maxiter=230, verbose=0, torque_tol=0.0, precision="f64",
                 fp64_finish=True):
        self.terms = list(terms)
        self.tol, self.maxiter, self.verbose = tol, maxiter, verbose
        self.torque_tol = torque_tol
        self.precision = precision
        self.fp64_finish = fp64_finish
        self.convergence_file = None
        self.last_stats = None

    def add_terms(self, *args):
        self.terms.extend(args)

    def delete_last_term(self):
        self.terms.pop()

    def minimize(self, state):
        ctx = state._context(self.terms, self.precision)
        finish = None
        if self.precision != "f64" and self.fp64_finish and self.torque_tol > 0:
            finish = state._context(self.terms, "f64")
        path = None if self.convergence_file is None else str(self.convergence_file).encode()
        check(lib().mafx_lbfgs_set_convergence_file(ctx.h, path))
        check(lib().mafx_lbfgs_set_finish_context(ctx.h, None if finish is None else finish.h))
        E = C.c_double()
        st = _abi.LbfgsStats()
        try:
            check(lib().mafx_lbfgs_minimize(ctx.h, C.c_void_p(state.m_dev.data_ptr()), self.tol, self.maxiter,
                                            self.verbose, self.torque_tol, C.byref(E), C.byref(st)))
        finally:
            check(lib().mafx_lbfgs_set_convergence_file(ctx.h, None))
            check(lib().mafx_lbfgs_set_finish_context(ctx.h, None))
        st = {k: getattr(st, k) for k, _ in _abi.LbfgsStats._fields_}
        if finish is not None:
            st["fp64_iterations"] = st["finish_iterations"]
            st["fp32_iterations"] = st["iterations"] - st["finish_iterations"]
        self.last_stats = st
        return E.value
