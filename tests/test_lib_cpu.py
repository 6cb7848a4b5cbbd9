"""CPU-side checks of libmafx: the C-ABI loads and exports every symbol include/mafx.h declares,
and the host Newell tensor setup reproduces the oracle (the reference's operation order) bit for
bit.  No kernel is launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    from magnumaf import _abi
    return _abi.lib()


def test_header_symbols_exported():
    from magnumaf import _abi
    hdr = open(os.path.join(ROOT, "include", "mafx.h")).read()
    declared = sorted(set(re.findall(r"\b(mafx_[A-Za-z0-9_]+)\s*\(", hdr)))
    L = _lib()
    for name in declared:
        assert hasattr(L, name), f"{name} declared in mafx.h but not exported"
    assert sorted(_abi.EXPORTS) == declared


def test_null_arguments_are_errors():
    from magnumaf import _abi
    L = _lib()
    rc = L.mafx_demag_calc_N(None, None, 1)
    assert rc == 1
    assert b"null" in L.mafx_last_error()
    m = _abi.Mesh(0, 1, 1, 1e-9, 1e-9, 1e-9)
    out = np.zeros(8)
    assert L.mafx_demag_calc_N(C.byref(m), out.ctypes.data_as(C.POINTER(C.c_double)), 1) == 1


@pytest.mark.parametrize("mesh", [(5, 4, 3, 1e-9, 1.2e-9, 0.8e-9), (7, 3, 1, 2e-9, 1e-9, 3e-9),
                                  (16, 8, 2, 2e-9, 2e-9, 2e-9), (40, 1, 1, 1e-9, 3e-9, 2e-9)])
def test_host_newell_tensor_bitwise_equals_oracle(mesh):
    from magnumaf import _abi
    nx, ny, nz = mesh[:3]
    n0, n1, n2 = 2 * nx, 2 * ny, (1 if nz == 1 else 2 * nz)
    out = np.empty(6 * n0 * n1 * n2)
    m = _abi.Mesh(*mesh)
    assert _lib().mafx_demag_calc_N(C.byref(m), out.ctypes.data_as(C.POINTER(C.c_double)), 3) == 0
    ref = O.demag_tensor(mesh).ravel()
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))


def test_context_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from magnumaf import _abi
    L = _lib()
    m = _abi.Mesh(4, 4, 1, 1e-9, 1e-9, 1e-9)
    t = (_abi.TermDesc * 1)()
    t[0].kind = _abi.EXCHANGE
    t[0].value = 1e-11
    h = C.c_void_p()
    rc = L.mafx_ctx_create(C.byref(m), 8e5, None, t, 1, 0, C.byref(h))
    assert rc != 0
    assert b"GPU" in L.mafx_last_error() or b"device" in L.mafx_last_error()


@pytest.mark.parametrize("mesh,prec,flags,dipole", [((128, 32, 1, 4e-9, 4e-9, 3e-9), 0, 0, 0),
                                                    ((1024, 1024, 4, 2e-9, 2e-9, 2e-9), 1, 0, 0),
                                                    ((1024, 1024, 4, 2e-9, 2e-9, 2e-9), 1, 4, 0),
                                                    ((1024, 1024, 4, 2e-9, 2e-9, 2e-9), 1, 32, 0),
                                                    ((1024, 1024, 4, 2e-9, 2e-9, 2e-9), 1, 128, 0),
                                                    ((256, 64, 2, 2e-9, 2e-9, 2e-9), 1, 128 | 32, 0),
                                                    ((64, 32, 2, 2e-9, 2e-9, 2e-9), 0, 32, 0),
                                                    ((512, 512, 1, 2.7e-10, 2.7e-10, 2.7e-10), 0, 0, 1)])
def test_demag_cache_name_is_the_reference_mesh_string(mesh, prec, flags, dipole):
    """The cache file name starts with the reference's to_string(mesh) (demag_field.cpp:185-190):
    expanded lengths and 1e9 * d printed by std::to_string (six decimals); the layout and the exact
    bits of dx, dy, dz follow, so meshes that print alike never share a file."""
    import struct
    from magnumaf import _abi
    nx, ny, nz, dx, dy, dz = mesh
    ref = (f"Nfft_n0exp_{2 * nx}_n1exp_{2 * ny}_n2exp_{1 if nz == 1 else 2 * nz}"
           f"_dx_{1e9 * dx:f}_dy_{1e9 * dy:f}_dz_{1e9 * dz:f}")
    buf = C.create_string_buffer(512)
    m = _abi.Mesh(*mesh)
    assert _lib().mafx_demag_cache_name(C.byref(m), prec, flags, dipole, buf, len(buf)) == 0
    name = buf.value.decode()
    # the layouts built from other real-space tensors (the whole host tensor on the FP32 path, whose
    # default is the host octant; the GPU-built one on either path) have names of their own
    layout = {(0, 0): "_mafx_c64_", (0, 32): "_mafx_c64_dev_", (1, 0): "_mafx_f32oct_", (1, 4): "_mafx_f32sym_host_",
              (1, 32): "_mafx_f32sym_dev_", (1, 128): "_mafx_f32res_host_", (1, 160): "_mafx_f32res_dev_"}
    assert name.startswith(ref + layout[(prec, flags)].rstrip("_"))
    assert (layout[(prec, flags)] in name) or (dipole and layout[(prec, flags)].rstrip("_") + "_dipole_" in name)
    bits = "".join(f"{struct.unpack('<Q', struct.pack('<d', v))[0]:016x}" for v in (dx, dy, dz))
    assert name.endswith(("_dipole_" if dipole else "_") + bits)
    other = _abi.Mesh(nx, ny, nz, dx * (1 + 1e-12), dy, dz)  # prints the same, differs in its bits
    buf2 = C.create_string_buffer(512)
    assert _lib().mafx_demag_cache_name(C.byref(other), prec, flags, dipole, buf2, len(buf2)) == 0
    assert buf2.value.decode().startswith(ref) and buf2.value != buf.value
    assert _lib().mafx_demag_cache_name(C.byref(m), prec, flags, dipole, buf, 8) == 1  # too small: error
