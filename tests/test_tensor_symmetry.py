"""The FP32 field path stores the demag tensor spectrum as a real quarter (ky <= n1/2, kz <= n2/2).

Pins the two facts that design rests on, with the oracle's reference-order Newell tensor
(demag_field.cpp:104-205) and its own H_demag as the checkers:
  * zeroing the unpaired planes x = nx, y = ny, z = nz of the padded tensor (pair distance -n,
    which no two cells of the mesh have) and keeping only the real part of the spectrum
    reproduces the cropped field up to the reference tensor's own rounding asymmetry
    (far-field Newell cancellation noise, <= 1e-9 normwise here; the FP32 path's own rounding is
    ~1e-7, and the FP64 path keeps the full complex spectrum);
  * the real spectrum is even in ky and kz for xx, yy, zz and odd along both axes of each
    off-diagonal component (xy: y; xz: z; yz: y and z), which is what lets one thread serve the
    columns ky and n1 - ky from one tensor read.
"""
import numpy as np
import pytest

import oracle_lib as O

SY = np.array([1, -1, 1, 1, -1, 1])
SZ = np.array([1, 1, -1, 1, -1, 1])


def _conv(Nt, m, Ms, shape, real_only):
    nx, ny, nz = shape
    n2, n1, n0 = Nt.shape[1:]
    Nk = np.fft.rfftn(Nt, axes=(1, 2, 3))
    if real_only:
        Nk = Nk.real.astype(complex)
    Mr = np.zeros((3, n2, n1, n0))
    for c in range(3):
        Mr[c, :nz, :ny, :nx] = (Ms * m[..., c]).transpose(2, 1, 0)
    Mk = np.fft.rfftn(Mr, axes=(1, 2, 3))
    xx, xy, xz, yy, yz, zz = Nk
    Hk = [xx * Mk[0] + xy * Mk[1] + xz * Mk[2], xy * Mk[0] + yy * Mk[1] + yz * Mk[2],
          xz * Mk[0] + yz * Mk[1] + zz * Mk[2]]
    return np.stack([np.fft.irfftn(h, s=(n2, n1, n0), axes=(0, 1, 2))[:nz, :ny, :nx].transpose(2, 1, 0)
                     for h in Hk], -1)


@pytest.mark.parametrize("mesh", [(32, 16, 4, 2e-9, 2e-9, 2e-9), (30, 18, 5, 2e-9, 3e-9, 1.5e-9),
                                  (48, 20, 1, 1e-9, 1e-9, 1e-9), (9, 7, 3, 1e-9, 2e-9, 1e-9)])
def test_symmetric_real_spectrum_reproduces_demag(mesh):
    nx, ny, nz = mesh[:3]
    rng = np.random.default_rng(3)
    m = rng.standard_normal((nx, ny, nz, 3))
    m /= np.linalg.norm(m, axis=-1, keepdims=True)
    Ms = 8e5
    H_ref = O.Sim(mesh, Ms, m, [O.term(O.DEMAG)]).heff()
    N = O.demag_tensor(mesh)
    n2, n1, n0 = N.shape[1:]
    nr = lambda a: np.linalg.norm(a.ravel())
    # the full complex spectrum is the reference algorithm itself
    assert nr(_conv(N, m, Ms, (nx, ny, nz), False) - H_ref) <= 1e-13 * nr(H_ref)
    Ns = N.copy()
    Ns[:, :, :, n0 // 2] = 0
    Ns[:, :, n1 // 2, :] = 0
    if n2 > 1:
        Ns[:, n2 // 2, :, :] = 0
    err = nr(_conv(Ns, m, Ms, (nx, ny, nz), True) - H_ref) / nr(H_ref)
    assert err <= 1e-9, err
    Nk = np.fft.rfftn(Ns, axes=(1, 2, 3))
    scale = np.abs(Nk.real).max()
    assert np.abs(Nk.imag).max() <= 1e-8 * scale
    R = Nk.real
    ky = (-np.arange(n1)) % n1
    kz = (-np.arange(n2)) % n2
    for c in range(6):
        assert np.abs(R[c][:, ky, :] - SY[c] * R[c]).max() <= 1e-8 * scale
        assert np.abs(R[c][kz, :, :] - SZ[c] * R[c]).max() <= 1e-8 * scale


def test_unpaired_planes_matter_only_to_the_uncropped_field():
    # a plane of garbage at the unpaired index changes nothing in the cropped convolution
    mesh = (12, 10, 3, 2e-9, 2e-9, 2e-9)
    nx, ny, nz = mesh[:3]
    rng = np.random.default_rng(5)
    m = rng.standard_normal((nx, ny, nz, 3))
    N = O.demag_tensor(mesh)
    N2 = N.copy()
    N2[:, :, :, nx] = rng.standard_normal(N2[:, :, :, nx].shape)
    N2[:, :, ny, :] = rng.standard_normal(N2[:, :, ny, :].shape)
    N2[:, nz, :, :] = rng.standard_normal(N2[:, nz, :, :].shape)
    a = _conv(N, m, 1.0, (nx, ny, nz), False)
    b = _conv(N2, m, 1.0, (nx, ny, nz), False)
    assert np.linalg.norm((a - b).ravel()) <= 1e-13 * np.linalg.norm(a.ravel())
