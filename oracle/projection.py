"""Rotated-scheme Galerkin projection Ghat and its application (oracle, FP64).

Restates ``SmallStrainFFT._get_projection_operator`` and
``SmallStrainFFT._apply_projection_tensor``
(``materialite/models/small_strain_fft/small_strain_fft.py:301-368, 376-382``).

Two builds of Ghat are kept:

* ``projection_literal`` follows the reference step by step (complex rotated
  frequencies, the 7-index A and B products, Nyquist zeroing, Mandel conversion,
  Ghat[0] = diag(mask)).  Its temporaries are N*81 complex numbers, so it is only
  used on small grids, to pin the closed form.
* ``projection_closed_form`` uses the fact that the normalised rotated frequency
  is a unit-modulus complex scalar times the real unit vector n = t/|t|,
  t_j = tan(pi q_j), so Ghat:eps = a(x)n + n(x)a - (n.a) n(x)n with a = eps.n.
"""

import numpy as np

from .tensors import mandel4_from_cartesian, to_mandel, from_mandel


def _freq_grid(dims):
    qs = [np.fft.fftfreq(int(n)) for n in dims]
    return np.array(np.meshgrid(*qs, indexing="ij"))


def _apply_mask_dc(ghat, stress_mask):
    mask = np.asarray(stress_mask, float)
    if not np.array_equal(mask, np.zeros(6)):
        ghat[0, 0, 0] = np.diag((mask != 0).astype(float))
    return ghat


def projection_literal(dims, stress_mask):
    """Literal restatement of small_strain_fft.py:301-368 -> (Nx,Ny,Nz,6,6)."""
    nx, ny, nz = (int(d) for d in dims)
    f = _freq_grid(dims)
    shift = lambda x: 1 + np.exp(2 * np.pi * 1j * x)
    const = shift(f[0]) * shift(f[1]) * shift(f[2])
    g = lambda x: 1j / 4 * np.tan(np.pi * x) * const
    rot = np.array([g(f[0]), g(f[1]), g(f[2])])
    norms = np.linalg.norm(rot, axis=0)
    q = np.divide(rot, norms, out=np.zeros_like(rot), where=norms != 0)
    a = np.real(np.einsum("im,jxyz,lxyz->xyzijlm", np.eye(3), q, np.conj(q)))
    b = np.real(np.einsum("ixyz,jxyz,lxyz,mxyz->xyzijlm", q, q, np.conj(q), np.conj(q)))
    g4 = (
        0.5
        * (
            a
            + np.einsum("xyzijlm->xyzijml", a)
            + np.einsum("xyzijlm->xyzjilm", a)
            + np.einsum("xyzijlm->xyzjiml", a)
        )
        - b
    )
    if nx % 2 == 0:
        g4[nx // 2] = 0.0
    if ny % 2 == 0:
        g4[:, ny // 2] = 0.0
    if nz % 2 == 0:
        g4[:, :, nz // 2] = 0.0
    ghat = mandel4_from_cartesian(g4)
    return _apply_mask_dc(ghat, stress_mask)


def unit_frequency(dims):
    """n(k) = t/|t| with t_j = tan(pi q_j); zero at k=0 and on even-N Nyquist planes."""
    f = _freq_grid(dims)
    t = np.tan(np.pi * f)
    nyquist = np.zeros(f.shape[1:], bool)
    for ax, n in enumerate(dims):
        if int(n) % 2 == 0:
            sl = [slice(None)] * 3
            sl[ax] = int(n) // 2
            nyquist[tuple(sl)] = True
    t[:, nyquist] = 0.0
    norm = np.sqrt(np.sum(t * t, axis=0))
    n = np.divide(t, norm, out=np.zeros_like(t), where=norm != 0)
    return np.moveaxis(n, 0, -1)


def projection_closed_form(dims, stress_mask):
    """Ghat from the closed form -> (Nx,Ny,Nz,6,6); equals projection_literal."""
    n = unit_frequency(dims)
    eye = np.eye(3)
    a = np.einsum("im,...j,...l->...ijlm", eye, n, n)
    g4 = (
        0.5
        * (
            a
            + np.swapaxes(a, -1, -2)
            + np.swapaxes(a, -3, -4)
            + np.swapaxes(np.swapaxes(a, -1, -2), -3, -4)
        )
        - np.einsum("...i,...j,...l,...m->...ijlm", n, n, n, n)
    )
    ghat = mandel4_from_cartesian(g4)
    return _apply_mask_dc(ghat, stress_mask)


def apply_closed_form_spectrum(spec, dims, stress_mask):
    """Ghat:X on a spectrum (Nx,Ny,Nz,6) without materialising the 6x6 table."""
    n = unit_frequency(dims)
    e = from_mandel(spec)  # (...,3,3) complex
    a = np.einsum("...ij,...j->...i", e, n)
    na = np.einsum("...i,...i->...", n, a)
    s = (
        a[..., :, None] * n[..., None, :]
        + n[..., :, None] * a[..., None, :]
        - na[..., None, None] * n[..., :, None] * n[..., None, :]
    )
    out = to_mandel(s)
    mask = np.asarray(stress_mask, float)
    out[0, 0, 0] = (mask != 0) * spec[0, 0, 0]
    return out


def apply_projection(field, dims, ghat=None, stress_mask=None):
    """F^-1[Ghat . F(field)].real for a (N,6) field  (small_strain_fft.py:376-382)."""
    grid = np.asarray(field, float).reshape(*[int(d) for d in dims], 6)
    spec = np.fft.fftn(grid, axes=(0, 1, 2))
    if ghat is not None:
        prod = np.einsum("xyzij,xyzj->xyzi", ghat, spec)
    else:
        prod = apply_closed_form_spectrum(spec, dims, stress_mask)
    return np.fft.ifftn(prod, axes=(0, 1, 2)).real.reshape(-1, 6)
