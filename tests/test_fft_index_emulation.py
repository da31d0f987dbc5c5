"""CPU emulation of the field kernel's transform (numericalflowiteration_amd/csrc/
field.hip: fft_axis / fft3): the same batched radix-2 Stockham index arithmetic
and buffer ping-pong, and the direct-DFT path for other lengths, checked
against numpy's FFT.  Also checks every index stays inside the buffers."""
import numpy as np
import pytest


def line_base(mode, line, Nx, Ny):
    if mode == 0:
        return line * Nx
    if mode == 1:
        return (line // Nx) * Nx * Ny + (line % Nx)
    return line


def log2_or_none(n):
    return n.bit_length() - 1 if n & (n - 1) == 0 else None


def fft_axis(inp, out, L, S, nlines, mode, Nx, Ny, inverse):
    N = inp.size
    tw = np.exp(-2j * np.pi * np.arange(L) / L)
    if L <= 1:
        return inp, out
    lg = log2_or_none(L)
    if lg is None:
        for e in range(nlines * L):
            line, k = divmod(e, L)
            base = line_base(mode, line, Nx, Ny)
            acc = 0j
            m = 0
            for j in range(L):
                w = tw[m].conjugate() if inverse else tw[m]
                idx = base + j * S
                assert 0 <= idx < N
                acc += inp[idx] * w
                m = (m + k) % L
            out[base + k * S] = acc
        return out, inp
    half = L >> 1
    lh = lg - 1
    Ns, log2Ns = 1, 0
    while Ns < L:
        twstep = L >> (log2Ns + 1)
        for e in range(nlines * half):
            line, j = e >> lh, e & (half - 1)
            base = line_base(mode, line, Nx, Ny)
            i0, i1 = base + j * S, base + (j + half) * S
            assert 0 <= i0 < N and 0 <= i1 < N
            x0, x1 = inp[i0], inp[i1]
            k = j & (Ns - 1)
            assert k * twstep < L
            w = tw[k * twstep]
            if inverse:
                w = w.conjugate()
            x1 = x1 * w
            o = ((j >> log2Ns) << (log2Ns + 1)) + k
            assert 0 <= base + (o + Ns) * S < N
            out[base + o * S] = x0 + x1
            out[base + (o + Ns) * S] = x0 - x1
        inp, out = out, inp
        Ns <<= 1
        log2Ns += 1
    return inp, out


def fft3(A, B, Nx, Ny, Nz, inverse):
    N = Nx * Ny * Nz
    r, o = fft_axis(A, B, Nx, 1, N // Nx, 0, Nx, Ny, inverse)
    r, o = fft_axis(r, o, Ny, Nx, N // Ny, 1, Nx, Ny, inverse)
    r, o = fft_axis(r, o, Nz, Nx * Ny, N // Nz, 2, Nx, Ny, inverse)
    return r


@pytest.mark.parametrize("shape", [(16, 1, 1), (64, 1, 1), (24, 1, 1), (8, 4, 2), (4, 4, 4), (8, 6, 4)])
def test_emulated_transform_matches_numpy(shape):
    Nx, Ny, Nz = shape
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(Nx * Ny * Nz) + 0j
    A, B = x.copy(), np.zeros_like(x)
    got = fft3(A, B, Nx, Ny, Nz, False)
    ref = np.fft.fftn(x.reshape(Nz, Ny, Nx)).ravel()
    assert np.allclose(got, ref, atol=1e-12)
    A2, B2 = got.copy(), np.zeros_like(got)
    back = fft3(A2, B2, Nx, Ny, Nz, True) / x.size
    assert np.allclose(back, x, atol=1e-13)
