"""CPU emulation of the field kernel's transform (numericalflowiteration_amd/csrc/
fft_lds.hpp: fft_axis, field.hip: fft3): the same batched Stockham index
arithmetic (a radix-2 stage when log2 L is odd, then radix-4 stages) and
buffer ping-pong, and the direct-DFT path for other lengths, checked against
numpy's FFT.  Also checks every index stays inside the buffers and every
twiddle read inside the table of L/2 entries."""
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
    table = tw[: L // 2]                    # the device table: exp(-2 pi i m / L), m < L/2

    def twid(m):                            # exp(-2 pi i m / L) for m < L from the half table
        assert 0 <= m < L
        w = table[m] if m < L // 2 else -table[m - L // 2]
        return w.conjugate() if inverse else w

    Ns, log2Ns = 1, 0
    if lg & 1:                              # one radix-2 stage
        half = L >> 1
        for e in range(nlines * half):
            line, j = e >> (lg - 1), e & (half - 1)
            base = line_base(mode, line, Nx, Ny)
            i0, i1 = base + j * S, base + (j + half) * S
            assert 0 <= i0 < N and 0 <= i1 < N
            x0, x1 = inp[i0], inp[i1]          # Ns = 1: k = 0, no twiddle
            out[base + 2 * j * S] = x0 + x1
            out[base + (2 * j + 1) * S] = x0 - x1
        inp, out = out, inp
        Ns, log2Ns = 2, 1
    quarter = L >> 2
    while Ns < L:
        twstep = L >> (log2Ns + 2)          # L / (4 Ns)
        for e in range(nlines * quarter):
            line, j = e >> (lg - 2), e & (quarter - 1)
            base = line_base(mode, line, Nx, Ny)
            idx = [base + (j + r * quarter) * S for r in range(4)]
            assert all(0 <= i < N for i in idx)
            k = j & (Ns - 1)
            b = [inp[idx[0]]] + [inp[idx[r]] * twid(r * k * twstep) for r in (1, 2, 3)]
            s02, d02 = b[0] + b[2], b[0] - b[2]
            s13, d13 = b[1] + b[3], b[1] - b[3]
            rot = 1j * d13 if inverse else -1j * d13        # -i (b1 - b3) forward
            y = [s02 + s13, d02 + rot, s02 - s13, d02 - rot]
            o = ((j >> log2Ns) << (log2Ns + 2)) + k
            for q in range(4):
                t = base + (o + q * Ns) * S
                assert 0 <= t < N
                out[t] = y[q]
        inp, out = out, inp
        Ns <<= 2
        log2Ns += 2
    return inp, out


def fft3(A, B, Nx, Ny, Nz, inverse):
    N = Nx * Ny * Nz
    r, o = fft_axis(A, B, Nx, 1, N // Nx, 0, Nx, Ny, inverse)
    r, o = fft_axis(r, o, Ny, Nx, N // Ny, 1, Nx, Ny, inverse)
    r, o = fft_axis(r, o, Nz, Nx * Ny, N // Nz, 2, Nx, Ny, inverse)
    return r


@pytest.mark.parametrize("shape", [(16, 1, 1), (64, 1, 1), (24, 1, 1), (8, 4, 2), (4, 4, 4), (8, 6, 4),
                                   (2, 1, 1), (32, 1, 1), (128, 1, 1), (256, 1, 1), (512, 1, 1), (16, 16, 16),
                                   (2, 8, 32)])
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
