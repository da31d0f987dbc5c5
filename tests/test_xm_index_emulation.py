"""CPU emulation of the x-major node layout of the 1D fast and FP32C rho
kernels (numericalflowiteration_amd/csrc/trace1d.hip: make_tile, make_xm_node,
xm_reduce; trace1d_xm_setup / trace1d_grid; the LDS size of
trace1d_lds_bytes) with the same integer arithmetic: every node of a q range
is traced by exactly one lane, the LDS f buffer and the partials stay inside
their allocations, every partial of the x range is written exactly once per
launch, and each partial sums exactly the nodes of its velocity group.  Also
the kick-polynomial table indices of the trace step stay inside one slice."""
import itertools

import numpy as np
import pytest

WAVE = 64
XM_GROUP = 16


def xm_setup(Nx, Nu, ix_count):
    xw = Nx if Nx < 64 else 64
    vpw = WAVE // xw
    n_xt = (ix_count + xw - 1) // xw
    ppx = (Nu + XM_GROUP - 1) // XM_GROUP
    return xw, vpw, n_xt, ppx


def shape_ok(block, npt):
    return (block in (64, 128, 256, 512, 1024) and npt in (1, 2, 4) and ((block // WAVE) * npt) % XM_GROUP == 0
            and not (block == 1024 and npt == 4))           # would exceed 128 VGPRs


def emulate(Nx, Nu, q_begin, q_end, block, npt):
    ix_first, ix_last = q_begin // Nu, (q_end - 1) // Nu
    ix_count = ix_last - ix_first + 1
    xw, vpw, n_xt, ppx = xm_setup(Nx, Nu, ix_count)
    WPB = block // WAVE
    vbw = npt * WPB * vpw
    grid = n_xt * ((Nu + vbw - 1) // vbw)
    fbuf_cap = block * npt                     # doubles (trace1d_lds_bytes lower bound)
    traced = np.zeros((Nx, Nu), dtype=np.int64)
    written = np.zeros((Nx, ppx), dtype=np.int64)
    sums = np.zeros((Nx, ppx))
    fval = np.arange(Nx * Nu, dtype=np.float64).reshape(Nx, Nu) + 1.0   # f of node (ix, iu)
    for wg in range(grid):
        xt, vb = wg % n_xt, wg // n_xt
        fbuf = np.full(vbw * xw, np.nan)
        for tid in range(block):
            lane, wave = tid % WAVE, tid // WAVE
            xo, vr = lane % xw, lane // xw
            lane_ok = vr < vpw and xt * xw + xo < ix_count
            for r in range(npt):
                vi = (r * WPB + wave) * vpw + vr
                ix = ix_first + xt * xw + xo if lane_ok else ix_first
                iu = vb * (npt * WPB * vpw) + vi
                q = ix * Nu + iu
                active = lane_ok and iu < Nu and q_begin <= q < q_end
                assert 0 <= ix < Nx
                if active:
                    traced[ix, iu] += 1
                if vr < vpw:
                    k = vi * xw + xo
                    assert 0 <= k < vbw * xw <= fbuf_cap
                    fbuf[k] = fval[ix, iu] if active else 0.0
        groups = vbw // XM_GROUP
        assert not np.isnan(fbuf).any(), "an f slot of the workgroup was not written"
        for k in range(groups * xw):
            xo, gi = k % xw, k // xw
            xl, gg = xt * xw + xo, vb * groups + gi
            if xl >= ix_count or gg >= ppx:
                continue
            col = [fbuf[(gi * XM_GROUP + v) * xw + xo] for v in range(XM_GROUP)]
            ix = ix_first + xl
            assert gg * Nx + ix < ppx * Nx             # partials[group*Nx + x]
            written[ix, gg] += 1
            sums[ix, gg] = sum(col)
    # every node of the range traced exactly once, nothing else
    q = np.arange(Nx * Nu).reshape(Nx, Nu)
    want = ((q >= q_begin) & (q < q_end)).astype(np.int64)
    assert np.array_equal(traced, want)
    # every partial of the x range written exactly once, holding its group's sum
    assert np.all(written[ix_first:ix_last + 1] == 1)
    assert np.all(written[:ix_first] == 0) and np.all(written[ix_last + 1:] == 0)
    for ix in range(ix_first, ix_last + 1):
        for g in range(ppx):
            iu = np.arange(g * XM_GROUP, min((g + 1) * XM_GROUP, Nu))
            m = want[ix, iu].astype(bool)
            assert sums[ix, g] == fval[ix, iu][m].sum()
    return grid


SHAPES = [(b, n) for b, n in itertools.product((64, 128, 256, 512, 1024), (1, 2, 4)) if shape_ok(b, n)]


def test_shape_rule():
    assert set(SHAPES) == {(256, 4), (512, 2), (512, 4), (1024, 1), (1024, 2)}


@pytest.mark.parametrize("Nx,Nu", [(256, 1024), (64, 256), (128, 512), (32, 128), (48, 200), (16, 32),
                                   (100, 64), (20, 48), (8, 16)])
def test_full_range(Nx, Nu):
    for block, npt in [(1024, 1), (256, 4), (512, 2)]:
        if Nx * Nu > 70000 and (block, npt) != (1024, 1):
            continue
        emulate(Nx, Nu, 0, Nx * Nu, block, npt)


@pytest.mark.parametrize("Nx,Nu", [(32, 128), (48, 200), (100, 64)])
def test_sub_ranges(Nx, Nu):
    N = Nx * Nu
    rng = np.random.default_rng(7)
    cuts = [(0, 1), (N - 1, N), (5, 5 + Nu), (Nu - 3, 2 * Nu + 5), (N // 3, N // 2)]
    cuts += [tuple(sorted(rng.integers(0, N, 2))) for _ in range(3)]
    for a, b in cuts:
        if a == b:
            continue
        emulate(Nx, Nu, int(a), int(b), 256, 4)


def test_table_indices_inside_a_slice():
    """poly_step reads the pair at 2(i+1), 2(i+1)+1 and Q2 at q2 + (i+1) with
    i in [-1, Nx] (exact wrap) or [0, Nx-1] (power-of-two mask)."""
    for Nx in (8, 16, 20, 48, 64, 100, 256, 1000):
        q2 = 2 * (Nx + 2)
        sq = (3 * (Nx + 2) + 3) & ~3
        assert sq % 4 == 0
        for i in range(-1, Nx + 1):
            assert 2 * (i + 1) + 1 < q2
            assert q2 + (i + 1) < sq


# ---- 3D x-major layout (trace3d.hip: make_xm3, xm3_velocity; launch_trace3d;
# xm3_vgroup in internal.hpp) ----------------------------------------------------

def xm3_vgroup(n_space, n_vel):
    g = 256
    tiles = (n_space + 63) // 64
    while g > 16 and tiles * ((n_vel + g - 1) // g) < 8192:
        g //= 2
    return g


def emulate3(n_space, n_vel, q_begin, q_end, block, npt):
    l_first, l_last = q_begin // n_vel, (q_end - 1) // n_vel
    n_count = l_last - l_first + 1
    vgroup = xm3_vgroup(n_space, n_vel)
    ppx = (n_vel + vgroup - 1) // vgroup
    assert vgroup % npt == 0
    wpb = block // WAVE
    n_st = (n_count + WAVE - 1) // WAVE
    grid = n_st * ((ppx + wpb - 1) // wpb)
    traced = np.zeros((n_space, n_vel), dtype=np.int64)
    written = np.zeros((n_space, ppx), dtype=np.int64)
    for b in range(grid):
        for tid in range(block):
            lane, wave = tid % WAVE, tid // WAVE
            ll = (b % n_st) * WAVE + lane
            g = (b // n_st) * wpb + wave
            ok = ll < n_count and g < ppx
            l = l_first + (ll if ll < n_count else 0)
            assert 0 <= l < n_space
            for pss in range(0, vgroup, npt):
                for r in range(npt):
                    iv = g * vgroup + pss + r
                    q = l * n_vel + iv
                    if ok and iv < n_vel and q_begin <= q < q_end:
                        traced[l, iv] += 1
            if ok:
                assert g * n_space + l < ppx * n_space  # partials[group*n_space + l]
                written[l, g] += 1
    q = np.arange(n_space * n_vel).reshape(n_space, n_vel)
    want = ((q >= q_begin) & (q < q_end)).astype(np.int64)
    assert np.array_equal(traced, want)
    assert np.all(written[l_first:l_last + 1] == 1)
    assert np.all(written[:l_first] == 0) and np.all(written[l_last + 1:] == 0)


def test_xm3_vgroup_rule():
    assert xm3_vgroup(16 ** 3, 32 ** 3) == 256          # 64 tiles x 128 groups = 8192 waves
    assert xm3_vgroup(8 ** 3, 8 ** 3) == 16
    assert xm3_vgroup(64, 100) == 16


@pytest.mark.parametrize("n_space,n_vel", [(512, 512), (64, 100), (100, 64), (27, 343)])
def test_xm3_full_and_sub_ranges(n_space, n_vel):
    N = n_space * n_vel
    for block, npt in [(256, 1), (256, 4), (512, 2), (512, 4)]:
        emulate3(n_space, n_vel, 0, N, block, npt)
    for a, b in [(0, 1), (N - 1, N), (n_vel - 3, 3 * n_vel + 5), (N // 3, N // 2)]:
        emulate3(n_space, n_vel, a, b, 256, 2)


def poly_chunk_plan(Nx, elem, req, lds_max=160 * 1024):
    """internal.hpp poly_chunk_plan: (chunk, double-buffered)."""
    row = ((3 * (Nx + 2) + 3) & ~3) * elem
    c = min(req, lds_max // (2 * row))
    if c >= 1:
        return c, True
    return max(1, min(req, lds_max // row)), False


@pytest.mark.parametrize("Nx", [1, 2, 16, 24, 256, 1000, 3000, 5000])
@pytest.mark.parametrize("n", [1, 2, 13, 14, 40])
def test_ready_table_copy_stays_in_bounds(Nx, n):
    """step1d.hpp poly_ready_pipeline with dma_copy16: the FP64 tables of slice
    m sit at qhist + m*sq (sq = poly_stride, a multiple of 4 doubles, so every
    slice starts 16-byte aligned).  Chunk k (slices lo .. lo+cnt-1, highest
    chunk first) is n16 = cnt*sq/2 words; word q = q0 + lane (q0 = wave*64 +
    i*block, q < n16) covers doubles 2q, 2q+1 of the buffer and of the chunk's
    rows.  Every word is copied once, inside the buffer (chunk*sq doubles), from
    the slices below n (the history holds Nt+1 >= n slices).  The FP32 tables
    (FP32C) have the same layout in floats: word q covers floats 4q .. 4q+3.
    Two buffers fit 160 KiB, else one buffer of the chunk does
    (poly_ready_direct, the same copy into one buffer)."""
    cells = Nx + 2
    sq = (3 * cells + 3) & ~3
    for elem in (8, 4):
        assert (sq * elem) % 16 == 0
        check_ready_copy(Nx, n, sq, elem)


def check_ready_copy(Nx, n, sq, elem):
    per = 16 // elem                       # table elements per 16-byte word
    for block in (256, 1024):
        for want in (1, 5, 13, 16, 1 << 20):
            chunk, pipe = poly_chunk_plan(Nx, elem, want)
            assert chunk <= want
            if pipe:
                assert 2 * chunk * sq * elem <= 160 * 1024
                assert chunk == min(want, 160 * 1024 // (2 * sq * elem))
            else:
                assert 2 * sq * elem > 160 * 1024
                assert chunk * sq * elem <= 160 * 1024 or chunk == 1
            top = n - 1
            K = (top + chunk) // chunk
            covered = set()
            for k in range(K):
                h = top - k * chunk
                lo = max(h - chunk + 1, 0)
                cnt = h - lo + 1
                n16 = cnt * sq // per
                seen = set()
                for w in range(block // 64):
                    q0 = w * 64
                    while q0 < n16:
                        for lane in range(64):
                            q = q0 + lane
                            if q < n16:
                                assert q not in seen
                                seen.add(q)
                                last = per * q + per - 1
                                assert last < chunk * sq                       # inside the buffer
                                assert lo * sq + last < (lo + cnt) * sq        # inside the chunk's rows
                                assert (lo + cnt) * sq <= n * sq               # slices below n
                        q0 += block
                assert len(seen) == n16
                covered.update(range(lo, lo + cnt))
            assert covered == set(range(n))


@pytest.mark.parametrize("block", [256, 512])
@pytest.mark.parametrize("n", [1, 3, 9])
def test_fp32c3_direct_lds_copy_stays_in_bounds(block, n):
    """trace3d.hip dma_copy16 and the FP32C double buffer: with st32 = st
    rounded up to 4 floats, chunk k's n16 = cnt*st32/4 words go to word
    q = q0 + lane of its buffer (q0 = wave*64 + i*block, lane < 64, q < n16):
    every word of the chunk is written once, inside the buffer, from inside
    the chunk's rows; the two buffers (chunk*st32 floats each) fit 160 KiB."""
    for N in (8, 16):
        st = (N + 3) ** 3
        st32 = (st + 3) & ~3
        for chunk in (1, 2, 16):
            dma = 2 * chunk * st32 * 4 <= 160 * 1024
            if not dma:
                continue
            top = n - 1
            K = (top + chunk) // chunk
            for k in range(K):
                h = top - k * chunk
                lo = max(h - chunk + 1, 0)
                cnt = h - lo + 1
                n16 = cnt * st32 // 4
                seen = set()
                for w in range(block // 64):
                    q0 = w * 64
                    while q0 < n16:
                        for lane in range(64):
                            q = q0 + lane
                            if q < n16:
                                assert q not in seen
                                seen.add(q)
                                assert 4 * q + 3 < chunk * st32          # inside the buffer
                                assert lo * st32 + 4 * q + 3 < (lo + cnt) * st32   # inside the chunk's rows
                        q0 += block
                assert len(seen) == n16
