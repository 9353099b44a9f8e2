"""The uv-coverage mask of the fused major cycle covers everything the binned
kernel touches (CPU check of the marking rule).

The transform passes skip the row chunks and column blocks of the paired
grid's upper half that no work item marks (casa-rs_amd/csrc/cycle.hip,
k_touch: a tile plus halo marks cells y <= gy/2 in place and cells y >= gy/2
or y == 0 at their Hermitian mirror ((gx - x) mod gx, (gy - y) mod gy)).
Skipping is exact only if the marks are a superset of the cells the binned
kernel reads from the model grid (y > gy/2 read as the conjugate of the
mirror) and adds to the residual grid (2H: y <= gy/2 in place, y >= gy/2 or
y == 0 also at the mirror), casa-rs_amd/csrc/binned.hip.  This test restates
the marking rule and enumerates those cells by brute force over every tile
of small grids (odd and even halves, tiles crossing every edge) and checks
that every touched (block, chunk) is marked and that every mark is inside
the upper blocks.
"""
import pytest

BLOCK, CHUNK = 8, 16


def marks(gx, gy, tile, halo, tx, ty):
    L = tile + 2 * halo
    x0, y0 = tx * tile - halo, ty * tile - halo
    xa, ya = max(x0, 0), max(y0, 0)
    xb, yb = min(x0 + L - 1, gx - 1), min(y0 + L - 1, gy - 1)
    out = set()
    if xa > xb or ya > yb:
        return out
    hy = gy // 2

    def mark(ylo, yhi, xlo, xhi):
        for b in range(ylo // BLOCK, yhi // BLOCK + 1):
            for ch in range(xlo // CHUNK, xhi // CHUNK + 1):
                out.add((b, ch))

    def mirror(ylo, yhi):
        if xa == 0:
            mark(ylo, yhi, 0, 0)
            if xb >= 1:
                mark(ylo, yhi, gx - xb, gx - 1)
        else:
            mark(ylo, yhi, gx - xb, gx - xa)

    if ya <= hy:
        mark(ya, min(yb, hy), xa, xb)
    if yb >= hy:
        mirror(gy - yb, gy - max(ya, hy))
    if ya == 0:
        mirror(0, 0)
    return out


def touched(gx, gy, tile, halo, tx, ty):
    L = tile + 2 * halo
    x0, y0 = tx * tile - halo, ty * tile - halo
    hy = gy // 2
    cells = set()
    for lx in range(L):
        for ly in range(L):
            x, y = x0 + lx, y0 + ly
            if x < 0 or y < 0 or x >= gx or y >= gy:
                continue
            mx = 0 if x == 0 else gx - x
            cells.add((mx, gy - y) if y > hy else (x, y))      # model read
            if y <= hy:
                cells.add((x, y))                               # residual flush, in place
            if y == 0 or y >= hy:
                cells.add((mx, 0 if y == 0 else gy - y))        # and at the mirror
    return {(y // BLOCK, x // CHUNK) for (x, y) in cells}


@pytest.mark.parametrize("gx,gy", [(80, 80), (96, 64), (120, 100), (36, 30), (250, 250)])
@pytest.mark.parametrize("tile,halo", [(8, 3), (16, 4), (32, 9)])
def test_marks_cover_the_binned_kernel(gx, gy, tile, halo):
    n_blocks = (gy // 2 + 1 + BLOCK - 1) // BLOCK
    n_chunks = (gx + CHUNK - 1) // CHUNK
    for tx in range((gx + tile - 1) // tile):
        for ty in range((gy + tile - 1) // tile):
            m = marks(gx, gy, tile, halo, tx, ty)
            missing = touched(gx, gy, tile, halo, tx, ty) - m
            assert not missing, (tx, ty, sorted(missing)[:5])
            assert all(b < n_blocks and ch < n_chunks for b, ch in m)
