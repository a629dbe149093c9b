"""A host-side model of the wave-private FP32 heat tile's store map
(npbench_amd/csrc/tiles.hip: tile_grid, col_range, heat_wave_tile), CPU only.

For every tile, wave and lane of one launch it lists the 16-byte vectors the
kernel stores -- plane q = P - D for the march steps P in [max(i0 - D, 0),
i1 - 1 + D] with i0 <= q < i1, rows jl + r for the output rows r in
[T, JR - T) below jhi, the lane's vector k0 = kl + 4 lane when it holds an
output point -- and checks, over the shapes the GPU tests run:

* every vector of interior points is stored by exactly one (tile, wave,
  lane) and nothing else is stored;
* every address lies in the grid and is 16-byte aligned;
* every buffer-descriptor access of the tile (a plane's descriptor with
  num_records = n*n*4 bytes, the row offset clamp(jl + r, 0, n-1)*n*4 as the
  scalar offset, the lane's byte offset 4*k0, or 0 outside the grid, as the
  vector offset) ends within num_records, for every row the wave holds,
  halo rows included.

So the tile's stores have one writer per vector per launch and its
descriptor offsets stay in range: wrong values seen with buffer stores at
two waves per SIMD cannot come from overlapping writers or out-of-range
offsets of this map.
"""

import numpy as np
import pytest

V = 4            # floats per 16-byte vector
KT = 64 * V      # columns a wave spans
WAVES = 4        # waves per workgroup


def cdiv(a, b):
    return -(-a // b)


def tile_grid(n, T, rows, seg):
    """tiles.hip tile_grid for an FP32 heat grid on wave-private tiles."""
    tiles_k = 1 if n <= KT else cdiv(n, KT - 2 * V)
    out_rows = WAVES * (rows - 2 * T)
    tiles_j = cdiv(n - 2, out_rows)
    tiles_i = cdiv(n - 2, seg)
    return tiles_k, tiles_j, tiles_i


def col_range(n, tiles_k, tk):
    if tiles_k == 1 and n <= KT:
        return 0, 1, n - 1
    kl = tk * (KT - 2 * V) - V
    return kl, max(kl + V, 1), min(kl + KT - V, n - 1)


def launch_map(n, T, rows, seg, D, row_stride=None):
    """Stores of one launch at depth D as keys (q*n + j)*n + k0 (one per
    stored 16-byte vector, repeats kept) and the largest byte offset a
    descriptor access reaches.  row_stride (rows between consecutive waves'
    first output rows) is the kernel's WJ; another value plants a mistake."""
    JR, WJ = rows, rows - 2 * T
    stride = WJ if row_stride is None else row_stride
    tiles_k, tiles_j, tiles_i = tile_grid(n, T, rows, seg)
    keys, reach = [], 0
    lanes = np.arange(64)
    for tile in range(tiles_k * tiles_j * tiles_i):
        tk, rest = tile % tiles_k, tile // tiles_k
        tj, ti = rest % tiles_j, rest // tiles_j
        kl, klo, khi = col_range(n, tiles_k, tk)
        i0 = 1 + ti * seg
        i1 = min(i0 + seg, n - 1)
        pbeg, pfin = max(i0 - D, 0), i1 - 1 + D
        planes = np.array([P - D for P in range(pbeg, pfin + 1) if i0 <= P - D < i1])
        assert sorted(planes.tolist()) == list(range(i0, i1)), "the march stores each plane once"
        k0 = kl + lanes * V
        vlive = (k0 >= 0) & (k0 + V <= n)
        lob = np.where(vlive, 4 * k0, 0)
        sto = vlive & (k0 + V > klo) & (k0 < khi)
        for w in range(WAVES):
            jl = 1 + (tj * WAVES + w) * stride - T
            jlo, jhi = jl + T, min(jl + T + WJ, n - 1)
            if jlo >= jhi:
                continue
            rowoff = np.clip(jl + np.arange(JR), 0, n - 1) * n * 4
            reach = max(reach, int(rowoff.max() + lob.max() + 16))
            out_rows = np.array([jl + r for r in range(T, JR - T) if jl + r < jhi])
            ks = k0[sto]
            if out_rows.size and ks.size:
                q, j, k = np.meshgrid(planes, out_rows, ks, indexing="ij")
                keys.append(((q * n + j) * n + k).ravel())
    return (np.concatenate(keys) if keys else np.zeros(0, np.int64)), reach


def interior_vectors(n):
    """Keys of the vectors holding at least one interior point."""
    ks = np.array([k0 for k0 in range(0, n - V + 1, V) if k0 + V > 1 and k0 < n - 1])
    q, j, k = np.meshgrid(np.arange(1, n - 1), np.arange(1, n - 1), ks, indexing="ij")
    return np.sort(((q * n + j) * n + k).ravel())


# shapes of tests/test_gpu_tiles.py (wave-private heat tiles) and the headline
SHAPES = [(2, 10), (2, 12), (3, 12), (3, 14)]
SIDES = [4, 8, 12, 20, 68, 264]
SEGS = [3, 32, 64]


@pytest.mark.parametrize("T,rows", SHAPES)
@pytest.mark.parametrize("n", SIDES)
@pytest.mark.parametrize("seg", SEGS)
def test_each_interior_vector_has_exactly_one_writer(T, rows, n, seg):
    for D in range(1, T + 1):
        keys, reach = launch_map(n, T, rows, seg, D)
        assert len(keys) == len(np.unique(keys)), f"a vector stored twice (n={n}, D={D})"
        # exactly the interior vectors: in the grid, aligned, none missing
        assert np.array_equal(np.sort(keys), interior_vectors(n))
        assert reach <= n * n * 4, "a descriptor access past num_records"


def test_headline_grid_has_one_writer_per_vector():
    """256^3 at the headline shape (T=3, 14 rows, 64 planes per tile) in its
    depth-3 rounds."""
    n = 256
    keys, reach = launch_map(n, 3, 14, 64, 3)
    assert len(keys) == len(np.unique(keys)) == (n - 2) ** 2 * (n // V)
    assert np.array_equal(np.sort(keys), interior_vectors(n))
    assert reach <= n * n * 4


def test_a_planted_overlap_is_caught():
    """With consecutive waves' rows one row closer than the kernel places
    them, adjacent waves store one row twice: the check sees it."""
    n, T, rows, seg = 68, 2, 12, 32
    keys, _ = launch_map(n, T, rows, seg, 2, row_stride=rows - 2 * T - 1)
    assert len(keys) != len(np.unique(keys))
