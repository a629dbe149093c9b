"""Temporal-blocked and batched sweeps on the GPU.

The tile kernels perform, per point, exactly the plain sweep's arithmetic,
so every result here must be bitwise equal to the plain per-grid sweeps
(tblock=1), whose FP64 form is itself bitwise the NumPy oracle (checked
again here for FP64).  Sizes cover one and several tiles along every axis,
odd and even sweep counts, both sweep parities, and depth-1 rounds.
"""

import numpy as np
import pytest

from oracle import stencils as ref

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

DT = {"f64": torch.float64, "f32": torch.float32, "bf16": torch.bfloat16}
TMAX = {"f64": 2, "f32": 4, "bf16": 2}


@pytest.fixture(scope="module")
def st():
    from npbench_amd import device, stencils
    device.require_device()
    return stencils


def grids(kind, n, dtype, seed):
    rng = np.random.default_rng(seed)
    shape = (n, n) if kind == "jacobi_2d" else (n, n, n)
    A = rng.random(shape)
    B = rng.random(shape)
    return A, B


def share_boundary(A, B):
    """B with A's global-boundary values (all faces), its interior kept."""
    B = B.copy()
    inner = tuple(slice(1, -1) for _ in A.shape)
    mask = np.ones(A.shape, bool)
    mask[inner] = False
    B[mask] = A[mask]
    return B


# one point of each face of a heat grid: (axis, side)
FACES = [(0, 0), (0, -1), (1, 0), (1, -1), (2, 0), (2, -1)]


def face_point(n, face, seed):
    """An index of a point on `face` away from its edges (a 3-D grid)."""
    axis, side = face
    rng = np.random.default_rng(seed)
    idx = [int(x) for x in rng.integers(1, n - 1, size=3)]
    idx[axis] = 0 if side == 0 else n - 1
    return tuple(idx)


def run(st, kind, A, B, dtype, nsweeps, first, tblock):
    from npbench_amd import device
    a, b = device.to_device(A, DT[dtype]), device.to_device(B, DT[dtype])
    f = st.jacobi2d_sweeps if kind == "jacobi_2d" else st.heat3d_sweeps
    f(a, b, nsweeps, first=first, tblock=tblock)
    torch.cuda.synchronize()
    return a, b


def same(x, y):
    return torch.equal(x.view(torch.int16) if x.dtype == torch.bfloat16 else x,
                       y.view(torch.int16) if y.dtype == torch.bfloat16 else y)


JACOBI_CASES = [("f64", 8), ("f64", 130), ("f64", 266), ("f32", 12), ("f32", 260), ("f32", 516),
                ("bf16", 16), ("bf16", 1032)]
HEAT_CASES = [("f64", 6), ("f64", 40), ("f64", 132), ("f32", 8), ("f32", 68), ("f32", 264),
              ("bf16", 16), ("bf16", 520)]


@pytest.mark.parametrize("dtype,n", JACOBI_CASES)
@pytest.mark.parametrize("nsweeps,first", [(1, 0), (2, 1), (5, 0), (9, 1), (14, 0)])
def test_jacobi_tiled_matches_plain(st, dtype, n, nsweeps, first):
    A, B = grids("jacobi_2d", n, dtype, n + nsweeps)
    pa, pb = run(st, "jacobi_2d", A, B, dtype, nsweeps, first, 1)
    for T in range(2, TMAX[dtype] + 1):
        print(f"jacobi tiled {dtype} n={n} sweeps={nsweeps} first={first} T={T}", flush=True)
        ta, tb = run(st, "jacobi_2d", A, B, dtype, nsweeps, first, T)
        assert same(ta, pa) and same(tb, pb), f"T={T}"
    if dtype == "f64":
        ra, rb = A.copy(), B.copy()
        for s in range(first, first + nsweeps):
            if s % 2 == 0:
                ref.jacobi_sweeps(ra, rb, 1)
            else:
                ref.jacobi_sweeps(rb, ra, 1)
        assert np.array_equal(pa.cpu().numpy(), ra) and np.array_equal(pb.cpu().numpy(), rb)


@pytest.mark.parametrize("dtype,n", HEAT_CASES)
@pytest.mark.parametrize("nsweeps,first", [(1, 1), (3, 0), (6, 1), (11, 0)])
def test_heat_tiled_matches_plain(st, dtype, n, nsweeps, first):
    A, B = grids("heat_3d", n, dtype, 3 * n + nsweeps)
    pa, pb = run(st, "heat_3d", A, B, dtype, nsweeps, first, 1)
    for T in range(2, TMAX[dtype] + 1):
        print(f"heat tiled {dtype} n={n} sweeps={nsweeps} first={first} T={T}", flush=True)
        ta, tb = run(st, "heat_3d", A, B, dtype, nsweeps, first, T)
        assert same(ta, pa) and same(tb, pb), f"T={T}"
    if dtype == "f64" and n <= 40:
        ra, rb = A.copy(), B.copy()
        for s in range(first, first + nsweeps):
            if s % 2 == 0:
                ref.heat_sweeps(ra, rb, 1)
            else:
                ref.heat_sweeps(rb, ra, 1)
        assert np.array_equal(pa.cpu().numpy(), ra) and np.array_equal(pb.cpu().numpy(), rb)


def _mixed_batch(dtype, seed):
    """Uneven grids of both kinds with different sweep ranges."""
    rng = np.random.default_rng(seed)
    specs = [("jacobi_2d", 64, 7, 0), ("heat_3d", 24, 4, 1), ("jacobi_2d", 520, 12, 1),
             ("heat_3d", 68, 9, 0), ("jacobi_2d", 8, 1, 0), ("heat_3d", 8, 2, 0),
             ("jacobi_2d", 1024, 3, 0), ("heat_3d", 132, 5, 1)]
    out = []
    vec = {"f64": 2, "f32": 4, "bf16": 8}[dtype]
    for i, (kind, n, ns, first) in enumerate(specs):
        n = -(-n // vec) * vec  # whole 16-byte vectors per row
        A, B = grids(kind, n, dtype, int(rng.integers(1 << 30)))
        if i % 4 == 3:  # heat grids of equal boundaries (as B = A.copy() gives)
            B = share_boundary(A, B)
        out.append((kind, A, B, ns, first))
    return out


@pytest.mark.parametrize("dtype", ["f32", "f64", "bf16"])
def test_plan_batch_matches_each_grid_alone(st, dtype):
    from npbench_amd import batch, device
    cases = _mixed_batch(dtype, 5)
    for T in range(2, TMAX[dtype] + 1):
        print(f"plan batch {dtype} T={T}", flush=True)
        gs = []
        for kind, A, B, ns, first in cases:
            a, b = device.to_device(A, DT[dtype]), device.to_device(B, DT[dtype])
            gs.append(batch.Grid(kind, a, b, ns, first))
        plan = batch.Plan(gs, tblock=T)
        info = plan.info()
        assert info["tiles"] > 0 and info["launches"] >= 2
        plan.launch()
        torch.cuda.synchronize()
        for g, (kind, A, B, ns, first) in zip(gs, cases):
            pa, pb = run(st, kind, A, B, dtype, ns, first, 1)
            assert same(g.A, pa) and same(g.B, pb), f"{kind} n={A.shape[0]}"
        plan.close()


def test_plan_graph_replays_match_eager(st):
    """Each replay of the captured graph, with eager work and syncs between
    replays, gives the eager result (fresh inputs copied in before each)."""
    from npbench_amd import batch, device
    cases = _mixed_batch("f32", 9)
    gs, inputs = [], []
    for kind, A, B, ns, first in cases:
        a, b = device.to_device(A, torch.float32), device.to_device(B, torch.float32)
        inputs.append((a.clone(), b.clone()))
        gs.append(batch.Grid(kind, a, b, ns, first))
    plan = batch.Plan(gs, tblock=3)
    plan.launch()
    torch.cuda.synchronize()
    eager = [(g.A.clone(), g.B.clone()) for g in gs]
    side = torch.cuda.Stream()
    for g, (a0, b0) in zip(gs, inputs):
        g.A.copy_(a0)
        g.B.copy_(b0)
    torch.cuda.synchronize()
    plan.capture(side)
    for rep in range(3):
        for g, (a0, b0) in zip(gs, inputs):
            g.A.copy_(a0)
            g.B.copy_(b0)
        torch.cuda.synchronize()
        plan.launch(side)
        side.synchronize()
        _ = torch.ones(16, device="cuda").sum().item()  # eager work between replays
        for g, (ea, eb) in zip(gs, eager):
            assert torch.equal(g.A, ea) and torch.equal(g.B, eb), f"replay {rep}"
    plan.close()


def test_plan_rejects_bad_grids(st):
    from npbench_amd import _native, batch, device
    a = device.to_device(np.zeros((10, 10)), torch.float32)  # 10 % 4 != 0
    b = a.clone()
    with pytest.raises(_native.NpbxError):
        batch.Plan([batch.Grid("jacobi_2d", a, b, 4)], tblock=2)
    a = device.to_device(np.zeros((12, 12)), torch.float64)
    with pytest.raises(_native.NpbxError):
        batch.Plan([batch.Grid("jacobi_2d", a, a.clone(), 4)], tblock=3)  # FP64 max T is 2


@pytest.fixture
def heat_rows():
    """Sets the library's heat.rows knob for one test and restores it."""
    from npbench_amd import _native
    old = _native.get_tuning("heat.rows")

    yield lambda rows: _native.set_tuning("heat.rows", rows)
    _native.set_tuning("heat.rows", old)


WAVE_HEAT_CASES = [4, 8, 12, 20, 68, 264, 520]


WAVE_SHAPES = [(2, 10), (2, 12), (3, 12), (3, 14)]  # (T, rows)


@pytest.mark.parametrize("T,rows", WAVE_SHAPES)
@pytest.mark.parametrize("n", WAVE_HEAT_CASES)
@pytest.mark.parametrize("nsweeps,first", [(1, 1), (2, 0), (3, 0), (6, 1), (11, 0)])
@pytest.mark.parametrize("bnd", ["different", "equal"])
def test_heat_wave_tiles_match_plain(st, heat_rows, T, rows, n, nsweeps, first, bnd):
    """The wave-private FP32 heat tile (heat.rows 10 .. 14, depth 2 and 3):
    every size class -- grids of a few rows (several boundary rows in one
    wave), one and several tiles along j, one and several along k, first and
    last planes, rounds of every depth -- with A and B of different
    boundaries (boundary values patched in from the packed columns and the
    arrays) and of equal ones (taken from registers), bitwise the plain
    sweeps."""
    A, B = grids("heat_3d", n, "f32", 7 * n + nsweeps)
    if bnd == "equal":
        B = share_boundary(A, B)
    pa, pb = run(st, "heat_3d", A, B, "f32", nsweeps, first, 1)
    heat_rows(rows)
    print(f"heat wave tile T={T} rows={rows} n={n} sweeps={nsweeps} first={first} {bnd}",
          flush=True)
    ta, tb = run(st, "heat_3d", A, B, "f32", nsweeps, first, T)
    assert same(ta, pa) and same(tb, pb)


@pytest.mark.parametrize("T,rows", [(2, 12), (3, 14)])
@pytest.mark.parametrize("n", [20, 68])
@pytest.mark.parametrize("face", FACES)
def test_heat_wave_tiles_one_boundary_point_differs(st, heat_rows, T, rows, n, face):
    """Boundaries equal but for one point of one face: the grid must take the
    patching tile, which reproduces that point's level-parity value."""
    A, B = grids("heat_3d", n, "f32", 11 * n)
    B = share_boundary(A, B)
    B[face_point(n, face, n)] += 0.5
    pa, pb = run(st, "heat_3d", A, B, "f32", 6, 0, 1)
    heat_rows(rows)
    print(f"heat wave tile T={T} rows={rows} n={n} face={face}", flush=True)
    ta, tb = run(st, "heat_3d", A, B, "f32", 6, 0, T)
    assert same(ta, pa) and same(tb, pb)


@pytest.mark.parametrize("T,rows", WAVE_SHAPES)
def test_plan_batch_wave_tiles(st, heat_rows, T, rows):
    from npbench_amd import batch, device
    heat_rows(rows)
    cases = _mixed_batch("f32", 11)
    for seg in (3, 32):
        print(f"plan batch wave tiles T={T} rows={rows} seg={seg}", flush=True)
        gs = []
        for kind, A, B, ns, first in cases:
            a, b = device.to_device(A, torch.float32), device.to_device(B, torch.float32)
            gs.append(batch.Grid(kind, a, b, ns, first))
        plan = batch.Plan(gs, tblock=T, seg=seg)
        plan.launch()
        torch.cuda.synchronize()
        for g, (kind, A, B, ns, first) in zip(gs, cases):
            pa, pb = run(st, kind, A, B, "f32", ns, first, 1)
            assert same(g.A, pa) and same(g.B, pb), f"{kind} n={A.shape[0]}"
        plan.close()


def test_pack_cols_layout(st):
    """npbx_pack_cols: cols[a][q][j][s] = (A, B)[a][q][j][(0, n-1)[s]], then
    one word that is 1 exactly when A's and B's boundaries are equal (each
    face checked with one point differing)."""
    import ctypes
    from npbench_amd import _native, device
    n = 20
    rng = np.random.default_rng(3)
    a = device.to_device(rng.random((n, n, n)), torch.float32)
    b = device.to_device(rng.random((n, n, n)), torch.float32)
    elems = ctypes.c_int64()
    _native.call("npbx_cols_elems", n, ctypes.byref(elems))
    assert elems.value == 4 * n * n + 1
    cols = torch.empty(elems.value, dtype=torch.float32, device="cuda")

    def pack(x, y):
        _native.call("npbx_pack_cols", device.dtype_code(x), x.data_ptr(), y.data_ptr(), n,
                     cols.data_ptr(), None)
        torch.cuda.synchronize()
        return int(cols[4 * n * n:].view(torch.int32).item())

    assert pack(a, b) == 0
    want = torch.stack([torch.stack([x[:, :, 0], x[:, :, n - 1]], dim=-1) for x in (a, b)])
    assert torch.equal(cols[:4 * n * n].view(2, n, n, 2), want)
    A = a.cpu().numpy()
    B = share_boundary(A, b.cpu().numpy())
    assert pack(a, device.to_device(B, torch.float32)) == 1
    assert pack(a, a) == 1
    for face in FACES:
        for seed in range(3):
            C = B.copy()
            C[face_point(n, face, seed)] += 1.0
            assert pack(a, device.to_device(C, torch.float32)) == 0, (face, seed)
        for edge in ((0, 0, 5), (n - 1, 5, n - 1), (4, n - 1, 0)):  # points on edges
            C = B.copy()
            C[edge] -= 1.0
            assert pack(a, device.to_device(C, torch.float32)) == 0, edge
    with pytest.raises(_native.NpbxError):
        _native.call("npbx_pack_cols", 0, a.data_ptr(), b.data_ptr(), n, cols.data_ptr(), None)
