"""cgfdm_grid_ranges_noalloc (allocation-free CFL ranges on device-resident
inputs) against cgfdm_grid_ranges (host inputs, library-allocated buffers),
through the C ABI: two calls in a row, the second on a smaller grid, so the
second call runs on the scratch array the first one left behind. Range values
are minima and maxima of the same FP64 quantities computed by the same kernels,
so the two entry points must agree bit for bit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(nx, ny, nz, seed):
    from cgfdm_amd import model as M
    dh = 100.0
    terrain = M.gaussian_hills(nx, ny, dh, 2, 3, seed, hmin=200.0, hmax=900.0, smin=500.0, smax=1500.0)
    coord = M.terrain_coord(nx, ny, nz, dh, terrain, dh * nz)
    rng = np.random.default_rng(seed)
    shape = coord.shape[1:]
    vs = rng.uniform(1500.0, 3500.0, shape)
    medium = np.stack([vs, vs * rng.uniform(1.6, 1.9, shape), rng.uniform(2000.0, 3000.0, shape)]).astype(np.float32)
    return np.ascontiguousarray(coord, np.float32), medium


def _noalloc(L, torch, nx, ny, nz, coord, medium, stream):
    dc = torch.from_numpy(coord).to("cuda:0")
    dm = torch.from_numpy(medium).to("cuda:0")
    out = torch.full((8,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    rc = L.cgfdm_grid_ranges_noalloc(nx, ny, nz, ctypes.c_void_p(dc.data_ptr()), ctypes.c_void_p(dm.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream.cuda_stream))
    assert rc == 0, L.cgfdm_last_error().decode()
    stream.synchronize()
    return out.cpu().numpy()


def test_noalloc_ranges_match_allocating_entry_point():
    import torch
    from cgfdm_amd import solver as S
    L = S.lib()
    stream = torch.cuda.Stream(device="cuda:0")
    # the second grid is smaller: its call reuses the scratch left by the first
    for nx, ny, nz, seed in [(70, 52, 30, 3), (23, 17, 11, 4)]:
        print(f"no-alloc ranges: grid {nx}x{ny}x{nz}, seed {seed}", flush=True)
        coord, medium = _grid(nx, ny, nz, seed)
        ref = np.empty(8, np.float64)
        rc = L.cgfdm_grid_ranges(nx, ny, nz, S._ptr(coord), S._ptr(medium), S._ptr(ref))
        assert rc == 0, L.cgfdm_last_error().decode()
        got = _noalloc(L, torch, nx, ny, nz, coord, medium, stream)
        np.testing.assert_array_equal(got, ref)
        # the medium ranges are the extrema of the padded arrays
        for c in range(3):
            assert got[2 + 2 * c] == float(medium[c].min()) and got[3 + 2 * c] == float(medium[c].max())
        assert 0.0 < got[0] <= got[1]


def test_noalloc_rejects_bad_arguments():
    import torch
    from cgfdm_amd import solver as S
    L = S.lib()
    x = torch.zeros(16, dtype=torch.float32, device="cuda:0")
    out = torch.zeros(8, dtype=torch.float64, device="cuda:0")
    p, o = ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr())
    assert L.cgfdm_grid_ranges_noalloc(0, 4, 4, p, p, o, None) == 2
    assert L.cgfdm_grid_ranges_noalloc(4, 4, 4, None, p, o, None) == 2
    assert L.cgfdm_grid_ranges_noalloc((1 << 30) + 1, 4, 4, p, p, o, None) == 2
    assert L.cgfdm_grid_ranges_noalloc(1 << 20, 1 << 20, 1 << 20, p, p, o, None) == 2
    assert b"2^40" in L.cgfdm_last_error()
