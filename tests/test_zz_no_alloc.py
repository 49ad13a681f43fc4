"""mtl_project_noalloc (projection with its spectrum and tables in a fixed
device array, no allocation) against mtl_solver_project of a handle of the
same grid and precision, through the C-ABI: two calls in a row, the second
on a smaller grid, in every precision form."""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")


def _noalloc(lib, dims, prec, x, y, mask):
    m = np.ascontiguousarray(mask, dtype=np.float64)
    return lib.mtl_project_noalloc(int(dims[0]), int(dims[1]), int(dims[2]), prec,
                                   ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                   m.ctypes.data_as(ctypes.c_void_p))


@pytest.mark.parametrize("precision", ["fp64", "mixed", "refined"])
def test_noalloc_projection_matches_handle(precision):
    from materialite_amd._native import PRECISIONS, NativeSolver, load_library

    lib = load_library()
    prec = PRECISIONS[precision]
    for dims in [(24, 20, 18), (12, 10, 9)]:  # the second call is smaller
        print(f"no-alloc projection {dims} {precision}", flush=True)
        n = int(np.prod(dims))
        gen = torch.Generator(device="cuda").manual_seed(n)
        x = torch.randn(6, n, dtype=torch.float64, device="cuda", generator=gen)
        for mask in (np.zeros(6), np.array([0, 1, 1, 1, 1, 1.0])):
            ref = torch.empty_like(x)
            sol = NativeSolver(dims, precision)
            torch.cuda.synchronize()
            sol.project_dev(x.data_ptr(), ref.data_ptr(), mask)
            sol.close()
            got = torch.full_like(x, float("nan"))
            torch.cuda.synchronize()
            assert _noalloc(lib, dims, prec, x, got, mask) == 0, lib.mtl_last_error()
            torch.cuda.synchronize()
            scale = ref.abs().max().item()
            err = (got - ref).abs().max().item()
            print(f"  mask {mask.tolist()}: max |noalloc - handle| / max |handle| = {err / scale:.3e}",
                  flush=True)
            # the same kernels on the same tables
            assert err <= 1e-14 * scale, err / scale


def test_noalloc_rejects_grids_beyond_its_scratch():
    from materialite_amd._native import MTL_FP64, load_library

    lib = load_library()
    x = torch.zeros(6, 8, dtype=torch.float64, device="cuda")
    y = torch.zeros_like(x)
    assert _noalloc(lib, (160, 160, 160), MTL_FP64, x, y, np.zeros(6)) != 0
    assert b"bytes of scratch" in lib.mtl_last_error()
