"""HIP path vs the oracle, through the C-ABI (needs an MI355X)."""

import numpy as np
import pytest
from numpy.testing import assert_allclose

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _soa(field):  # (N,6) -> [6][N]
    return np.ascontiguousarray(np.asarray(field).T)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")


GRIDS = [(3, 4, 5), (4, 4, 4), (2, 3, 6), (1, 4, 3), (6, 5, 8), (8, 8, 8), (7, 9, 10),
         (16, 16, 16), (12, 10, 14), (32, 32, 32), (48, 48, 48)]


@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_projection_matches_oracle(dims, precision):
    from materialite_amd._native import NativeSolver
    from oracle.projection import apply_projection

    print(f"projection {dims} {precision}", flush=True)
    rng = np.random.default_rng(sum(dims))
    n = int(np.prod(dims))
    field = rng.normal(size=(n, 6))
    for mask in (np.zeros(6), np.array([0, 1, 1, 1, 1, 1.0])):
        ref = apply_projection(field, dims, stress_mask=mask)
        sol = NativeSolver(dims, precision)
        x = _dev(_soa(field))
        y = torch.empty_like(x)
        sol.project_dev(x.data_ptr(), y.data_ptr(), mask)
        torch.cuda.synchronize()
        got = y.cpu().numpy().T
        sol.close()
        scale = np.abs(ref).max()
        tol = 1e-13 if precision == "fp64" else 2e-6
        assert np.abs(got - ref).max() <= tol * scale, (dims, precision, np.abs(got - ref).max() / scale)
