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


# The long-x grids (nx >= 250) take the x pass that streams the components
# through LDS one at a time (six resident components do not fit two kz
# columns there), with full, partial (256 x 2 x 40: nzc = 20 in tiles of 8) and
# odd-width (256 x 4 x 6: nzc = 3) tiles and a radix-5 length (250); in FP32,
# 250 x 3 x 4 and 256 x 3 x 32 (whole even-width tiles) move pairs of values.
GRIDS = [(3, 4, 5), (4, 4, 4), (2, 3, 6), (1, 4, 3), (6, 5, 8), (8, 8, 8), (7, 9, 10),
         (16, 16, 16), (12, 10, 14), (32, 32, 32), (48, 48, 48),
         (256, 4, 6), (256, 2, 40), (250, 3, 4), (512, 2, 6), (256, 3, 32),
         (11, 13, 4), (22, 26, 12), (13, 11, 26),  # radix-11/13 passes on every axis
         (17, 4, 6), (19, 19, 19), (2, 23, 34), (31, 6, 37)]  # generic prime passes


@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("precision", ["fp64", "mixed", "refined"])
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
        # refined: spectrum kept to 32 significant bits between the passes
        tol = {"fp64": 1e-13, "refined": 1e-8, "mixed": 2e-6}[precision]
        assert np.abs(got - ref).max() <= tol * scale, (dims, precision, np.abs(got - ref).max() / scale)


def test_prime_grid_solve_matches_oracle():
    """A 17^3 EVP solve (prime length on every axis: the generic FFT pass)
    against the oracle's numpy.fft path."""
    from materialite_amd import serving
    from materialite_amd.models.small_strain_fft import LoadSchedule, SmallStrainFFT
    from oracle.solver import LoadSchedule as OLS
    from oracle.solver import SmallStrainFFT as OSolver
    from tests.test_gpu_cutback import _oracle_evp

    print("17^3 EVP solve", flush=True)
    material = serving.grain_material(17, 6, 20260017)
    rot = material.extract("orientation").rotation_matrix.reshape(-1, 3, 3)
    ref = OSolver(OLS.from_constant_uniaxial_strain_rate(1e-3, "x"), end_time=0.6,
                  initial_time_increment=0.2).run([17] * 3, rot, _oracle_evp())
    for precision, tol in (("fp64", 1e-9), ("refined", 1e-6)):
        out = SmallStrainFFT(LoadSchedule.from_constant_uniaxial_strain_rate(1e-3, "x"),
                             end_time=0.6, initial_time_increment=0.2,
                             constitutive_model=serving.config3_evp(), precision=precision)(material)
        err = np.abs(out.extract("stress").components - ref["stress"]).max() / np.abs(ref["stress"]).max()
        print(f"  {precision}: {err:.3e}", flush=True)
        assert err <= tol, (precision, err)


def test_invalid_model_arguments_are_errors():
    """Model registration rejects what the kernels cannot run, at the call
    that passes it: no slip systems, a slip-system count the EVP kernels are
    not compiled for, an unknown hardening law."""
    from materialite_amd._native import NativeSolver

    sol = NativeSolver((4, 4, 4), "fp64")
    try:
        c = np.eye(6) * 100.0
        with pytest.raises(Exception, match="at least one slip system"):
            sol.add_evp(c, np.zeros((0, 6)), 1e-3, 20.0, 50.0, 0, [])
        with pytest.raises(Exception, match="5 slip systems is not built"):
            sol.add_evp(c, np.ones((5, 6)), 1e-3, 20.0, 50.0, 0, [])
        with pytest.raises(Exception, match="unknown hardening law"):
            sol.add_evp(c, np.ones((12, 6)), 1e-3, 20.0, 50.0, 7, [])
        assert sol.add_evp(c, np.ones((12, 6)), 1e-3, 20.0, 50.0, 0, []) == 0
    finally:
        sol.close()


@pytest.mark.parametrize("dims,precision", [((128, 128, 128), "fp64"), ((128, 128, 128), "mixed"),
                                             ((128, 128, 128), "refined"),
                                             ((256, 256, 256), "mixed")])
def test_projection_properties_full_size(dims, precision):
    """At the workload sizes (128^3, and 256^3 where the x pass streams the
    components): P is an orthogonal projector -- P(P x) = P x and
    <P x, w> = <x, P w> -- and P(a x + b w) = a P x + b P w."""
    from materialite_amd._native import NativeSolver

    print(f"projection properties {dims} {precision}", flush=True)
    n = int(np.prod(dims))
    gen = torch.Generator(device="cuda").manual_seed(sum(dims))
    x = torch.randn(6, n, dtype=torch.float64, device="cuda", generator=gen)
    w = torch.randn(6, n, dtype=torch.float64, device="cuda", generator=gen)
    mask = np.array([0, 1, 1, 1, 1, 1.0])
    sol = NativeSolver(dims, precision)

    def proj(a):
        out = torch.empty_like(a)
        torch.cuda.synchronize()
        sol.project_dev(a.data_ptr(), out.data_ptr(), mask)
        torch.cuda.synchronize()
        return out

    px, pw = proj(x), proj(w)
    ppx = proj(px)
    pxw = proj(2.0 * x - 3.0 * w)
    sol.close()
    tol = {"fp64": 1e-12, "refined": 1e-8, "mixed": 2e-6}[precision]
    scale = px.abs().max().item()
    assert (ppx - px).abs().max().item() <= tol * scale
    assert (pxw - (2.0 * px - 3.0 * pw)).abs().max().item() <= 5 * tol * scale
    lhs, rhs = (px * w).sum().item(), (x * pw).sum().item()
    assert abs(lhs - rhs) <= tol * 10 * (px.norm() * w.norm()).item()


def _pore_case():
    from oracle import constitutive as C
    from tests import fixtures as F

    rot, pore, _ = F.voronoi_8()
    idx = [np.nonzero(~pore)[0], np.nonzero(pore)[0]]
    mp = C.Multiphase([C.Elastic(F.TI_STIFFNESS), C.Elastic(F.TI_STIFFNESS * 0)], idx, 512)
    return rot, pore, mp


def _check_minres(precision, x, itn, x_ref, itn_ref, lhs, b, bounds):
    """FP64 and refined: the oracle's iteration count (+-1) and iterate.  Every
    form: the device iterate against the oracle's iterate after the same number
    of iterations, within `bounds[precision]` of its largest value (a Krylov
    bug moves it by order one; rounding by far less: the bounds are a few times
    the deviations measured on MI355X)."""
    from oracle.minres import minres

    scale = np.abs(x_ref).max()
    x_k, itn_k, _ = minres(lhs, b, rtol=1e-30, maxiter=itn)
    assert itn_k == itn
    dev_k = np.abs(x - x_k).max() / np.abs(x_k).max()
    print(f"iterations {itn} (oracle {itn_ref}); iterate vs oracle's {itn}th: {dev_k:.3e}",
          flush=True)
    if precision != "mixed":
        assert abs(itn - itn_ref) <= 1, (itn, itn_ref)
    if precision == "fp64":
        assert np.abs(x - x_ref).max() <= 1e-9 * scale
    assert dev_k <= bounds[precision], (precision, dev_k)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "refined"])
def test_minres_matches_oracle(precision):
    from materialite_amd._native import NativeSolver
    from oracle.minres import minres
    from oracle.projection import apply_projection
    from tests import fixtures as F

    print(f"minres {precision}", flush=True)
    rot, pore, mp = _pore_case()
    tangent = mp.initialize(rot)
    mask = np.array([1, 1, 0, 1, 1, 1.0])
    sig = np.einsum("pmn,n->pm", tangent, [0, 0, 5e-5, 0, 0, 0])
    b = apply_projection(-sig, (8, 8, 8), stress_mask=mask)

    def lhs(x):
        return apply_projection(np.einsum("pmn,pn->pm", tangent, x.reshape(512, 6)), (8, 8, 8),
                                stress_mask=mask).ravel()

    rtol = 1e-5 if precision == "mixed" else 1e-7
    x_ref, itn_ref, _ = minres(lhs, b.ravel(), rtol=rtol, maxiter=5 * 6 * 512)
    sol = NativeSolver((8, 8, 8), precision)
    sol.add_elastic(F.TI_STIFFNESS)
    sol.add_elastic(F.TI_STIFFNESS * 0)
    sol.set_microstructure(rot, pore.astype(np.int32))
    sol.initialize()
    bd = _dev(_soa(b))
    xd = torch.empty_like(bd)
    itn = sol.minres_dev(bd.data_ptr(), xd.data_ptr(), mask, rtol)
    x = xd.cpu().numpy().T.ravel()
    sol.close()
    # the pore makes the system singular: FP32 Lanczos vectors drift most here
    # (measured 5.0e-4 after 45 iterations; refined 7.5e-6 after 85, where the
    # FP32 w and x vectors dominate)
    _check_minres(precision, x, itn, x_ref, itn_ref, lhs, b.ravel(),
                  {"fp64": 1e-9, "refined": 3e-5, "mixed": 2e-3})


@pytest.mark.parametrize("precision", ["fp64", "mixed", "refined"])
def test_minres_partial_point_block(precision):
    """5 x 6 x 7 = 210 points: three 64-point blocks and a partial one.  The
    Krylov vectors are stored blocked by 64 points; the padding of the last
    block must stay out of every inner product and of the returned solution."""
    from materialite_amd._native import NativeSolver
    from oracle import constitutive as C
    from oracle import tensors as OT
    from oracle.minres import minres
    from oracle.projection import apply_projection
    from tests import fixtures as F

    dims, n = (5, 6, 7), 210
    print(f"minres partial block {precision}", flush=True)
    rng = np.random.default_rng(210)
    rot = OT.rotation_from_euler(OT.random_euler(n, rng))
    tangent = C.Elastic(F.TI_STIFFNESS).initialize(rot)
    mask = np.zeros(6)
    sig = np.einsum("pmn,pn->pm", tangent, rng.normal(scale=1e-4, size=(n, 6)))
    b = apply_projection(-sig, dims, stress_mask=mask)

    def lhs(x):
        return apply_projection(np.einsum("pmn,pn->pm", tangent, x.reshape(n, 6)), dims,
                                stress_mask=mask).ravel()

    rtol = 1e-5 if precision == "mixed" else 1e-7
    x_ref, itn_ref, _ = minres(lhs, b.ravel(), rtol=rtol, maxiter=5 * 6 * n)
    sol = NativeSolver(dims, precision)
    sol.add_elastic(F.TI_STIFFNESS)
    sol.set_microstructure(rot, np.zeros(n, dtype=np.int32))
    sol.initialize()
    bd = _dev(_soa(b))
    xd = torch.empty_like(bd)
    itn = sol.minres_dev(bd.data_ptr(), xd.data_ptr(), mask, rtol)
    x = xd.cpu().numpy().T.ravel()
    sol.close()
    _check_minres(precision, x, itn, x_ref, itn_ref, lhs, b.ravel(),
                  {"fp64": 1e-9, "refined": 1e-6, "mixed": 1e-4})


@pytest.mark.parametrize("precision", ["fp64", "mixed", "refined"])
@pytest.mark.parametrize("rate_exponent", [10.0, 10.5])
def test_evp_update_matches_oracle(precision, rate_exponent):
    """rate exponent 10: |tau/g|^9 by squaring; 10.5: the general pow()."""
    from materialite_amd._native import NativeSolver
    from oracle import tensors as T
    from tests import fixtures as F

    print(f"evp update {precision} n={rate_exponent}", flush=True)
    rot, _, _ = F.voronoi_8()
    props = {"tau_1": 100.0, "theta_0": 20000.0, "theta_1": 10.0}
    model = F.evp_model("voce", props, rate_exponent)
    tangent = model.initialize(rot)
    rng = np.random.default_rng(5)
    strain = np.tile([-0.0012, -0.0012, 0.004, 0, 0, 0], (512, 1)) + 2e-4 * rng.normal(size=(512, 6))
    guess = np.einsum("pmn,pn->pm", tangent, strain) * 0.5
    stress_ref, tangent_ref, it_ref = model.calculate_stress_and_tangent(strain, guess, 1e-3)
    assert it_ref
    sol = NativeSolver((8, 8, 8), precision)
    schmid = T.schmid_mandel(*T.octahedral_slip_systems())
    sol.add_evp(F.TI_STIFFNESS, schmid, 1.0, rate_exponent, 300.0, 2,
                [props["theta_0"], props["theta_1"], props["tau_1"]])
    sol.set_microstructure(rot, np.zeros(512, np.int32))
    sol.initialize()
    sd, gd = _dev(_soa(strain)), _dev(_soa(guess))
    out = torch.empty_like(sd)
    it = sol.constitutive_dev(sd.data_ptr(), gd.data_ptr(), 1e-3, out.data_ptr())
    stress = out.cpu().numpy().T
    tan = sol.get_field("tangent").reshape(512, 6, 6)
    g = sol.get_field("slip_resistances")
    sol.close()
    serr = np.abs(stress - stress_ref).max() / np.abs(stress_ref).max()
    terr = np.abs(tan - tangent_ref).max() / np.abs(tangent_ref).max()
    print(f"iterations {it} (oracle {it_ref}), stress err {serr:.3e}, tangent err {terr:.3e}",
          flush=True)
    assert it == it_ref, (it, it_ref)
    assert np.abs(stress - stress_ref).max() <= 1e-10 * np.abs(stress_ref).max()
    ttol = 1e-10 if precision == "fp64" else 1e-6
    assert np.abs(tan - tangent_ref).max() <= ttol * np.abs(tangent_ref).max()
    assert np.abs(g - model.g).max() <= 1e-10 * np.abs(model.g).max()
