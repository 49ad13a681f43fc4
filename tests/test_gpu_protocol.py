"""The constitutive models' own protocol on the device, against the oracle.

The reference's SmallStrainFFT drives its models through initialize,
calculate_stress_and_tangent (returning (None, None, False) on failure),
update_state_variables, generate_outputs and postprocess
(small_strain_fft/elastic.py:23-46, elastic_viscoplastic.py:52-294,
isotropic_elastic_plastic.py:55-192, multiphase.py:31-95).  Here each call
runs the HIP kernels on a point handle; the oracle restates the same calls
in NumPy.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from materialite_amd import Order2SymmetricTensor, Order4SymmetricTensor, Orientation, SlipSystem  # noqa: E402
from materialite_amd.models.small_strain_fft import (  # noqa: E402
    Elastic, ElasticViscoplastic, IsotropicElasticPlastic, Multiphase, linear, voce,
)
from oracle import constitutive as OC  # noqa: E402
from oracle import tensors as OT  # noqa: E402
from tests import fixtures as F  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _evp_pair(hardening=voce, name="voce", props=None, n=10.0):
    props = props or {"tau_1": 100.0, "theta_0": 20000.0, "theta_1": 10.0}
    dev = ElasticViscoplastic(Order4SymmetricTensor.from_voigt(F.TI_VOIGT), SlipSystem.octahedral(),
                              1.0, n, 300.0, hardening, props)
    ref = F.evp_model(name, props, n)
    return dev, ref


def _strain_path(npts, rng):
    base = np.tile([-0.0012, -0.0012, 0.004, 0, 0, 0], (npts, 1))
    return base + 2e-4 * rng.normal(size=(npts, 6))


def test_evp_protocol_matches_oracle():
    """initialize -> calculate -> update -> calculate on 512 points (Voce):
    iteration counts, stress, tangent and every state variable."""
    print("EVP protocol", flush=True)
    rot, _, _ = F.voronoi_8()
    dev, ref = _evp_pair()
    t_dev = dev.initialize(Orientation(rot, "p"))
    t_ref = ref.initialize(rot)
    assert _rel(t_dev.components, t_ref) <= 1e-14
    rng = np.random.default_rng(5)
    strain = _strain_path(512, rng)
    guess = np.einsum("pmn,pn->pm", t_ref, strain) * 0.5
    for step in range(2):
        e = strain * (1 + 0.2 * step)
        s_dev, d_dev, it_dev = dev.calculate_stress_and_tangent(
            Order2SymmetricTensor(e, "p"), Order2SymmetricTensor(guess, "p"), 1e-3)
        s_ref, d_ref, it_ref = ref.calculate_stress_and_tangent(e, guess, 1e-3)
        print(f"  step {step}: iterations {it_dev} (oracle {it_ref}), stress "
              f"{_rel(s_dev.components, s_ref):.2e}, tangent {_rel(d_dev.components, d_ref):.2e}",
              flush=True)
        assert it_dev == it_ref
        assert _rel(s_dev.components, s_ref) <= 1e-10
        assert _rel(d_dev.components, d_ref) <= 1e-10
        assert _rel(dev.state_variables["slip_resistances"].components, ref.g) <= 1e-10
        assert _rel(dev.state_variables["slip_system_shear_strains"].components, ref.gamma) <= 1e-9
        dev.update_state_variables()
        ref.update_state_variables()
        assert _rel(dev.state_variables["plastic_strains"].components, ref.ep_specimen) <= 1e-9
        guess = s_ref
    out = dev.postprocess(["slip_resistances", "plastic_strains", "slip_system_shear_strains"])
    assert _rel(out["slip_resistances"].components, ref.g) <= 1e-10
    assert _rel(out["plastic_strains"].components, ref.ep_specimen) <= 1e-9
    assert out["slip_resistances"].dims_str == "ps"
    assert dev.state_variables == {}


def test_evp_protocol_reports_failure():
    """A strain far beyond what 60 implicit iterations can reach from the guess:
    the reference returns (None, None, False) (elastic_viscoplastic.py:166-168)."""
    print("EVP protocol failure", flush=True)
    rot, _, _ = F.voronoi_8()
    dev, ref = _evp_pair(hardening=linear, name="linear", props={"hardening_rate": 10.0}, n=10.0)
    t_ref = ref.initialize(rot)
    dev.initialize(Orientation(rot, "p"))
    strain = np.tile([-0.05, -0.05, 0.1, 0, 0, 0], (512, 1))
    guess = np.zeros((512, 6))
    s_ref, d_ref, it_ref = ref.calculate_stress_and_tangent(strain, guess, 10.0)
    s_dev, d_dev, it_dev = dev.calculate_stress_and_tangent(
        Order2SymmetricTensor(strain, "p"), Order2SymmetricTensor(guess, "p"), 10.0)
    assert s_ref is None and not it_ref
    assert s_dev is None and d_dev is None and it_dev is False
    dev.postprocess()


def test_elastic_protocol_matches_oracle():
    print("elastic protocol", flush=True)
    rng = np.random.default_rng(9)
    rot = OT.rotation_from_euler(OT.random_euler(100, rng))
    dev = Elastic(Order4SymmetricTensor.from_voigt(F.TI_VOIGT))
    ref = OC.Elastic(F.TI_STIFFNESS)
    t_dev = dev.initialize(Orientation(rot, "p"))
    t_ref = ref.initialize(rot)
    assert _rel(t_dev.components, t_ref) <= 1e-14
    assert _rel(dev.state_variables["stiffnesses"].components, t_ref) <= 1e-14
    e = rng.normal(scale=1e-3, size=(100, 6))
    s_dev, d_dev, it = dev.calculate_stress_and_tangent(Order2SymmetricTensor(e, "p"),
                                                        Order2SymmetricTensor(e, "p"), 1.0)
    s_ref, _, _ = ref.calculate_stress_and_tangent(e, e, 1.0)
    assert it == 1
    assert _rel(s_dev.components, s_ref) <= 1e-14
    assert dev.generate_outputs(None) == {}


def test_j2_protocol_matches_oracle():
    print("J2 protocol", flush=True)
    n = 64
    rot = np.broadcast_to(np.eye(3), (n, 3, 3))
    props = {"hardening_rate": 1000.0}
    dev = IsotropicElasticPlastic(F.J2_MODULUS, F.J2_SHEAR_MODULUS, F.J2_YIELD, linear, props)
    ref = OC.IsotropicElasticPlastic(F.J2_MODULUS, F.J2_SHEAR_MODULUS, F.J2_YIELD, "linear", props)
    dev.initialize(Orientation(rot, "p"))
    ref.initialize(rot)
    rng = np.random.default_rng(4)
    base = np.array([-0.0005, -0.0005, 0.002, 0, 0, 0])
    for step in range(3):
        e = (1 + step) * base + 1e-4 * rng.normal(size=(n, 6))
        s_dev, d_dev, it_dev = dev.calculate_stress_and_tangent(
            Order2SymmetricTensor(e, "p"), Order2SymmetricTensor(np.zeros((n, 6)), "p"), 1e-4)
        s_ref, d_ref, it_ref = ref.calculate_stress_and_tangent(e, np.zeros((n, 6)), 1e-4)
        print(f"  step {step}: iterations {it_dev} (oracle {it_ref})", flush=True)
        assert it_dev == it_ref
        assert _rel(s_dev.components, s_ref) <= 1e-10
        assert _rel(d_dev.components, d_ref) <= 1e-10
        assert _rel(dev.state_variables["eq_plastic_strains"].components, ref.eq) <= 1e-9
        dev.update_state_variables()
        ref.update_state_variables()
    dev.postprocess()


def test_multiphase_protocol_matches_oracle():
    """EVP grains with an elastic, zero-stiffness pore (the reference's pore
    material, tests/test_elastic_viscoplastic_model.py:120-132)."""
    print("multiphase protocol", flush=True)
    rot, pore, _ = F.voronoi_8()
    idx = [np.nonzero(~pore)[0], np.nonzero(pore)[0]]
    evp, evp_ref = _evp_pair()
    hole = Elastic(Order4SymmetricTensor.from_voigt(F.TI_VOIGT) * 0.0)
    dev = Multiphase([1, 2], [evp, hole], idx, 512)
    ref = OC.Multiphase([evp_ref, OC.Elastic(F.TI_STIFFNESS * 0)], idx, 512)
    t_dev = dev.initialize(Orientation(rot, "p"))
    t_ref = ref.initialize(rot)
    assert _rel(t_dev.components, t_ref) <= 1e-14
    strain = _strain_path(512, np.random.default_rng(6))
    guess = np.einsum("pmn,pn->pm", t_ref, strain) * 0.5
    s_dev, d_dev, it_dev = dev.calculate_stress_and_tangent(
        Order2SymmetricTensor(strain, "p"), Order2SymmetricTensor(guess, "p"), 1e-3)
    s_ref, d_ref, it_ref = ref.calculate_stress_and_tangent(strain, guess, 1e-3)
    assert it_dev == it_ref
    assert _rel(s_dev.components, s_ref) <= 1e-10
    assert _rel(d_dev.components, d_ref) <= 1e-10
    dev.update_state_variables()
    ref.update_state_variables()
    out = dev.postprocess(["slip_resistances", "plastic_strains", "slip_system_shear_strains"])
    g = out["slip_resistances"].components
    assert _rel(g[idx[0]], evp_ref.g) <= 1e-10
    assert np.all(g[idx[1]] == 0.0)
    assert _rel(out["plastic_strains"].components[idx[0]], evp_ref.ep_specimen) <= 1e-9
    # the EVP model's points do not fill whole 64-point blocks of its state
    assert len(idx[0]) % 64 != 0
    assert _rel(out["slip_system_shear_strains"].components[idx[0]], evp_ref.gamma) <= 1e-9
    with pytest.raises(ValueError, match="not available"):
        Multiphase([1], [hole], [np.arange(512)], 512).generate_outputs(["slip_resistances"])


def test_multiphase_object_drives_small_strain_fft():
    """A Multiphase passed as the constitutive model gives the solve that the
    phase-table path gives (small_strain_fft.py:384-398 accepts either)."""
    import pandas as pd

    from materialite_amd.models.small_strain_fft import LoadSchedule, SmallStrainFFT
    from tests.test_gpu_solver import voronoi_material

    print("multiphase object in SmallStrainFFT", flush=True)
    material = voronoi_material()
    mid = material.sizes[2] / 2
    from materialite_amd import Sphere

    material = material.insert_feature(Sphere(radius=1, centroid=[mid, mid, mid]), fields={"phase": 2})
    evp, _ = _evp_pair()
    hole = Elastic(evp.stiffness * 0.0)
    ls = LoadSchedule.from_constant_uniaxial_strain_rate(direction="z")
    table = pd.DataFrame({"phase": [1, 2], "constitutive_model": [evp, hole]})
    by_label = SmallStrainFFT(ls, end_time=4e-3, initial_time_increment=1e-3)(
        material.create_regional_fields("phase", table), phase_label="phase")
    phase = material.extract("phase")
    idx = [np.nonzero(phase == 1)[0], np.nonzero(phase == 2)[0]]
    mp = Multiphase([1, 2], [evp, hole], idx, material.num_points)
    by_object = SmallStrainFFT(ls, end_time=4e-3, initial_time_increment=1e-3,
                               constitutive_model=mp)(material)
    a, b = by_object.extract("stress").components, by_label.extract("stress").components
    assert np.array_equal(a, b)
