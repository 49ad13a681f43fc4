"""Same-size solves advanced together (mtl_solver_step_batch): each result is
the one its solve gives alone, also when the solves of a batch take different
numbers of Newton and Krylov iterations, cut back at different times or end
at different times, and when requests join a serving lane while others run."""

import numpy as np
import pytest

from materialite_amd import serving

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")


def _jobs(precision, n=24):
    """Three jobs on one n^3 grid: EVP under uniaxial x, elastic under xy
    shear, EVP from an oversized first increment (it cuts back) under yz shear
    with a longer schedule."""
    from materialite_amd.models.small_strain_fft import SmallStrainFFT

    out = []
    specs = [(31, 10, "evp", "xx", 0.5, 3.0), (32, 14, "elastic", "xy", 0.5, 2.0),
             (33, 8, "evp", "yz", 2.0, 4.0)]
    for seed, grains, model, direction, dt, end in specs:
        material = serving.grain_material(n, grains, seed)
        cm = serving.config3_evp() if model == "evp" else serving.cubic_elastic()
        solver = SmallStrainFFT(serving.stream_load_schedule(direction), end_time=end,
                                initial_time_increment=dt, max_time_increment=0.5,
                                constitutive_model=cm, precision=precision)
        out.append((solver, material, {"output_variables": ["plastic_strains"]}
                    if model == "evp" else {}))
    return out


@pytest.mark.parametrize("graphs", ["1", "0"])
@pytest.mark.parametrize("precision", ["fp64", "refined", "mixed"])
def test_batched_runs_match_single_runs(precision, graphs, monkeypatch):
    """With the batch iterations launched one by one (the default) and
    replayed from recorded graphs."""
    from materialite_amd.models.small_strain_fft import SmallStrainFFT

    monkeypatch.setenv("MTL_BATCH_GRAPHS", graphs)
    print(f"batch {precision} graphs {graphs}", flush=True)
    alone = []
    for solver, material, kw in _jobs(precision):
        out = solver.run(material, **kw)
        alone.append((out, solver.stats["history"]))
    jobs = _jobs(precision)
    together = SmallStrainFFT.run_batch(jobs)
    histories = [j[0].stats["history"] for j in jobs]
    assert any(not h[2] for h in alone[2][1]), "the third job should cut back"
    for (a, ha), b, hb in zip(alone, together, histories):
        print("  alone:", [h[2:5] for h in ha], "\n  batch:", [h[2:5] for h in hb], flush=True)
        assert ha == hb
        sa, sb = a.extract("stress").components, b.extract("stress").components
        assert np.abs(sa - sb).max() <= 1e-12 * np.abs(sa).max()
    pa = alone[0][0].extract("plastic_strains").components
    pb = together[0].extract("plastic_strains").components
    assert np.abs(pa - pb).max() <= 1e-12 * np.abs(pa).max()


@pytest.mark.parametrize("lanes", [1, 2])
def test_serving_lane_matches_alone(lanes):
    """A batching server: same-size requests share lanes (one, or two side by
    side) and join them as they arrive; every result equals the request's
    solve alone."""
    specs = [serving.StreamSpec(seed=41 + k, n=32, grains=20 + 10 * k,
                                model="evp" if k % 2 == 0 else "elastic",
                                direction=serving.DIRECTIONS[k], steps=4) for k in range(4)]
    specs.append(serving.StreamSpec(seed=50, n=48, grains=30, model="evp", direction="zz", steps=3))
    alone = [serving.build_stream_request(sp, keep_material=True).run() for sp in specs]
    print(f"serving lanes {lanes}", flush=True)
    server = serving.RveServer(hbm_budget=torch.cuda.mem_get_info()[0] // 2, max_slots=8,
                               batching=True, lanes_per_size=lanes)
    futs = [server.submit(serving.build_stream_request(sp, keep_material=True)) for sp in specs]
    got = [f.result(timeout=300) for f in futs]
    server.close()
    assert server.peak_in_flight >= 2
    for sp, a, g in zip(specs, alone, got):
        assert a["krylov"] == g["krylov"], sp
        assert np.abs(a["stress"] - g["stress"]).max() <= 1e-12 * np.abs(a["stress"]).max(), sp
