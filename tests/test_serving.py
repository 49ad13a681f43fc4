"""Request loop of the mixed-size serving stream: admission by HBM footprint,
slot limit, FIFO order and latency bookkeeping (host logic, CPU), and on the
GPU that every solve gives the same result alone and inside a concurrent mix."""

import threading
import time

import numpy as np
import pytest

from materialite_amd import serving


def _fake(duration, footprint, log, lock, tag):
    def run():
        with lock:
            log.append(("start", tag))
        time.sleep(duration)
        with lock:
            log.append(("end", tag))
        return tag
    return serving.Request(run=run, footprint=footprint, tag=tag)


def test_admission_respects_budget_and_slots():
    log, lock = [], threading.Lock()
    server = serving.RveServer(hbm_budget=100, max_slots=2)
    sizes = [60, 30, 50, 20, 90, 10]
    futs = [server.submit(_fake(0.05, s, log, lock, i)) for i, s in enumerate(sizes)]
    results = [f.result(timeout=10) for f in futs]
    server.close()
    assert results == list(range(len(sizes)))
    assert server.peak_in_flight <= 2
    assert server.peak_bytes <= 100
    # replay the log: bytes and slots in flight never exceed the limits
    in_flight, used = set(), 0
    for ev, tag in log:
        if ev == "start":
            in_flight.add(tag)
            used += sizes[tag]
            assert len(in_flight) <= 2 and used <= 100
        else:
            in_flight.remove(tag)
            used -= sizes[tag]
    # FIFO: requests start in arrival order
    starts = [tag for ev, tag in log if ev == "start"]
    assert starts == sorted(starts)


def test_oversized_request_rejected():
    server = serving.RveServer(hbm_budget=10, max_slots=1)
    with pytest.raises(ValueError):
        server.submit(serving.Request(run=lambda: None, footprint=11))
    server.close()


def test_errors_reach_the_future_and_latency_summary():
    server = serving.RveServer(hbm_budget=10, max_slots=2)

    def bad():
        raise RuntimeError("solve failed")
    f1 = server.submit(serving.Request(run=bad, footprint=1))
    f2 = server.submit(serving.Request(run=lambda: 7, footprint=1))
    with pytest.raises(RuntimeError):
        f1.result(timeout=10)
    assert f2.result(timeout=10) == 7
    server.close()
    summary = serving.latency_summary(server.completed)
    assert summary["count"] == 1 and summary["p99_s"] >= summary["p50_s"] >= 0.0


def test_footprint_grows_with_grid_and_plasticity():
    a = serving.footprint_bytes([32] * 3, "mixed", 0)
    b = serving.footprint_bytes([32] * 3, "mixed", 32 ** 3)
    c = serving.footprint_bytes([64] * 3, "mixed", 0)
    d = serving.footprint_bytes([32] * 3, "fp64", 0)
    assert a < b and a < c and a < d


def test_stream_specs_follow_the_workload():
    specs = [serving.draw_stream_spec(20264000 + j) for j in range(200)]
    assert {s.n for s in specs} == {32, 48, 64, 96}
    assert {s.model for s in specs} == {"elastic", "evp"}
    assert set(s.direction for s in specs) == set(serving.DIRECTIONS)
    assert all(50 <= s.grains <= 500 for s in specs)
    assert serving.draw_stream_spec(5) == serving.draw_stream_spec(5)


def test_shear_schedule_controls_one_component():
    ls = serving.stream_load_schedule("xz")
    d = ls.strain_increment(0.0, 0.5).components
    assert np.allclose(d, [0, 0, 0, 0, np.sqrt(2) * 0.5e-3, 0])
    assert list(ls.stress_mask) == [1, 1, 1, 1, 0, 1]


@pytest.mark.gpu
def test_results_same_alone_and_in_mix():
    import torch

    specs = [serving.StreamSpec(seed=11, n=16, grains=12, model="evp", direction="xx", steps=3),
             serving.StreamSpec(seed=12, n=24, grains=20, model="elastic", direction="xy", steps=3),
             serving.StreamSpec(seed=13, n=16, grains=8, model="evp", direction="yz", steps=3),
             serving.StreamSpec(seed=14, n=20, grains=15, model="elastic", direction="zz", steps=3)]
    alone = []
    for sp in specs:
        print(f"alone: {sp}", flush=True)
        alone.append(serving.build_stream_request(sp, keep_material=True).run())
    server = serving.RveServer(hbm_budget=torch.cuda.mem_get_info()[0] // 2, max_slots=4)
    reqs = [serving.build_stream_request(sp, keep_material=True) for sp in specs]
    print("mix: submitting", flush=True)
    futs = [server.submit(r) for r in reqs]
    mixed = [f.result(timeout=300) for f in futs]
    server.close()
    assert server.peak_in_flight >= 2
    for sp, a, m in zip(specs, alone, mixed):
        assert a["krylov"] == m["krylov"], sp
        scale = np.abs(a["stress"]).max()
        assert np.abs(a["stress"] - m["stress"]).max() <= 1e-12 * scale, sp


@pytest.mark.gpu
def test_footprint_estimate_bounds_device_memory():
    import torch

    from materialite_amd._native import NativeSolver

    for n, evp in [(32, False), (48, True)]:
        sp = serving.StreamSpec(seed=3, n=n, grains=10, model="evp" if evp else "elastic",
                                direction="xx", steps=1)
        req = serving.build_stream_request(sp)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        native = NativeSolver([n] * 3, "mixed")
        model = serving.config3_evp() if evp else serving.cubic_elastic()
        from materialite_amd.models.small_strain_fft.small_strain_fft import SmallStrainFFT
        SmallStrainFFT._register(native, model)
        rot = np.broadcast_to(np.eye(3), (n ** 3, 3, 3))
        native.set_microstructure(rot, np.zeros(n ** 3, np.int32))
        native.initialize()
        torch.cuda.synchronize()
        used = free0 - torch.cuda.mem_get_info()[0]
        native.close()
        print(f"n={n} evp={evp} used={used} estimate={req.footprint}", flush=True)
        assert used <= req.footprint
        assert used >= 0.5 * req.footprint
