"""Serving stream: concurrent RVE homogenization solves on one GPU.

A ``RveServer`` takes independent ``SmallStrainFFT`` solves as they arrive and
runs them with continuous batching:

* every admitted request owns a solver handle: its fields, tangents, Krylov
  vectors and hardening state stay resident in HBM for all of its load steps;
* a request is admitted (FIFO) as soon as a slot is free and its HBM footprint
  fits the budget left by the requests in flight; requests that arrive while
  others run join them without waiting for the batch to drain;
* with ``batching``, requests of the same grid size and precision
  (``Request.job`` set) share lanes: one worker thread per lane advances all
  of its requests one load increment per device call
  (mtl_solver_step_batch), so the MINRES iterations of every solve in the
  lane run as one launch of each kernel; a request joins a lane at the next
  increment and leaves it when its own schedule ends.  Up to
  ``lanes_per_size`` lanes of one size run side by side (each on its own
  streams: the device takes kernels from several hardware queues at once),
  and a new request goes to a new lane while that cap allows, else to the
  least loaded one;
* other requests run in worker threads of their own; the library's calls
  release the GIL (ctypes), so the device work of all lanes and threads
  overlaps on the GPU.

A solve's arithmetic does not depend on what else is in flight (per-solve
buffers, scalars and stopping tests, also inside a batch), so a result is the
same alone and in a mix.  Per-request times (arrival, admission, completion)
give the latency distribution reported beside throughput.

The reference has no serving layer; this module is the request loop that the
mixed-size stream workload runs on.
"""

import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

# HBM bytes per voxel of one solver handle (allocations in csrc/solver.hip,
# spectral.hip, minres.hip, constitutive.hip), h = nzc / nz ~ 1/2:
#   FP64 strain/stress fields and work vectors: 7 x 6 doubles + rotation 9 doubles
#   FP64 spectrum (residual projection): 6 complex doubles x h
#   matvec spectrum (FP32 forms only): 6 complex values x h, 8 bytes each
#     (mixed) or 12 (refined: FP32 value plus bf16 remainder per real)
#   MINRES: 2 Lanczos vectors (Mandel-6; 4, 6 or 8 bytes per value) and 4 w / x
#     vectors of the tangent type
#   tangent and committed tangent: 2 x 21 packed of the tangent type
_FIELDS = 7 * 6 * 8 + 9 * 8
# per point of a model: elastic index; EVP iteration, committed, pending and
# trial state (EvpModel in csrc/constitutive.hip): per slip system the slip
# shear of the committed, pending and trial states; per point 45 doubles
# (crystal strain, |guess|, 2 stresses, 2 slip increments, 3 x (plastic
# strain, accumulated slip, slip resistance)), a flag, the last iteration and
# the point index
_ELASTIC_POINT = 4
_EVP_POINT_PER_SYSTEM = 3 * 8
_EVP_POINT_FIXED = 45 * 8 + 1 + 4 + 4


def footprint_bytes(dims, precision="mixed", evp_points=0, num_slip_systems=12):
    """Upper estimate of the device bytes one solve holds (checked against the
    device's free-memory change in the GPU tests)."""
    n = int(np.prod(dims))
    nz = int(dims[2])
    h = ((nz + 1) // 2) / nz
    k = 8 if precision == "fp64" else 4  # tangent and w / x vectors
    lanczos = {"fp64": 8, "mixed": 4, "refined": 6}[precision]
    spec = {"fp64": 0, "mixed": 8, "refined": 12}[precision]
    per_voxel = _FIELDS + 6 * 16 * h + 6 * spec * h
    per_voxel += 2 * 6 * lanczos + 4 * 6 * k + 2 * 21 * k + _ELASTIC_POINT
    evp = evp_points * (_EVP_POINT_FIXED + num_slip_systems * _EVP_POINT_PER_SYSTEM)
    # allocation granularity (about 30 device allocations per handle) and
    # per-handle fixed buffers
    return int(1.05 * (n * per_voxel + evp)) + (32 << 20)


@dataclass
class Request:
    """One RVE solve: ``run()`` performs it (e.g. a bound ``SmallStrainFFT.run``).
    With ``job = (solver, material, run keywords)`` and ``finish`` (the
    returned Material -> the request's result) the server may batch it with
    other requests of its grid size and precision instead."""
    run: callable
    footprint: int
    tag: object = None
    job: tuple = None
    finish: callable = None
    arrival: float = None
    admitted: float = None
    finished: float = None
    result: object = None
    error: BaseException = None
    future: Future = field(default_factory=Future)

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class _Lane:
    """Same-size requests advanced together by one worker thread."""

    def __init__(self, key):
        self.key = key
        self.joining = []  # admitted, not yet started
        self.load = 0      # requests in the lane (joining or running)


class RveServer:
    """Continuous-batching request loop with admission by HBM footprint; with
    ``batching`` same-size requests share device launches (see the module
    docstring)."""

    def __init__(self, hbm_budget, max_slots=8, batching=False, lanes_per_size=4):
        if max_slots < 1:
            raise ValueError("max_slots must be >= 1")
        if lanes_per_size < 1:
            raise ValueError("lanes_per_size must be >= 1")
        self.batching = bool(batching)
        self.lanes_per_size = int(lanes_per_size)
        self._lanes = {}  # (dims, precision) -> running _Lane objects
        self.hbm_budget = int(hbm_budget)
        self.max_slots = int(max_slots)
        self._queue = deque()
        self._in_flight = 0
        self._bytes = 0
        self.peak_bytes = 0
        self.peak_in_flight = 0
        self._cv = threading.Condition()
        self._closed = False
        self._threads = []
        self.completed = []
        self._dispatcher = threading.Thread(target=self._dispatch, daemon=True)
        self._dispatcher.start()

    # --------------------------------------------------------------- client
    def submit(self, request):
        if request.footprint > self.hbm_budget:
            raise ValueError(f"request needs {request.footprint} bytes, budget is {self.hbm_budget}")
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            request.arrival = time.perf_counter() if request.arrival is None else request.arrival
            self._queue.append(request)
            self._cv.notify_all()
        return request.future

    def close(self, wait=True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            self._dispatcher.join()
            for t in list(self._threads):
                t.join()

    # --------------------------------------------------------------- loop
    def _fits(self, req):
        return (self._in_flight < self.max_slots
                and self._bytes + req.footprint <= self.hbm_budget)

    def _dispatch(self):
        while True:
            with self._cv:
                while not (self._queue and self._fits(self._queue[0])):
                    if self._closed and not self._queue:
                        return
                    self._cv.wait()
                req = self._queue.popleft()
                self._in_flight += 1
                self._bytes += req.footprint
                self.peak_bytes = max(self.peak_bytes, self._bytes)
                self.peak_in_flight = max(self.peak_in_flight, self._in_flight)
            req.admitted = time.perf_counter()
            if self.batching and req.job is not None:
                self._to_lane(req)
                continue
            t = threading.Thread(target=self._work, args=(req,), daemon=True)
            self._threads.append(t)
            t.start()

    # --------------------------------------------------------------- lanes
    @staticmethod
    def _lane_key(req):
        solver, material, _ = req.job
        return (tuple(int(d) for d in material.dimensions), solver.precision)

    def _to_lane(self, req):
        key = self._lane_key(req)
        with self._cv:
            lanes = self._lanes.setdefault(key, [])
            lane = min(lanes, key=lambda ln: ln.load, default=None)
            if lane is not None and (lane.load == 0 or len(lanes) >= self.lanes_per_size):
                lane.joining.append(req)  # picked up at the lane's next increment
                lane.load += 1
                return
            lane = _Lane(key)
            lane.joining.append(req)
            lane.load = 1
            lanes.append(lane)
        t = threading.Thread(target=self._lane_work, args=(lane,), daemon=True)
        self._threads.append(t)
        t.start()

    def _lane_work(self, lane):
        from .models.small_strain_fft.small_strain_fft import NativeSolver

        live = []  # (request, run)
        while True:
            with self._cv:
                joining, lane.joining = lane.joining, []
                if not joining and not live:
                    self._lanes[lane.key].remove(lane)
                    return
            for req in joining:
                solver, material, kw = req.job
                try:
                    live.append((req, solver.start(material, **kw)))
                except BaseException as e:
                    self._lane_done(lane, req, error=e)
            stepping = [(q, r) for q, r in live if not r.finished]
            if stepping:
                try:
                    params = [r.next_params() for _, r in stepping]
                    reps = NativeSolver.step_batch([r.native for _, r in stepping], params)
                except BaseException as e:  # every run of the failed device call ends with it
                    reps = None
                    for _, r in stepping:
                        r.stopped, r.error = True, e
                for (q, r), rp in zip(stepping, reps or []):
                    try:
                        r.apply(rp)
                    except BaseException as e:
                        r.stopped, r.error = True, e
            keep = []
            for q, r in live:
                if not r.finished:
                    keep.append((q, r))
                    continue
                err = getattr(r, "error", None)
                if err is not None:
                    r.close()
                    self._lane_done(lane, q, error=err)
                    continue
                try:
                    self._lane_done(lane, q, result=q.finish(r.result()))
                except BaseException as e:
                    self._lane_done(lane, q, error=e)
                finally:
                    r.close()
            live = keep

    def _lane_done(self, lane, req, result=None, error=None):
        with self._cv:
            lane.load -= 1
        self._done(req, result=result, error=error)

    def _done(self, req, result=None, error=None):
        req.result, req.error = result, error
        req.finished = time.perf_counter()
        with self._cv:
            self._in_flight -= 1
            self._bytes -= req.footprint
            self.completed.append(req)
            self._cv.notify_all()
        if error is not None:
            req.future.set_exception(error)
        else:
            req.future.set_result(result)

    def _work(self, req):
        try:
            result = req.run()
        except BaseException as e:  # reported through the future
            self._done(req, error=e)
            return
        self._done(req, result=result)


def latency_summary(requests):
    """p50/p99 latency (s), mean queueing delay, solves/s over finished
    requests (first arrival to last completion) and the offered rate (arrivals
    per second between the first and the last arrival)."""
    done = [r for r in requests if r.finished is not None and r.error is None]
    if not done:
        return {"count": 0}
    lat = np.array([r.latency for r in done])
    wait = np.array([r.admitted - r.arrival for r in done])
    span = max(r.finished for r in done) - min(r.arrival for r in done)
    arrivals = sorted(r.arrival for r in requests)
    window = arrivals[-1] - arrivals[0]
    return {
        "count": len(done),
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "mean_queue_s": float(wait.mean()),
        "solves_per_s": len(done) / span if span > 0 else float("inf"),
        "offered_rate_per_s": (len(arrivals) - 1) / window if window > 0 else float("inf"),
    }


# ------------------------------------------------------------------ workload
_SHEAR_SLOT = {"yz": 3, "xz": 4, "xy": 5}
DIRECTIONS = ("xx", "yy", "zz", "xy", "yz", "xz")


def stream_load_schedule(direction, rate=1.0e-3):
    """Uniaxial strain rate along a normal direction (stress-free laterally),
    or a single shear Mandel component sqrt(2) * rate with the other five
    stress components held at zero."""
    from .models.small_strain_fft import LoadSchedule
    from .tensor import Order2SymmetricTensor

    if direction in ("xx", "yy", "zz"):
        return LoadSchedule.from_constant_uniaxial_strain_rate(rate, direction[0])
    slot = _SHEAR_SLOT[direction]
    strain_rate = np.zeros(6)
    strain_rate[slot] = np.sqrt(2.0) * rate
    mask = np.ones(6)
    mask[slot] = 0.0
    return LoadSchedule.from_constant_rates(Order2SymmetricTensor(strain_rate),
                                            Order2SymmetricTensor.zero(), mask)


def config3_evp():
    """FCC crystal plasticity of the 128^3 workload: cubic C11, C12, C44 =
    250, 150, 120 GPa (in MPa), 12 octahedral systems, reference rate 1e-3/s,
    n = 20, g0 = 50 MPa, Voce tau_1 = 100, theta_0 = 500, theta_1 = 0."""
    from .models.small_strain_fft import ElasticViscoplastic, voce
    from .slip_system import SlipSystem
    from .tensor import Order4SymmetricTensor

    c = Order4SymmetricTensor.from_cubic_constants(250e3, 150e3, 120e3)
    return ElasticViscoplastic(c, SlipSystem.octahedral(), reference_slip_rate=1e-3,
                               rate_exponent=20.0, slip_resistance=50.0, hardening_function=voce,
                               hardening_properties={"tau_1": 100.0, "theta_0": 500.0,
                                                     "theta_1": 0.0})


def cubic_elastic():
    from .models.small_strain_fft import Elastic
    from .tensor import Order4SymmetricTensor

    return Elastic(Order4SymmetricTensor.from_cubic_constants(250e3, 150e3, 120e3))


def grain_material(n, grains, seed):
    """n^3 periodic-Voronoi polycrystal (voxel i at coordinate i, period n) with
    one Haar-random orientation per grain, both drawn from default_rng(seed)
    (grain seeds first, then orientations)."""
    from .material import Material
    from .microstructure import grain_rotations, periodic_voronoi
    from .tensor import Orientation

    rng = np.random.default_rng(seed)
    grain_of_point, _ = periodic_voronoi([n] * 3, grains, rng)
    rot = grain_rotations(grains, rng)
    m = Material(dimensions=[n] * 3).create_fields({"grain": grain_of_point})
    return m.create_regional_fields(
        "grain", {"grain": np.arange(grains), "orientation": Orientation(rot, "p")})


def two_phase_material(n=256, grains=1000, seed=20260005, radius=8.0, volume_fraction=0.10):
    """The two-phase workload: a periodic-Voronoi FCC matrix (config3_evp) with
    randomly placed non-overlapping periodic spheres of the given radius up to
    the volume fraction, elastic and isotropic with bulk and shear moduli 10x the
    matrix's Voigt averages (phase 1).  Draws grains, orientations, then sphere
    centres from default_rng(seed).  Returns (material, number of spheres)."""
    import pandas as pd

    from .material import Material
    from .microstructure import grain_rotations, periodic_spheres, periodic_voronoi
    from .models.small_strain_fft import Elastic
    from .tensor import Order4SymmetricTensor, Orientation

    rng = np.random.default_rng(seed)
    grain_of_point, _ = periodic_voronoi([n] * 3, grains, rng)
    rot = grain_rotations(grains, rng)
    inclusion, centres = periodic_spheres([n] * 3, radius, volume_fraction, rng)
    c11, c12, c44 = 250e3, 150e3, 120e3
    k = 10 * (c11 + 2 * c12) / 3
    mu = 10 * (c11 - c12 + 3 * c44) / 5
    e = 9 * k * mu / (3 * k + mu)
    m = Material(dimensions=[n] * 3).create_fields(
        {"grain": grain_of_point, "phase": inclusion.astype(np.int64)})
    m = m.create_regional_fields("grain", {"grain": np.arange(grains),
                                           "orientation": Orientation(rot, "p")})
    table = pd.DataFrame({"phase": [0, 1], "constitutive_model": [
        config3_evp(), Elastic(Order4SymmetricTensor.from_isotropic_constants(e, mu))]})
    return m.create_regional_fields("phase", table), len(centres)


@dataclass
class StreamSpec:
    """One request of the mixed-size stream, drawn from its own seed."""
    seed: int
    n: int
    grains: int
    model: str       # "elastic" or "evp"
    direction: str
    steps: int = 20
    dt: float = 0.5


def draw_stream_spec(seed, sizes=(32, 48, 64, 96), grains=(50, 500), steps=20, dt=0.5):
    rng = np.random.default_rng(seed)
    n = int(rng.choice(sizes))
    g = int(rng.integers(grains[0], grains[1] + 1))
    model = "evp" if rng.random() < 0.5 else "elastic"
    direction = DIRECTIONS[int(rng.integers(len(DIRECTIONS)))]
    return StreamSpec(seed=seed, n=n, grains=g, model=model, direction=direction, steps=steps, dt=dt)


def build_stream_request(spec, precision="refined", keep_material=False):
    """Material (periodic Voronoi, Haar-random orientations from the spec's
    seed), model and schedule of one request; returns a Request whose run()
    performs the solve and returns the volume-averaged stress (Mandel, FP64)
    and the Krylov iteration count, plus the stress field when asked."""
    from .material import Material
    from .microstructure import grain_rotations, periodic_voronoi
    from .models.small_strain_fft import SmallStrainFFT
    from .tensor import Orientation

    rng = np.random.default_rng(spec.seed + 1)
    grain_of_point, _ = periodic_voronoi([spec.n] * 3, spec.grains, rng)
    rot = grain_rotations(spec.grains, rng)
    material = Material(dimensions=[spec.n] * 3).create_fields({"grain": grain_of_point})
    material = material.create_regional_fields(
        "grain", {"grain": np.arange(spec.grains), "orientation": Orientation(rot, "p")})
    model = config3_evp() if spec.model == "evp" else cubic_elastic()
    solver = SmallStrainFFT(stream_load_schedule(spec.direction), end_time=spec.steps * spec.dt,
                            initial_time_increment=spec.dt, max_time_increment=spec.dt,
                            constitutive_model=model, precision=precision)

    def finish(out):
        stress = out.extract("stress").components
        res = {"mean_stress": stress.mean(axis=0), "krylov": solver.stats["krylov_iterations"],
               "steps": solver.stats["steps"]}
        if keep_material:
            res["stress"] = stress
        return res

    def run():
        return finish(solver.run(material))

    evp_points = spec.n ** 3 if spec.model == "evp" else 0
    return Request(run=run, footprint=footprint_bytes([spec.n] * 3, precision, evp_points),
                   tag=spec, job=(solver, material, {}), finish=finish)
