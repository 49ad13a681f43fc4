"""CPU FP64 oracle for the SmallStrainFFT hot path — TEST INFRASTRUCTURE ONLY.

This package is a NumPy restatement of materialite's ``SmallStrainFFT`` solver
(``materialite/models/small_strain_fft/*.py``) used as the checker for the HIP
implementation in ``materialite_amd``.  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg may import it, and only as the checker or
as the timed CPU baseline.  The product path never imports it and has no CPU
fallback.

Pinning (how each function is checked, see ``tests/test_oracle_*.py``):

* ``tensors``       Mandel algebra / rotations / stiffness constructors: textbook
                    identities (Voigt->Mandel sqrt(2) scaling, cartesian 4th-order
                    rotation), reference known answers (directional modulus
                    183.0467131980609, seeded Euler angles, octahedral systems).
* ``projection``    closed-form Ghat vs. the literal complex-arithmetic build of the
                    reference; symmetry, idempotence, trace 3, evenness, Nyquist
                    zeros, Ghat[0] = diag(mask); brute-force DFT on tiny grids.
* ``minres``        iterate-for-iterate against ``scipy.sparse.linalg.minres``.
* ``constitutive``  closed-form hardening identities; single-crystal closed forms;
                    end-to-end reference goldens (EVP 8^3: 361.276725, 541.043871,
                    539.345343, 442.608112, 426.923816/428.328875; J2 8^3: 150.0,
                    150.3980086, 171.694205 and the reference's step relations
                    between axial stress and equivalent plastic strain; J2
                    two-phase 150.3980086).
                    Per-system initial slip resistances: a system whose
                    resistance is effectively infinite never slips, so the model
                    equals the one without those systems; equal per-system values
                    equal the scalar model.
* ``solver``        end-to-end reference goldens (elastic: 6.5890252, 6.0,
                    4.547568053, homogeneous closed forms, periodic-shift
                    invariance); the load-schedule factories against the
                    reference's known answers (tests/test_load_schedule.py).
* ``microstructure`` reference known answers for seeded Voronoi region ids and
                    Haar-random orientations.
"""
