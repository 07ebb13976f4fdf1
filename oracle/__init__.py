"""CPU oracle for the palinstrophy DNS time step -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy + scipy.fft, the CPU branches of the
reference solver ``palinstrophy/turbo_simulator.py`` (each function cites the
reference lines it follows).  It is the checker for the HIP implementation in
``palinstrophy_amd``: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path never
calls into this package.

Precision-generic: ``dtype=np.float32`` reproduces the reference's FP32 path,
``dtype=np.float64`` is the FP64 restatement used as the parity oracle for the
FP64 GPU path (the reference itself has no FP64 mode).

Pinning status (how each part is checked, see ``tests/test_oracle_*.py``):
  * LCG / RANVEC / PAO amplitudes / phases: reference log KATs
    (seed-out 149, Q2, W2, time scale, CFLM0, DT0 at N=512, seed 1).
  * CNAB2 + NEXTDT + flow_eddy_metrics: reference 1001-step log trajectory
    (CFLM, DT, CN, T, TAU_L at 1e-5 relative).
  * FFT layer: scipy.fft is the reference's own FFT; checked against a
    brute-force DFT on tiny sizes.
  * Linear part (CNAB2, LS-IMEX-RK3): closed-form single-mode decay factors.
  * Nonlinear term: energy/enstrophy conservation invariants and a
    brute-force physical-space Jacobian on N=8.
  * KM3 init, LS-IMEX-RK3 trajectories, spectrum bins, field getters:
    golden vectors written by ``tests/golden/make_golden.py`` from the
    reference module, plus closed-form properties (energy normalisation,
    divergence, Hermitian kx=0 column).
"""
