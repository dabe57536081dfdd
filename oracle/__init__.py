"""CPU FP64 oracle for the MPS contraction hot path — TEST INFRASTRUCTURE ONLY.

This package restates, in numpy/LAPACK FP64, the algorithms of lcauser/TensorNetworks.jl
that the HIP library replaces (each function cites the reference file:line it
follows). It is imported only by tests/, __graft_entry__.smoke() and bench.py's CPU
legs (cpu_baseline and the CPU-port timings beside the GPU legs), to compare against —
never by the product library.

The reference is Julia and cannot run here (no Julia toolchain; its dependencies
TensorOperations.jl, KrylovKit.jl are absent), and its test suite holds no golden
vectors. Each oracle function is therefore pinned by checks the mathematics fixes:
brute-force dense contractions for tiny bond dimension, equality of the reference
contraction order with the optimal order, Hermiticity, LAPACK singular values,
dense exact diagonalisation and closed forms (critical transverse-field Ising:
E0(N) = 1 - csc(pi/(4N+2))). See tests/test_oracle.py.

Every function is pinned (tests/test_oracle.py, tests/test_oracle_pins.py,
tests/test_oracle_qjmc.py, tests/test_models.py). What stays "parity unpinned" is a
property of third-party code that cannot run here, not of a function of this package:
  * that KrylovKit keeps exactly one Ritz vector per restart at krylovdim <= 3 (taken from
    its published algorithm; krylov.eigsolve_ground's own per-cycle Ritz values are pinned
    as Rayleigh-Ritz values on the Krylov spaces K_k(H, x0), K_k(H, y1), ...);
  * Julia's global RNG stream (randomMPS, QJMC's rand()): inputs are drawn here instead.
"""
