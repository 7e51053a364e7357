"""CPU oracle of the rooftop-MoM iterative-solve path.

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
package, and only as the checker (or the timed CPU baseline). The product path (librooftop.so and the Python mirror in
rooftop-mom_amd/) never imports it and fails loudly when its HIP library is missing.

This package restates, in numpy/scipy FP64 (complex128), the algorithms of the reference MATLAB package rooftopMoM 3.5.0.
Every function cites the reference file:line it follows. Modules:

  quadrature  Gauss-Legendre (Golub-Welsch) and the reduced 2D moment rule.
  singular    closed-form touching moments of R^p and the extraction remainder.
  assembly    the MomentAssembler: near pass, far lattice pass, touching closed forms, accumulation into the three
              Toeplitz generators.
  spectra     circulant-embedding / subband spectra, the FFT and split-FFT products, the dense matrix by definition and
              the two-sided spatial blocks.
  solvers     COCG (plain and preconditioned), restarted left-preconditioned GMRES with CGS2 and complex Givens, and
              the IterativeMethod wrapper (zero right-hand side, true residual).
  excitation  plane wave, delta gap, middle port, input impedance and radar cross section.
  nearfield   the 3x3-cell block-Jacobi near-field preconditioner and the Jacobi diagonal. The reference has no
              near-field preconditioner; this one is defined by this project and pinned by construction (restriction
              of the dense matrix, inverse by LAPACK).

Pinning (what checks each function beyond its own formula) is recorded next to each function and exercised by
tests/test_oracle_*.py: reference.mat products and plane waves at 1e-13, exactMoments.csv at 1e-13/1e-12, the dense
matrix at 1e-12, the thesis self terms, the Taylor series of the remainder, brute-force tensor Gauss-Legendre on
separated cells, lattice == per-offset integration, grid-size independence, transpose symmetry, and backslash on the
synthetic solver systems. No function is "parity unpinned".
"""
