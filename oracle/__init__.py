"""CPU float64 oracle of the cdft Kohn-Sham hot path -- TEST INFRASTRUCTURE ONLY.

This package is the checker, never the product. It is a from-scratch numpy restatement of the
reference algorithm (cusp-factorised staggered divergence-form Hamiltonian, CheFSI, Rayleigh-Ritz,
cusp-aware quadrature, Coulomb-cutoff Poisson, VWN LDA, the SCF loop), each function citing the
reference ``file:line`` it follows. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it, and only to check or to time a CPU baseline. The GPU package
never imports it and has no CPU fallback.

How each part is pinned (anything not listed is covered by a test named in the module):

* ``core``: staggered/central FD weights against the closed forms of
  ``src/cdft/operators/divergence.py:44-47`` and ``laplacian.py:44-49``; ``D^T == D.T`` exactly on
  assembled matrices; the checkerboard Rayleigh quotients and self-adjointness order pinned in the
  reference's ``tests/golden/values.json``; ``W == -Z^2/2`` for one nucleus; identity factor gives
  the plain Laplacian.
* ``model``: the plain Laplacian exact on a cubic away from the zero halo; the sampled sine an
  eigenvector of the odd-reflection operator with its closed-form eigenvalue, which is the
  reference's recorded ``box_L10`` eigenvalue to 1.3e-14 (``tests/test_model.py``).
* ``eigen``: bare H at 33^3 against the reference golden eigenvalue ``-0.4999999999997764`` and
  kinetic energy ``0.5000000000415459``; Chebyshev recurrence against the explicit polynomial.
* ``quadrature``: ``integral f^2 = pi/Z^3`` (and the two-centre closed form), Lagrange reproduction
  of degree-7 polynomials, ``lift == interpolate^T``, ``E_H[1s] = 5Z/16``.
* ``physics``: Gaussian Hartree energy ``Q^2/(2 sqrt(pi) sigma)``; LDA derivative against finite
  differences and torch autograd of the energy density; the He LDA SCF on the 16-bohr box at
  h = 0.5 against the reference ladder value ``-2.830936240732186``.
"""
