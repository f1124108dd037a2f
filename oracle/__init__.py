"""CPU oracle for TORCH's Cartesian finite-volume step.

TEST INFRASTRUCTURE ONLY. This package is a from-scratch NumPy restatement of
the reference arithmetic (BFG508/TORCH, python/cfdSolver/*.py). It is the
checker the HIP path is compared against, and the CPU baseline timed beside
it; it is never the thing measured or shipped. Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it. The product path (``torch_amd``) never imports this package and
fails loudly when its HIP library is missing.

Modules and the reference code they restate:

- ``flux``     torchFlux.py:6-77, 125-248 (HLLC, grid driver)
- ``solver3d`` torchSolver3D.py:27-256 (ghosts, slopes, predictor, HLLC
               faces, update, first-order fallback, species, chemistry)
- ``solver2d`` torchSolver2D.py:38-91, 124-530, 647-667
- ``solver1d`` torchSolver.py:71-142, 164-404, 412-508, 534-700, 708-789,
               1038-1132 (legacy 1D scheme and kinetics)
- ``session``  torchSession.py:333-522 (Strang-split step driver)
- ``metric``   torchFiniteVolume.py:16-276, torchGeometry.py:93-213,
               torchBoundaries.py:17-246, torchFlux.py:80-122 (metric/face scheme)
- ``legacy``   torchSolver.py:918-1418, torchSolver2D.py:670-946,
               torchSolver3D.py:289-471 (stand-alone drivers, time histories)
- ``adaptive`` torchAdaptive.py:14-206 (adaptive leaves: face values, bounded
               gradients, rhs, cfl, advance)
- ``presets``  the reference's named initial conditions
- ``exact``    an exact Riemann solver (Toro, "Riemann Solvers and
               Numerical Methods for Fluid Dynamics", ch. 4); test oracle
               for the 1D/3D Sod profiles, independent of the reference.

Pinning (how each function is checked, besides retyping its formula):
golden vectors written by the reference itself (tests/golden/*.npz, made by
tests/golden/make_fixtures.py) cover hllc, hllc_grid, slopes of all five
limiters, the 3D step for 5 limiters x 4 boundary sets, variable gamma, both
fallback paths, species transport, both kinetic models (default and
nondimensional constants), the 2D step (5 limiters x 4 boundary sets incl.
Inflow) and its CFL, and multi-step sessions in 1D/2D/3D (inert, Arrhenius,
ZND, reactive dt retry). Closed forms and invariants (HLLC(U,U)=F(U),
contact transport of tangential momenta, uniform-state preservation,
symmetry of the explosion, the README's published L1 errors) are checked in
tests/test_oracle_*.py.
"""
