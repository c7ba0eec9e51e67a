"""FP64 CPU oracle of the Dirichlet-splat hot path (test infrastructure only).

This package restates, in numpy float64, the arithmetic of the reference kernels
(Asixa/Dirichlet-Splatting, `dsplat/kernels/*.slang`) and of its float64 test
oracles (`tests/reference/*.py`, `dsplat/groundtruth.py`). It is the checker of the
HIP path, never the thing measured or shipped: only `tests/`,
`__graft_entry__.smoke()` and the `cpu_baseline` leg of `bench.py` may import it.
The product path (`dsplat`) never imports it and fails loudly when its HIP
library is missing.

Modules
    kernel    Dirichlet kernel D(d), D'(d), windowed kernels, brute-force DFT sums
    surfel    scattering, chirp carrier, FMCW columns, render/JVP/VJP, design,
              Jacobian, correlations, certificate, ADC beat-signal synthesis
    spectral  separable windowed atoms on 1-3-D grids
    search    range gate, greedy separated peaks, quasi-Newton direction/trial
    linalg    dense normal equations, energy, FP64 range responses

Every function cites the reference file:line it follows. How each one is pinned
(brute force, closed form, physics, finite differences, adjoint identity) is
listed in the module docstrings and exercised by tests/test_oracle*.py.
"""
