"""FP64 CPU oracle for the VMC inner loop (TEST INFRASTRUCTURE ONLY).

This package restates, in NumPy / plain Python, the algorithm of the
reference's VMC hot path (sjdu10/vmc_torch):

  * lattice, Hubbard Hamiltonian, connected configurations and the
    Jordan-Wigner (symmray-order) sign            -> oracle.hubbard
  * Slater and neural-backflow amplitudes, exact enumeration,
    per-sample log-derivatives                     -> oracle.slater, oracle.nnbf
  * the batched exchange/hop Metropolis sweep      -> oracle.sampler
  * minSR / dense SR / MINRES and the SGD update   -> oracle.sr
  * integrated autocorrelation time and ESS        -> oracle.stats
  * Philox4x32-10 counter-based RNG                -> oracle.philox
  * point-group projection with fermionic signs    -> oracle.symmetry

Each function cites the reference file:line it follows (paths relative to
the reference checkout).

It is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it.  The product path (``vmc_torch_amd``) never imports it and
fails loudly when its HIP library is missing.

Parity status (what pins each part; see the tests named):
  hubbard   pinned: 2-site closed form, dense-H symmetry, spectrum equal to
            the netket-phase Hamiltonian, reference checkpoint energies
            (tests/test_oracle.py)
  slater    pinned: slogdet closed form, reference checkpoint energies via
            exact enumeration, finite differences for the log-derivative
  sampler   pinned: detailed balance / exact distribution on 2x2 by long
            chains; proposal table checked case by case
  sr        pinned: Woodbury identity (minSR == param-space SR), scipy
            MINRES on the same operator
  stats     pinned: AR(1) closed-form tau_int
  philox    pinned: Random123 known-answer vectors
  symmetry  pinned: the signed permutation operators form a representation
            of C4v / C2v and commute with the dense Hubbard H; ground states
            fall in one irrep; planted sign mistakes caught
            (tests/test_symmetry_oracle.py)
"""
