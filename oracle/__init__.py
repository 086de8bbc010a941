"""CPU oracle: a restatement of impurityModel's exact-diagonalization hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``impuritymodel_amd`` imports this
package; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and there only as the checker (or, for
``cpu_baseline``, as the timed CPU port).  The product path runs on the GPU and
fails loudly when its HIP library is missing.

Every function follows a reference function and cites it as
``impurityModel/ed/<file>.py:<lines>``.  Product states are Python ``int`` in
the reference's integer form (spin-orbital ``i`` of ``n`` at bit ``n-1-i``,
``create.py:79``), so ``int`` order equals the order of the reference's
``bytes`` keys (big-endian bitarray, ``product_state_representation.py:131-161``).

How each part is pinned (the tests named here live in ``tests/``):

* ``psr``, ``fermion``: the reference's own known-answer tests
  (``test_product_state_representation.py``, ``test_create.py``,
  ``test_remove.py``) and a Jordan-Wigner matrix brute force
  (``test_oracle_pins.py``).
* ``applyop``: the known-answer tests of ``test_applyOp.py`` and the
  Jordan-Wigner brute force, anticommutators and hermiticity.
* ``basis``: closed-form binomial counts (1770 = 45 + 500 + 1225 for the
  50-bath NiO layout, ``output.txt:19``) and itertools order.
* ``hamiltonian`` (fixed basis, BFS closure, CSR): dense Jordan-Wigner H
  projected on the basis; closure invariants (closed under H, restrictions
  hold, seeds first, layers ascending).
* ``lanczos``/``spectra``: exact resolvent from dense ``eigh`` when the Krylov
  space is the whole space, Hankel moments, and the sum rule
  ``int A dw = pi |T psi|^2``; the compositions (``get_spectra`` with its
  per-operator memo, the photoemission negation, thermal averages,
  operator sums) by the reference's own golden 50-bath run: ``XAS.dat``,
  ``NIXS.dat``, ``XPS.dat`` and ``PS.dat`` and every BFS basis size of its
  log (``test_refout_oracle.py``).
* ``eigen``: dense ``eigh`` brute force.
* ``average``: closed form of the Boltzmann average.
* ``spectra.rixs_map``: dense Kramers-Heisenberg amplitudes in the full Fock
  space of a small system (``test_oracle_pins.py``).
* ``mlp`` (the surrogate network, no reference counterpart): float64
  gradients against central finite differences; the BF16 emulation mode
  against the float64 mode within BF16 rounding.

* ``basis``/``eigen``/``observables`` on the golden run too: basis 1770,
  E0 = -903.8237, the E - E0 and expectation-value tables and the
  eigenvectors' product-state counts of ``output.txt``.

The golden run's inputs (the pickled h0 and the ``hOp``/``restrictions`` of
``data.npz``) are read by ``impuritymodel_amd.readers``' allow-list
unpickler, which builds nothing but numpy values, ordered dicts, frozensets
and complex numbers, and are kept as fixtures in ``tests/golden/``.
"""
