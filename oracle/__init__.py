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
  ``int A dw = pi |T psi|^2``.
* ``eigen``: dense ``eigh`` brute force.
* ``average``: closed form of the Boltzmann average.
* ``spectra.rixs_map``: dense Kramers-Heisenberg amplitudes in the full Fock
  space of a small system (``test_oracle_pins.py``).
* ``mlp`` (the surrogate network, no reference counterpart): float64
  gradients against central finite differences; the BF16 emulation mode
  against the float64 mode within BF16 rounding.

The reference's own pickled data (``h0/*.pickle``, the object arrays of
``referenceOutput/.../data.npz``) is not loaded: the safe loader
(``numpy.load(allow_pickle=False)``) refuses it.  The end-to-end golden spectra
of the 50-bath run therefore cannot be reproduced here; see the tests for the
pins used instead.
"""
