"""CPU FP64 oracle for the U(1) block-sparse contraction path.

TEST INFRASTRUCTURE ONLY.  This package restates, in numpy, the reference
algorithm (amiragha/mps, Julia) for the two-site DMRG hot path: sector
enumeration, fuse/split/permute, the sector-matched block product,
``contract``, ``_applymps2site``, the environment updates, ``svdtrunc`` and
the two-site sweep.  It is imported only by ``tests/``, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``,
and only as the checker.  The product path (``mps_amd``) never imports it.

Pinning (what keeps this restatement honest):
  * exact reference known-answer tests: ``test/symtensortests.jl:41-53``
    (split of a U(1) identity), ``:76-92`` (SymMatrix == fuselegs o fuselegs,
    also under a permutation), ``:94-110`` (integer contract == dense
    product), ``test/mpotests.jl:3-59`` (XXZ MPO == dense/ED Hamiltonian,
    per magnetisation block, exactly);
  * dense equivalence: every block-sparse contraction equals numpy einsum
    of the ``array()`` embeddings (``test/symtensortests.jl:112-137``);
  * physics closed forms: H_eff hermiticity on canonical environments,
    <psi|H_eff|psi> == measure(mps, mpo), DMRG energies == scipy ED.
Functions without such a check are marked "parity unpinned" in their
docstring (only the Lanczos trajectory, see ``tensornetalgs.eigsolve``).
"""
