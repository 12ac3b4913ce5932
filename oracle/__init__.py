"""CPU oracle for the MI355X cyclic modal solve.

TEST INFRASTRUCTURE ONLY: a numpy/scipy restatement of the reference
turbomodal algorithms (element integrals, assembly, cyclic reduction, static
prestress, shift-invert Lanczos).  It is the checker the HIP path is compared
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; no product code does.

Pinning (what fixes each function independently of this restatement) is
listed in the header of each module's tests: closed forms (consistent mass
matrix, D matrix values, Keast-14 exactness on polynomials), invariants
(rigid-body null space, patch tests, Hermitian/real K(k)), library routines
(scipy eigh/eigsh) and a full-annulus brute force on the reference's wedge
fixture.  Parity unpinned: the exact Lanczos start vector (the reference draws
it from libstdc++'s mt19937/normal_distribution with unspecified argument
evaluation order) and Spectra's iteration history for k = 0 and N/2 (Spectra
is absent from the reference tree); only converged eigenpairs are compared.
"""
