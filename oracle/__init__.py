"""CPU oracle for the sketch-and-factor path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (the ``randnla_amd``
package and its HIP library) imports, links or runs this package; only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, always as the checker, never as the thing measured or
shipped.

The oracle restates, in numpy, the reference Rust crate ``randblas`` 0.1.0
(divijkhaitan/randnla) and its vendored ``rust-random123``.  Every function
cites the reference file:line it follows.  The reference cannot be built here
(no Rust toolchain, crates.io dependencies not vendored), so ``oracle/_ref``
does not exist; the oracle is pinned instead by the reference's own
known-answer tests, closed forms and LAPACK cross-checks (see ``tests/``).

Parity status per function is listed in each module's header.
"""
