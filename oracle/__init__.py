"""CPU FP64 oracle for the TeNeS iTPS CTMRG path -- TEST INFRASTRUCTURE ONLY.

This package restates, in numpy (LAPACK through numpy.linalg), the algorithm
of the TeNeS functions the GPU library replaces, following the reference's
contraction order and control flow function by function (each function cites
the TeNeS file:line it follows).  It is the checker of the GPU library: only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import it,
and never as part of the product path.

Pinning (what checks each part besides retyped formulas):
  * rng.py          std::mt19937 + libstdc++ uniform_real_distribution,
                    pinned by the C++ standard's 10000th-output check and a
                    gcc-compiled probe (tests/test_oracle_goldens.py).
  * simple_update   test/data/simple_update.dat golden (G7) and the
                    TE_TFI end-to-end goldens.
  * ctm / contract  TE_TFI end-to-end goldens (SU + CTM chi=10 + observables,
                    complex) and the AntiferroHeisenberg_real goldens (with
                    full_update), closed-form product states, lattice-symmetry
                    and gauge invariants (tests/test_oracle_*.py).
  * full_update     test/data/full_update.dat golden (G6).
  * transfer        dense eigendecomposition of the full transfer matrix and
                    the correlation_length.dat goldens; the Arnoldi restarts
                    (no golden runs them) by the power-step, similarity,
                    Arnoldi-relation and convergence checks of
                    tests/test_oracle_arnoldi.py.
  * contract (clusters up to 4x4)  the contraction trees of the 1x1, 1x2,
                    2x1 and 2x2 clusters, a D = 1 closed form and the
                    correlation functions (tests/test_oracle_cluster.py);
                    multisite_obs_4.dat.
  * ctm_single / contract_density / finite_temperature  the FT_TFI_square
                    and FT_Kitaev end-to-end goldens (tests/test_oracle_ft.py).
  * adjoint         the contract.py tree values, central differences and the
                    closed form of an embedded product state
                    (tests/test_oracle_adjoint.py).
The mptensor randomized SVD (Use_RSVD) is not restated: its sampling is not
visible in the reference tree; the oracle always uses the exact SVD with the
Use_RSVD branch's cut.  That branch (cut and projector formulas, up-down
blocks) is pinned by the RSVD end-to-end goldens, whose sizes make the
randomized SVD exact; at sizes where it is not exact the randomized SVD itself
is parity unpinned.
"""
