"""CPU oracle for the lattice-Boltzmann elasticity / Darcy / Biot path.

TEST INFRASTRUCTURE ONLY.  This package is a plain torch-CPU restatement of the
reference solvers (s-lunowa/xlb-biot, ``src/elasticity.py``, ``src/darcy.py``,
``src/poroelasticity.py``, ``src/boundary_conditions.py``, ``src/lattice.py``,
``src/base.py``).  It follows the reference operation for operation, including
its quirks (duplicate-index scatters inside the halfway boundary conditions,
row-major mask pairing of the elastic anti-bounce-back, last-BC-wins, the
halo marked solid in the grid mask, ``t_max + 1`` steps in ``run``).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import it, and only as the checker (or as the timed CPU restatement of the
reference algorithm).  The product in ``xlb-biot_amd/src`` never imports it.

Arrays use the reference layout ``(nx, ny, q)`` (cell-major), in the dtype of the
precision policy (float64 by default).  Every function cites the reference
file:line it restates.

Parity pins (what checks each part; ``tests/test_oracle.py``,
``tests/test_bc_tables.py``, ``tests/test_orders.py``):

* lattice tables: closed-form constants, moment identities;
* transfer matrices: ``(m_to_f @ f_to_m)[:5] = I``;
* equilibrium/macroscopic: exact round trip;
* collision: one-cell brute force against the moment relaxation written out;
* streaming: ``numpy.roll`` brute force, per-direction sums conserved;
* restrict/prolong: 2x2 sum / injection, ``restrict(prolong(c)) = 4c``;
* V-cycle: ``levels=1`` reduces to ``_step``; residual equals ``|f - S(f)|``;
* elasticity: manufactured periodic solution converges at O(nx^-2); the
  reference drivers' Dirichlet walls at second order (displacement) and first
  (stress), traction walls at first order;
* Darcy: mass conservation, ``rho*w`` fixed point, analytic mode decay, O(nx^-2);
  the Dirichlet-wall driver at first order;
* gradient: exact on linear fields / spectral symbol of the D2Q9 stencil;
* boundary conditions: hand-derived wall tables (missing directions, their
  partners and the added traction / pressure terms) of the elastic and ADR
  anti-bounce-back conditions on 4 x 6 grids;
* Biot: the manufactured periodic solution (``examples/poroelasticity/periodic.py``)
  converges at O(dx^2), Terzaghi's consolidation (``Terzaghi.py``) at O(dx); a
  wrong sign planted in the lattice gradient breaks the former.

Fine-level Darcy wall indices applied to coarse multigrid arrays (multigrid + Biot +
Darcy walls) follow JAX's out-of-bounds semantics: clamp on gather, drop on scatter.
The clamp is pinned as irrelevant (tests/test_oracle_oob.py: the ADR wall rules are
row-local, so the gradient is bit-identical when the out-of-range rows gather NaN,
wrapped or zero indices instead, and cells named by no in-range row keep the streamed
stencil).  The drop is JAX's documented default for ``.at[].set`` ("The Sharp Bits",
out-of-bounds indexing); JAX is absent here, so no reference run checks that one
convention.
"""
