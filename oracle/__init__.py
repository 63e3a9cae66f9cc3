"""CPU oracle for the EdelweissFE implicit static (NIST) Newton step.

TEST INFRASTRUCTURE ONLY.  This package restates, on the CPU, the reference
algorithm of the path that ``edelweissfe_amd`` implements on the GPU.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker or as the reported CPU
baseline; the product path never calls into it.

Modules and the reference files they follow (paths relative to the reference
checkout):

* ``rng``        project-owned SplitMix64 / Box-Muller streams for synthetic inputs.
* ``mesh``       edelweissfe/generators/boxgen.py:91-448 (node and element order,
                 node/element sets), plus the seeded interior-node jitter.
* ``element``    edelweissfe/elements/displacementelement/element.py:290-371 and
                 _elementcomputationmatrices.py:106-212,306-540,679-1060 (hexa8,
                 hexa20; quadrature tables edelweissfe/elements/library.py:212-291).
* ``materials``  edelweissfe/materials/linearelastic/linearelastic.py:95-210,
                 edelweissfe/materials/vonmises/vonmises.py:105-254; the
                 non-associative Drucker-Prager model has no reference
                 implementation (parity unpinned by the reference, pinned by the
                 textbook checks in tests/test_oracle_materials.py).
* ``assembly``   edelweissfe/numerics/dofmanager.py:259-300,445-471,522-557 and
                 edelweissfe/numerics/csrgenerator.pyx:47-115.
* ``nist``       edelweissfe/solvers/nonlinearimplicitstatic.py:110-458,559-725,947-986,
                 edelweissfe/timesteppers/adaptivetimestepper.py:90-183,
                 edelweissfe/steps/adaptivestep.py:56-113,
                 edelweissfe/stepactions/{dirichlet,nodeforces,bodyforce}.py.
* ``csrc``       C restatement of the element loop, the COO->CSR scatter and the
                 CPU PCG used as the timed CPU baseline (OpenMP over elements as
                 edelweissfe/solvers/nonlinearimplicitstaticparallelmk2.pyx:157-185).

Parity status of each function is recorded next to it.
"""
