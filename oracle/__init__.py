"""CPU oracle for the RealQM grid-solver hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import anything from this package, and
only as the checker or as the timed CPU baseline -- never as part of the
product path.  The product (``realmolecule_amd``) never imports it and fails
loudly when its HIP library is missing.

Contents
--------
``s1``      restatement of ``cloud/solver.py`` (CuPy/CUDA engine): trial init,
            label-Neumann ITP update, Jacobi Poisson, per-domain normalise,
            nuclear-score boundary flip, energy, Hellmann-Feynman forces,
            K refresh and the damped nuclear integrator.
``mg``      the multigrid Poisson V-cycle (new work: the reference has at most a
            two-grid correction) and the exact FP64 coarsest-level solve.
``s2``      restatement of the ``molecule.js`` WGSL kernels on the hot path:
            level-set boundary update, damped smoother and two-grid correction,
            exclusion/dipole energy, sphere-averaged grad-P force, trial init,
            the per-atom fields (Z_nuc, initZeff, initRcut, angular splits),
            unique-nucleus E_KK, the frame potential with SIC and the page's
            direct per-electron Pother.
``adjoint`` the S1 fixed-step trajectory energy (frozen labels) in torch
            FP64, whose autograd gives dE/dR with respect to atom positions.
``cstep``   (``cstep.c`` + ``cstep.py``) the S1 step with the V-cycle in C +
            OpenMP: the all-core CPU baseline, pinned against ``s1`` + ``mg``.
``molfast`` restatement of the ``mol_fast.js`` fused per-electron engine:
            set-up in float64, the WGSL update, reduction, normalisation,
            per-electron two-grid correction, forces, K refresh and extract
            in FP32 operand order, the host loop's nuclear terms in float64.

Every function takes an explicit dtype: ``np.float64`` is the FP64 oracle and
``np.float32`` is the FP32-faithful mode, which performs the same IEEE
single-precision operations in the same association order as the HIP kernels
(compiled with ``-ffp-contract=off``) so that per-kernel results can be
compared bit for bit.

Parity pinning (what checks each function against something other than its
own formula) is listed in each module header; functions without such a check
say "parity unpinned".
"""
