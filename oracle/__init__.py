"""Oracle for the table-mode Kaiser-Bessel NUFFT path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (``torchkbnufft_amd``) imports, calls or links anything in
this package.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker (or as the timed CPU
baseline), never as the thing measured or shipped.

Contents
--------
``table_path``  CPU restatement of torchkbnufft's table-mode path in torch CPU ops,
                following the reference op order (including its float32 index
                arithmetic) so complex64 results match the reference's.
                Follows ``torchkbnufft/_nufft/interp.py:88-203,315-403,406-726``,
                ``torchkbnufft/_nufft/fft.py:9-118`` and
                ``torchkbnufft/modules/kbnufft.py:165-228,347-410``.
``kb``          FP64 closed forms: Kaiser-Bessel kernel, its table (with the
                reference's linear phase), its Fourier transform (apodisation),
                and the kernel derivative.  Independent of the reference's
                sparse-matrix construction (``_nufft/utils.py:16-254``).
``nudft``       FP64 direct non-uniform DFT, its adjoint and its derivative with
                respect to the trajectory, plus an exact-KB (no table) FP64
                interpolation operator used for finite differences.

Parity pins (what checks each function; see ``tests/test_oracle.py``)
---------------------------------------------------------------------
* ``kb.kb_table`` / ``kb.apodization``: the reference's own ``table_i`` and
  ``scaling_coef`` buffers written by ``tests/golden/make_golden.py``; spot values
  (``table[J*L/2] == 1``, ``table[0] == table[J*L] == 0``); numpy ``i0``.
* ``table_path`` forward/adjoint/NUFFT: golden vectors from the reference in
  complex64 and float64 (``tests/golden/*.npz``); dot-product adjointness;
  FP64 NUDFT within the table-quantisation error; dense brute force on tiny sizes.
* ``nudft``: samples on the uniform grid reduce it to numpy's FFT times the shift
  phase, and its adjoint to the unnormalised inverse FFT (``test_nudft_on_grid_is_shifted_fft``,
  1e-12); ``nudft_domega`` against FP64 central differences of ``nudft``.
* ``nudft.exact_kb_interp``: agreement with the table path to the table-quantisation
  error (``test_table_interp_matches_exact_kb``).
* ``nudft.exact_kb_interp_domega``: per-coordinate FP64 central differences of
  ``exact_kb_interp`` in 1-D/2-D/3-D away from footprint edges (1e-6,
  ``test_exact_kb_domega_matches_fd``); the directional finite-difference protocol of
  the differentiation workload against it (``test_directional_fd_protocol_matches_analytic_gradient``).
"""
