"""Oracle for the periodic-box P(k)/B(k) estimator -- TEST INFRASTRUCTURE.

This package is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product path (``pyspectrum_amd``) never imports, links or runs
anything here and fails loudly when its HIP library is missing.

Contents
--------
``spectrum``  FP64 numpy restatement of the reference's periodic-box path
              and survey monopole (``pyspectrum/pyspectrum.py``,
              ``pyspectrum/estimator.f``, ``pyspectrum/util.py``); every
              function cites the reference lines it follows.
``ref``       ctypes binding to ``oracle/_ref/libestimator.so``, the reference's
              own Fortran kernels (``assign_quad``, ``fcomb_periodic``,
              ``fcomb_survey``, ``pk_pbox_rsd``, ``FiveDelta2g_1/2``,
              ``build_quad``) compiled from
              ``/root/reference`` by ``oracle/Makefile``.  Used only to pin
              ``spectrum`` in tests.

Pins (how each oracle function is checked; see ``tests/test_oracle.py``,
``tests/test_survey_oracle.py`` and ``tests/test_quad_oracle.py``):

* ``assign_quad``       -- the reference Fortran (``ref``), mass/normalisation
                           invariants, the 1-4-1 node stencil.
* ``fft_packed``        -- textbook DFT (direct sum) on tiny grids.
* ``fcomb_periodic``    -- the reference Fortran (``ref``) on random grids,
                           closed form for one particle, delta(k=0) = 1.
* ``reflect_delta``     -- Hermitian symmetry, brute-force index mapping.
* ``shell_index``       -- the float formula at the reference lines, shell
                           sizes 370/1402/3154 (step 3).
* ``counts_Bk123``      -- the reference's FP64 count caches (committed
                           fixtures), integer brute force on tiny grids.
* ``bk_from_shells``    -- direct k-space triple sum on tiny grids.
* ``pk_pbox_rsd``       -- the reference Fortran (``ref``), bit-exact bins.
* ``Pk_periodic``       -- Parseval identity against the shell fields.
* ``assign_quad(offset)`` -- the reference Fortran with offset Ngrid/2 and
                           negative (observer-centred) positions.
* ``fcomb_survey``      -- the reference Fortran ``fcomb_survey``.
* ``comoving_distance`` -- closed forms (Omega_m = 1 and 0).
* ``radecz_to_cartesian`` -- axis points.
* ``B0_survey_from_sums`` -- closed form when data = randoms (delta = 0).
* ``FFT_survey_mono`` / ``B0_survey`` -- compositions of the pinned pieces.
* ``five_delta2g_1/2``, ``build_quad`` -- the reference Fortran bit for bit,
                           closed forms (factor 5 / -2.5, equal inputs).
* ``quad_weights``      -- the reference Fortran ``assign_quad`` with
                           (ia, ib, ic, id) != 0, traces sum to w.
"""
