"""CPU oracle for the correlation-energy hot path (TEST INFRASTRUCTURE ONLY).

This package is a plain-numpy restatement of the reference algorithms of
modrzejewski/beyond-rpa for the direct-RPA energy and its exchange (SOSEX)
correction.  It is the checker that the GPU library is compared against and
the timed CPU baseline of ``bench.py``.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product path (``brpa``) never does and fails loudly when the
HIP library is missing.

Modules and the reference routines each one restates:

* ``grid``     -- ``rpa_DaiHistogram``/``rpa_NCore`` (src/rpa/OrbDiffHist.f90:42-146),
                  ``rpa_OptQuadScaling``/``rpa_FreqQuadError``/``rpa_OptimizeFreqQuad``
                  (src/rpa/OptFreqQuad.f90:12-169), ``quad_GaussLegendre`` and
                  ``quad_CasimirPolder`` (src/rpa/quadratures.f90:418-506, 570-643).
* ``erpa``     -- ``rpa_GPiUG_MO`` (src/rpa/rpa_core_MO.f90:347-406),
                  ``rpa_CC_EcRPA_Analytic`` (src/rpa/rpa_CC.f90:412-471),
                  ``real_LogDet`` (src/linear-algebra/real_linalg.f90:793-825).
* ``t2``       -- ``rpa_CC_T2_v5`` and helpers (src/rpa/rpa_CC_Doubles.f90:80-91, 254-642),
                  ``rpa_CC_Diagonalize_PiU`` (src/rpa/rpa_CC.f90:1108-1129),
                  ``rpa_T2_EigRPA`` (src/rpa/rpa_CC_Doubles.f90:987-1009).
* ``sosex``    -- ``rpa_CC_Exchange_SOSEX``/``rpa_CC_EcExchange1``
                  (src/rpa/rpa_CC_Exchange.f90:36-116).
* ``adjoint``  -- dE_RPA/dR and dE_RPA/d(eps); no reference counterpart (the
                  reference has no RPA gradients).
* ``closed_forms`` -- independent closed forms used only to pin the above
                  (plasmon formula, direct MP2 limit, Riccati amplitudes,
                  lambda-integral identity of src/rpa/rpa_CC.f90:341-361).

Parity status of each function is written in its docstring ("pinned by ...").
"""
