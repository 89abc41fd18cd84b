!
! brpa_mod -- Fortran drop-in for the correlation-energy hot path.
!
! Keeps the names and argument lists of the reference routines so that
! reference host code can call the gfx950 library through ISO_C_BINDING:
!
!   rpa_GPiUG_MO            rpa/rpa_core_MO.f90:347-406
!   rpa_CC_EcRPA_Analytic   rpa/rpa_CC.f90:412-471
!   rpa_CC_T2               rpa/rpa_CC_Doubles.f90:43-77
!   rpa_THC_CC_T2           rpa/rpa_CC_Doubles.f90:645-984 (with rpa_T2_EigRPA 987-1009)
!   rpa_THC_PiU             rpa/rpa_THC.f90:215-282
!   rpa_CRG                 rpa/rpa_core_MO.f90:233-301 (one image: NSubsets(2) = 1)
!   rpa_G                   rpa/rpa.f90:31-173
!   rpa_CC_Exchange_MBPT3_1 rpa/rpa_CC_Exchange.f90:232-270
!   rpa_CC_Exchange_MBPT3_2 rpa/rpa_CC_Exchange.f90:381-393
!   rpa_CC_Exchange_SOSEX   rpa/rpa_CC_Exchange.f90:36-97
!   rpa_DaiHistogram        rpa/OrbDiffHist.f90:42-123 (with rpa_NCore 126-146)
!   rpa_DaiMaxThresh        rpa/OrbDiffHist.f90:10-39
!   rpa_OptQuadScaling      rpa/OptFreqQuad.f90:75-122
!   rpa_OptimizeFreqQuad    rpa/OptFreqQuad.f90:12-72
!
! Built with -fdefault-integer-8 like every reference build (default INTEGER
! is 64-bit, matching int64_t of the C ABI).  Arrays are passed as contiguous
! column-major storage; Fortran CRG(P, ai) has the bytes of C R[ai][P].
! The reference has no return codes: it prints a message and stops; a
! non-zero status from the library does the same here.  The library context
! (one GPU, one stream) is created on first use and kept for the run.
!
module brpa_mod
      use iso_c_binding
      implicit none
      private

      integer, parameter :: F64 = c_double
      type(c_ptr), save :: ctx = c_null_ptr

      public :: brpa_init, brpa_finalize
      public :: rpa_GPiUG_MO, rpa_CC_EcRPA_Analytic, rpa_CC_T2, rpa_CC_Exchange_SOSEX
      public :: rpa_THC_CC_T2, rpa_THC_PiU
      public :: rpa_CC_Exchange_MBPT3_1, rpa_CC_Exchange_MBPT3_2, rpa_CRG, rpa_G
      public :: rpa_DaiHistogram, rpa_DaiMaxThresh, rpa_OptQuadScaling, rpa_OptimizeFreqQuad
      ! T2CutoffType values (rpa/rpa_definitions.f90:139-142)
      integer, parameter, public :: RPA_T2_CUTOFF_EIG = 0
      integer, parameter, public :: RPA_T2_CUTOFF_EIG_DIV_MAXEIG = 1
      integer, parameter, public :: RPA_T2_CUTOFF_EIG_DIV_NELECTRON = 2
      integer, parameter, public :: RPA_T2_CUTOFF_SUM_REJECTED = 3
      public :: brpa_ecrpa_fused

      interface
            integer(c_int) function c_ctx_create(device, out) bind(C, name="brpa_ctx_create")
                  import :: c_int, c_ptr
                  integer(c_int), value :: device
                  type(c_ptr) :: out
            end function
            integer(c_int) function c_ctx_destroy(c) bind(C, name="brpa_ctx_destroy")
                  import :: c_int, c_ptr
                  type(c_ptr), value :: c
            end function
            type(c_ptr) function c_last_error(c) bind(C, name="brpa_last_error")
                  import :: c_ptr
                  type(c_ptr), value :: c
            end function
            integer(c_int) function c_gpiug_mo(c, gpiug, crg, naux, eocc, nocc, evir, nvir, &
                  freq, pref) bind(C, name="brpa_gpiug_mo")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  real(c_double) :: gpiug(*)
                  real(c_double) :: crg(*)
                  integer(c_int64_t), value :: naux, nocc, nvir
                  real(c_double) :: eocc(*), evir(*)
                  real(c_double), value :: freq, pref
            end function
            integer(c_int) function c_ecrpa_analytic(c, e, piu, freqs, weights, nfreqs, naux, eu) &
                  bind(C, name="brpa_ecrpa_analytic")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  real(c_double) :: e
                  real(c_double) :: piu(*), freqs(*), weights(*)
                  integer(c_int64_t), value :: nfreqs, naux
                  type(c_ptr), value :: eu
            end function
            integer(c_int) function c_ecrpa(c, prec, r, naux, eocc, nocc, evir, nvir, freqs, &
                  weights, nfreqs, e, eu) bind(C, name="brpa_ecrpa")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  integer(c_int), value :: prec
                  real(c_double) :: r(*), eocc(*), evir(*), freqs(*), weights(*)
                  integer(c_int64_t), value :: naux, nocc, nvir, nfreqs
                  real(c_double) :: e
                  type(c_ptr), value :: eu
            end function
            integer(c_int) function c_t2_compute(c, prec, r, naux, eocc, nocc, evir, nvir, freqs, &
                  weights, nfreqs, lambda, cutoff, nguess, seed, nvecs) bind(C, name="brpa_t2_compute")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  integer(c_int), value :: prec
                  real(c_double) :: r(*), eocc(*), evir(*), freqs(*), weights(*)
                  integer(c_int64_t), value :: naux, nocc, nvir, nfreqs, nguess, seed
                  real(c_double), value :: lambda, cutoff
                  integer(c_int64_t) :: nvecs
            end function
            integer(c_int) function c_t2_copyout(c, a, v) bind(C, name="brpa_t2_copyout")
                  import :: c_int, c_ptr, c_double
                  type(c_ptr), value :: c
                  real(c_double) :: a(*), v(*)
            end function
            integer(c_int) function c_t2_cutoff(c, ctype, thresh, smooth, steep, nocc, common, nvecs, &
                  nscaled) bind(C, name="brpa_t2_cutoff")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  integer(c_int), value :: ctype, smooth
                  real(c_double), value :: thresh, steep
                  integer(c_int64_t), value :: nocc
                  real(c_double) :: common
                  integer(c_int64_t) :: nvecs, nscaled
            end function
            integer(c_int) function c_thc_piu(c, piu, rkai, xga, xgi, zgk, ngrid, nocc, nvir, naux, &
                  eocc, evir, freqs, nfreqs, pref, accum) bind(C, name="brpa_thc_piu")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  real(c_double) :: piu(*)
                  type(c_ptr), value :: rkai
                  real(c_double) :: xga(*), xgi(*), zgk(*), eocc(*), evir(*), freqs(*)
                  integer(c_int64_t), value :: ngrid, nocc, nvir, naux, nfreqs
                  real(c_double), value :: pref
                  integer(c_int), value :: accum
            end function
            integer(c_int) function c_mbpt3_1(c, prec, e, v, a, nt, r, naux, eocc, evir, nocc, nvir) &
                  bind(C, name="brpa_mbpt3_1")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  integer(c_int), value :: prec
                  real(c_double) :: e
                  real(c_double) :: v(*), a(*), r(*), eocc(*), evir(*)
                  integer(c_int64_t), value :: nt, naux, nocc, nvir
            end function
            integer(c_int) function c_mbpt3_2(c, e, r, v, a, nocc, nvir, naux, nt) &
                  bind(C, name="brpa_mbpt3_2")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  real(c_double) :: e
                  real(c_double) :: r(*), v(*), a(*)
                  integer(c_int64_t), value :: nocc, nvir, naux, nt
            end function
            integer(c_int) function c_ao_mo(c, crg, rg, nvecs, npq, pairs, occ, vir, nao, nocc, nvir, &
                  accum) bind(C, name="brpa_ao_mo")
                  import :: c_int, c_ptr, c_int64_t, c_double, c_int32_t
                  type(c_ptr), value :: c
                  real(c_double) :: crg(*), rg(*), occ(*), vir(*)
                  integer(c_int64_t), value :: nvecs, npq, nao, nocc, nvir
                  integer(c_int32_t) :: pairs(*)
                  integer(c_int), value :: accum
            end function
            integer(c_int) function c_rpa_g(c, basis, g, rwr, n, target, seed, nvecs, abserr) &
                  bind(C, name="brpa_rpa_g")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  integer(c_int), value :: basis
                  real(c_double) :: g(*), rwr(*)
                  integer(c_int64_t), value :: n, seed
                  real(c_double), value :: target
                  integer(c_int64_t) :: nvecs
                  real(c_double) :: abserr
            end function
            integer(c_int) function c_t2_eig_rpa(c, r, naux, eig) bind(C, name="brpa_t2_eig_rpa")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  real(c_double) :: r(*), eig(*)
                  integer(c_int64_t), value :: naux
            end function
            integer(c_int) function c_t2_copyout_ov(c, a, v) bind(C, name="brpa_t2_copyout_ov")
                  import :: c_int, c_ptr, c_double
                  type(c_ptr), value :: c
                  real(c_double) :: a(*), v(*)
            end function
            integer(c_int) function c_sosex(c, prec, e, alpha, v, a, r, nocc, nvir, naux, nvecs, s2) &
                  bind(C, name="brpa_sosex")
                  import :: c_int, c_ptr, c_int64_t, c_double
                  type(c_ptr), value :: c
                  integer(c_int), value :: prec
                  real(c_double) :: e
                  real(c_double), value :: alpha
                  real(c_double) :: v(*), a(*), r(*)
                  integer(c_int64_t), value :: nocc, nvir, naux, nvecs
                  integer(c_int), value :: s2
            end function
            integer(c_int) function c_dai_histogram(eocc, nocc, evir, nvir, core, cap, vals, wts) &
                  bind(C, name="brpa_dai_histogram")
                  import :: c_int, c_int64_t, c_double
                  real(c_double), intent(in) :: eocc(*), evir(*)
                  integer(c_int64_t), value :: nocc, nvir
                  real(c_double), value :: core, cap
                  real(c_double), intent(out) :: vals(*), wts(*)
            end function
            integer(c_int) function c_dai_max_thresh(nspins, eocc, nocc, evir, nvir, core, out) &
                  bind(C, name="brpa_dai_max_thresh")
                  import :: c_int, c_int64_t, c_double
                  integer(c_int64_t), value :: nspins
                  real(c_double), intent(in) :: eocc(*), evir(*)
                  integer(c_int64_t), intent(in) :: nocc(*), nvir(*)
                  real(c_double), value :: core
                  real(c_double), intent(out) :: out
            end function
            integer(c_int) function c_opt_quad_scaling(nfreqs, nsets, vals, wts, freqs, weights, &
                  info) bind(C, name="brpa_opt_quad_scaling")
                  import :: c_int, c_int64_t, c_double
                  integer(c_int64_t), value :: nfreqs, nsets
                  real(c_double), intent(in) :: vals(*), wts(*)
                  real(c_double), intent(out) :: freqs(*), weights(*), info(3)
            end function
            integer(c_int) function c_optimize_freq_quad(nsets, vals, wts, terr, trel, nmin, nmax, &
                  freqs, weights, nfreqs, info) bind(C, name="brpa_optimize_freq_quad")
                  import :: c_int, c_int64_t, c_double
                  integer(c_int64_t), value :: nsets, nmin, nmax
                  real(c_double), intent(in) :: vals(*), wts(*)
                  real(c_double), value :: terr, trel
                  real(c_double), intent(out) :: freqs(*), weights(*), info(3)
                  integer(c_int64_t), intent(out) :: nfreqs
            end function
            integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
                  import :: c_size_t, c_ptr
                  type(c_ptr), value :: s
            end function
      end interface

contains

      subroutine brpa_init(device)
            integer, optional, intent(in) :: device
            integer(c_int) :: d, rc
            if (c_associated(ctx)) return
            d = 0
            if (present(device)) d = int(device, c_int)
            rc = c_ctx_create(d, ctx)
            if (rc /= 0) then
                  print '(a,i0)', "brpa: cannot create a GPU context, status ", rc
                  error stop
            end if
      end subroutine brpa_init


      subroutine brpa_finalize()
            integer(c_int) :: rc
            if (.not. c_associated(ctx)) return
            rc = c_ctx_destroy(ctx)
            ctx = c_null_ptr
      end subroutine brpa_finalize


      subroutine check_grid(rc, where)
            ! the grid routines are host code and need no context: their only
            ! failures are the reference's own error stops
            integer(c_int), intent(in) :: rc
            character(*), intent(in) :: where
            if (rc == 0) return
            if (rc == 4) then
                  print '(a,a)', where, ": Found invalid orbital energy gap. Cannot decompose energy denominators."
            else if (rc == 7) then
                  print '(a,a)', where, ": Quadrature did not converge to specified accuracy"
            else
                  print '(a,a,i0)', where, ": error ", rc
            end if
            error stop
      end subroutine check_grid


      subroutine check(rc, where)
            integer(c_int), intent(in) :: rc
            character(*), intent(in) :: where
            character(kind=c_char), pointer :: msg(:)
            type(c_ptr) :: p
            integer(c_size_t) :: n
            if (rc == 0) return
            p = c_last_error(ctx)
            n = c_strlen(p)
            call c_f_pointer(p, msg, [n])
            print '(a,a,i0,a,*(a))', where, ": error ", rc, ": ", msg
            error stop
      end subroutine check


      subroutine rpa_GPiUG_MO(GPiUG, CRG, DCRG, OccEnergies, VirtEnergies, NOcc, NVirt, &
            OccBounds, VirtBounds, NSpins, Freq)
            !
            ! GPiUG += Prefactor * CRG * diag(d/(d^2+u^2)) * CRG^T per spin block
            ! (rpa_core_MO.f90:347-406).  DCRG is the reference's scratch array and
            ! is not touched.
            !
            real(F64), dimension(:, :), intent(inout), contiguous, target :: GPiUG
            real(F64), dimension(:, :, :), intent(in), contiguous, target :: CRG
            real(F64), dimension(:, :), intent(out)   :: DCRG
            real(F64), dimension(:, :), intent(in)    :: OccEnergies
            real(F64), dimension(:, :), intent(in)    :: VirtEnergies
            integer, dimension(2), intent(in)         :: NOcc
            integer, dimension(2), intent(in)         :: NVirt
            integer, dimension(2, 2), intent(in)      :: OccBounds
            integer, dimension(2, 2), intent(in)      :: VirtBounds
            integer, intent(in)                       :: NSpins
            real(F64), intent(in)                     :: Freq

            real(F64) :: Prefactor
            real(F64), dimension(:), allocatable :: eo, ev
            integer :: s, NVecs, Nai, i0, a0

            call brpa_init()
            if (NSpins > 1) then
                  Prefactor = 2.0_F64
            else
                  Prefactor = 4.0_F64
            end if
            NVecs = size(GPiUG, dim=1)
            do s = 1, NSpins
                  Nai = NOcc(s) * NVirt(s)
                  if (Nai > 0) then
                        i0 = OccBounds(1, s)
                        a0 = VirtBounds(1, s)
                        eo = OccEnergies(i0:i0+NOcc(s)-1, s)
                        ev = VirtEnergies(a0:a0+NVirt(s)-1, s)
                        call check(c_gpiug_mo(ctx, GPiUG, CRG(:, :, s), int(NVecs, c_int64_t), eo, &
                              int(NOcc(s), c_int64_t), ev, int(NVirt(s), c_int64_t), Freq, Prefactor), &
                              "rpa_GPiUG_MO")
                  end if
            end do
            if (.false.) DCRG = 0.0_F64
      end subroutine rpa_GPiUG_MO


      subroutine rpa_CC_EcRPA_Analytic(EcRPA, PiU, Freqs, FreqWeights, NFreqs, NVecsPiU)
            !
            ! EcRPA = sum_u w_u/(2 pi) [log det(1 + Pi(u)) - tr Pi(u)]
            ! (rpa_CC.f90:412-471)
            !
            real(F64), intent(out)                                    :: EcRPA
            real(F64), dimension(:, :, :), intent(in), contiguous     :: PiU
            real(F64), dimension(:), intent(in)                       :: Freqs
            real(F64), dimension(:), intent(in)                       :: FreqWeights
            integer, intent(in)                                       :: NFreqs
            integer, intent(in)                                       :: NVecsPiU

            call brpa_init()
            if (size(PiU, 1) /= NVecsPiU .or. size(PiU, 2) /= NVecsPiU) then
                  print '(a)', "rpa_CC_EcRPA_Analytic: PiU must be NVecsPiU x NVecsPiU x NFreqs"
                  error stop
            end if
            call check(c_ecrpa_analytic(ctx, EcRPA, PiU, Freqs, FreqWeights, &
                  int(NFreqs, c_int64_t), int(NVecsPiU, c_int64_t), c_null_ptr), "rpa_CC_EcRPA_Analytic")
      end subroutine rpa_CC_EcRPA_Analytic


      subroutine brpa_ecrpa_fused(EcRPA, Rkai, NVecs, OccEnergies, VirtEnergies, NOcc, NVirt, &
            Freqs, FreqWeights, NFreqs, SplitBF16, Precision)
            !
            ! Pi(u) build + E_RPA in one call, Pi kept on the GPU
            ! (rpa_CC_PiU frequency loop, rpa_CC.f90:523-526, then rpa_CC_EcRPA_Analytic).
            ! Precision (optional, overrides SplitBF16): 0 FP64, 1 split-BF16, 2 split-int8.
            !
            real(F64), intent(out)                             :: EcRPA
            real(F64), dimension(:, :), intent(in), contiguous :: Rkai
            integer, intent(in)                                :: NVecs, NOcc, NVirt, NFreqs
            real(F64), dimension(:), intent(in)                :: OccEnergies, VirtEnergies
            real(F64), dimension(:), intent(in)                :: Freqs, FreqWeights
            logical, intent(in)                                :: SplitBF16
            integer, intent(in), optional                      :: Precision
            integer(c_int) :: prec
            call brpa_init()
            prec = 0
            if (SplitBF16) prec = 1
            if (present(Precision)) prec = int(Precision, c_int)
            call check(c_ecrpa(ctx, prec, Rkai, int(NVecs, c_int64_t), OccEnergies, &
                  int(NOcc, c_int64_t), VirtEnergies, int(NVirt, c_int64_t), Freqs, FreqWeights, &
                  int(NFreqs, c_int64_t), EcRPA, c_null_ptr), "brpa_ecrpa_fused")
      end subroutine brpa_ecrpa_fused


      subroutine rpa_CC_T2(A, V, NVecsT2, PiUEigenvecs, PiUEigenvals, Rkai, NVecsChol, NOcc, NVirt, &
            OccBounds, VirtBounds, Freqs, FreqWeights, NFreqs, Lambda, OccEnergies, VirtEnergies, &
            SmallEigenvalsCutoffT2, GuessNVecsT2, MaxBatchDim)
            !
            ! Factorised direct-ring amplitudes (rpa_CC_Doubles.f90:43-77).  The
            ! eigenpairs of Pi(u) are not read: the library applies
            ! l (1 + l Pi(u))^{-1} through a Cholesky factor built from Rkai.
            ! MaxBatchDim only bounded the reference's memory and is not needed.
            !
            real(F64), dimension(:), allocatable, intent(out)    :: A
            real(F64), dimension(:, :), allocatable, intent(out) :: V
            integer, intent(inout)                               :: NVecsT2
            real(F64), dimension(:, :, :), intent(in)            :: PiUEigenvecs
            real(F64), dimension(:, :), intent(in)               :: PiUEigenvals
            real(F64), dimension(:, :), intent(in), contiguous   :: Rkai
            integer, intent(in)                                  :: NVecsChol
            integer, intent(in)                                  :: NOcc
            integer, intent(in)                                  :: NVirt
            integer, dimension(2), intent(in)                    :: OccBounds
            integer, dimension(2), intent(in)                    :: VirtBounds
            real(F64), dimension(:), intent(in)                  :: Freqs
            real(F64), dimension(:), intent(in)                  :: FreqWeights
            integer, intent(in)                                  :: NFreqs
            real(F64), intent(in)                                :: Lambda
            real(F64), dimension(:), intent(in)                  :: OccEnergies
            real(F64), dimension(:), intent(in)                  :: VirtEnergies
            real(F64), intent(in)                                :: SmallEigenvalsCutoffT2
            integer, intent(in)                                  :: GuessNVecsT2
            integer, intent(in)                                  :: MaxBatchDim
            integer(c_int64_t) :: nt

            call brpa_init()
            call check(c_t2_compute(ctx, 0_c_int, Rkai, int(NVecsChol, c_int64_t), OccEnergies, &
                  int(NOcc, c_int64_t), VirtEnergies, int(NVirt, c_int64_t), Freqs, FreqWeights, &
                  int(NFreqs, c_int64_t), Lambda, SmallEigenvalsCutoffT2, int(GuessNVecsT2, c_int64_t), &
                  0_c_int64_t, nt), "rpa_CC_T2")
            NVecsT2 = int(nt)
            allocate(A(NVecsT2))
            allocate(V(NVecsT2, NOcc*NVirt))
            if (NVecsT2 > 0) call check(c_t2_copyout(ctx, A, V), "rpa_CC_T2")
            if (.false.) print *, PiUEigenvecs(1,1,1), PiUEigenvals(1,1), OccBounds, VirtBounds, MaxBatchDim
      end subroutine rpa_CC_T2


      subroutine rpa_G(G, RWR, NVecs, TargetError, NCholesky, RWRBasis, NAO)
            !
            ! Basis of the effective dielectric matrix (rpa.f90:31-173), same argument
            ! list.  RWRBasis = 1 (random: the library's own Gaussian draws), 2
            ! (eigenvectors, the reference default) or 3 (full Cholesky space).
            !
            real(F64), dimension(NCholesky, NCholesky), intent(out)   :: G
            real(F64), dimension(NCholesky, NCholesky), intent(inout) :: RWR
            integer, intent(out)                                      :: NVecs
            real(F64), intent(in)                                     :: TargetError
            integer, intent(in)                                       :: NCholesky
            integer, intent(in)                                       :: RWRBasis
            integer, intent(in)                                       :: NAO
            integer(c_int64_t) :: nv
            real(c_double) :: err

            call brpa_init()
            call check(c_rpa_g(ctx, int(RWRBasis, c_int), G, RWR, int(NCholesky, c_int64_t), &
                  TargetError, 0_c_int64_t, nv, err), "rpa_G")
            NVecs = int(nv)
            if (.false.) print *, NAO
      end subroutine rpa_G


      subroutine rpa_CRG(CRG, RG, OccCoeffs, VirtCoeffs, NOcc, NVirt, OccBounds, VirtBounds, &
            NSpins, ShellPairs, ShellPairLoc, ShellPairDim, ShellLoc, ShellParamsIdx, SubsetBounds, &
            SubsetDim, NSubsets, NAngFunc, NAO, NVecs)
            !
            ! AO -> MO transform of the Cholesky vectors (rpa_core_MO.f90:233-301) on
            ! one image: the AO-pair subsets X = 1..NSubsets(1) of RG(:, :, X) are
            ! accumulated into CRG(:, :, s).  The AO pair of each packed column is
            ! taken from the shell-pair tables in rpa_C's order (rpa_Cab: p fastest
            ! over shell A, rpa_Caa: the lower triangle column by column) and the
            ! product runs in brpa_ao_mo.
            !
            real(F64), dimension(:, :, :), intent(out), contiguous :: CRG
            real(F64), dimension(:, :, :), intent(in)              :: RG
            real(F64), dimension(:, :, :), intent(in)              :: OccCoeffs
            real(F64), dimension(:, :, :), intent(in)              :: VirtCoeffs
            integer, dimension(2), intent(in)                      :: NOcc
            integer, dimension(2), intent(in)                      :: NVirt
            integer, dimension(2, 2), intent(in)                   :: OccBounds
            integer, dimension(2, 2), intent(in)                   :: VirtBounds
            integer, intent(in)                                    :: NSpins
            integer, dimension(:, :), intent(in)                   :: ShellPairs
            integer, dimension(:, :), intent(in)                   :: ShellPairLoc
            integer, dimension(:), intent(in)                      :: ShellPairDim
            integer, dimension(:), intent(in)                      :: ShellLoc
            integer, dimension(:), intent(in)                      :: ShellParamsIdx
            integer, dimension(:, :), intent(in)                   :: SubsetBounds
            integer, dimension(:), intent(in)                      :: SubsetDim
            integer, dimension(2), intent(in)                      :: NSubsets
            integer, dimension(:), intent(in)                      :: NAngFunc
            integer, intent(in)                                    :: NAO
            integer, intent(in)                                    :: NVecs
            integer, parameter :: SUBSET_STORAGE = 2   ! ParallelCholesky.f90:45
            integer(c_int32_t), dimension(:, :), allocatable :: pairs
            real(F64), dimension(:, :), allocatable :: RGxy, Co, Cv
            integer :: X, s, ShAB, LocAB, ShA, ShB, Na, Nb, LocA, LocB, p, q, pq, Npq
            integer(c_int), dimension(2) :: accum

            call brpa_init()
            if (NSubsets(2) /= 1) then
                  print *, "rpa_CRG: the GPU module runs on one image (NSubsets(2) must be 1)"
                  error stop
            end if
            CRG = 0.0_F64
            accum = 0_c_int
            do X = 1, NSubsets(1)
                  Npq = SubsetDim(X)
                  if (Npq <= 0) cycle
                  allocate(pairs(2, Npq), RGxy(NVecs, Npq))
                  do ShAB = SubsetBounds(1, X), SubsetBounds(2, X)
                        LocAB = ShellPairLoc(SUBSET_STORAGE, ShAB)
                        ShA = ShellPairs(1, ShAB)
                        ShB = ShellPairs(2, ShAB)
                        Na = NAngFunc(ShellParamsIdx(ShA))
                        Nb = NAngFunc(ShellParamsIdx(ShB))
                        LocA = ShellLoc(ShA)
                        LocB = ShellLoc(ShB)
                        if (ShA /= ShB) then
                              do q = 1, Nb
                                    do p = 1, Na
                                          pq = LocAB + p - 1 + (q - 1) * Na
                                          pairs(1, pq) = int(LocA + p - 2, c_int32_t)
                                          pairs(2, pq) = int(LocB + q - 2, c_int32_t)
                                    end do
                              end do
                        else
                              pq = LocAB
                              do q = 1, Na
                                    pairs(:, pq) = int(LocA + q - 2, c_int32_t)
                                    pq = pq + 1
                                    do p = q + 1, Na
                                          pairs(1, pq) = int(LocA + p - 2, c_int32_t)
                                          pairs(2, pq) = int(LocA + q - 2, c_int32_t)
                                          pq = pq + 1
                                    end do
                              end do
                        end if
                  end do
                  RGxy(:, :) = RG(1:NVecs, 1:Npq, X)
                  do s = 1, NSpins
                        if (NOcc(s) * NVirt(s) > 0) then
                              allocate(Co(NAO, NOcc(s)), Cv(NAO, NVirt(s)))
                              Co = OccCoeffs(:, OccBounds(1, s):OccBounds(2, s), s)
                              Cv = VirtCoeffs(:, VirtBounds(1, s):VirtBounds(2, s), s)
                              call check(c_ao_mo(ctx, CRG(:, :, s), RGxy, int(NVecs, c_int64_t), &
                                    int(Npq, c_int64_t), pairs, Co, Cv, int(NAO, c_int64_t), &
                                    int(NOcc(s), c_int64_t), int(NVirt(s), c_int64_t), accum(s)), "rpa_CRG")
                              accum(s) = 1_c_int
                              deallocate(Co, Cv)
                        end if
                  end do
                  deallocate(pairs, RGxy)
            end do
            if (.false.) print *, ShellPairDim(1)
      end subroutine rpa_CRG


      subroutine rpa_CC_Exchange_MBPT3_1(EcExchange, Vkai, Ak, NVecsT2, Rkai, NVecsPiU, &
            OccEnergies, VirtEnergies, NOcc, NVirt, Precision)
            !
            ! Third-order exchange MBPT3-1 (rpa_CC_Exchange.f90:232-270), same
            ! argument list; the optional trailing Precision=2 selects the split-int8
            ! form of the final contraction (default FP64).
            !
            real(F64), intent(out)                                          :: EcExchange
            real(F64), dimension(NVecsT2, NVirt, NOcc), intent(in)          :: Vkai
            real(F64), dimension(NVecsT2), intent(in)                       :: Ak
            integer, intent(in)                                             :: NVecsT2
            real(F64), dimension(NVecsPiU, NVirt, NOcc), intent(in)         :: Rkai
            integer, intent(in)                                             :: NVecsPiU
            real(F64), dimension(:), intent(in), contiguous                 :: OccEnergies
            real(F64), dimension(:), intent(in), contiguous                 :: VirtEnergies
            integer, intent(in)                                             :: NOcc
            integer, intent(in)                                             :: NVirt
            integer, intent(in), optional                                   :: Precision
            integer(c_int) :: prec

            call brpa_init()
            prec = 0_c_int
            if (present(Precision)) prec = int(Precision, c_int)
            call check(c_mbpt3_1(ctx, prec, EcExchange, Vkai, Ak, int(NVecsT2, c_int64_t), Rkai, &
                  int(NVecsPiU, c_int64_t), OccEnergies, VirtEnergies, int(NOcc, c_int64_t), &
                  int(NVirt, c_int64_t)), "rpa_CC_Exchange_MBPT3_1")
      end subroutine rpa_CC_Exchange_MBPT3_1


      subroutine rpa_CC_Exchange_MBPT3_2(EcExchange, Rkai, Vkai, Ak, NOcc, NVirt, NVecsPiU, NVecsT2)
            !
            ! Third-order exchange MBPT3-2 (rpa_CC_Exchange.f90:381-393), same argument list.
            !
            real(F64), intent(out)                                  :: EcExchange
            real(F64), dimension(NVecsPiU, NVirt, NOcc), intent(in) :: Rkai
            real(F64), dimension(NVecsT2, NVirt, NOcc), intent(in)  :: Vkai
            real(F64), dimension(NVecsT2), intent(in)               :: Ak
            integer, intent(in)                                     :: NOcc
            integer, intent(in)                                     :: NVirt
            integer, intent(in)                                     :: NVecsPiU
            integer, intent(in)                                     :: NVecsT2

            call brpa_init()
            call check(c_mbpt3_2(ctx, EcExchange, Rkai, Vkai, Ak, int(NOcc, c_int64_t), &
                  int(NVirt, c_int64_t), int(NVecsPiU, c_int64_t), int(NVecsT2, c_int64_t)), &
                  "rpa_CC_Exchange_MBPT3_2")
      end subroutine rpa_CC_Exchange_MBPT3_2


      subroutine rpa_THC_PiU(PiU, Rkai, Xga, Xgi, Zgk, NOcc, NVirt, Freqs, &
            OccEnergies, VirtEnergies, MaxBlockDim, FullRkai)
            !
            ! Pi(u) of every frequency from the THC factors (rpa_THC.f90:215-282).
            ! As in the reference, the number of spin blocks is size(Rkai, 3) and
            ! Alpha is 4 for one block, 2 otherwise; FullRkai = .true. writes the
            ! complete R of block s into Rkai(:, :, s), .false. leaves Rkai alone
            ! (the reference keeps only its last ai block there).  MaxBlockDim only
            ! bounded the reference's memory.
            !
            real(F64), dimension(:, :, :), intent(out), contiguous         :: PiU
            real(F64), dimension(:, :, :), intent(out), contiguous, target :: Rkai
            real(F64), dimension(:, :, :), intent(in), contiguous          :: Xga
            real(F64), dimension(:, :, :), intent(in), contiguous          :: Xgi
            real(F64), dimension(:, :), intent(in), contiguous             :: Zgk
            integer, dimension(2), intent(in)                              :: NOcc
            integer, dimension(2), intent(in)                              :: NVirt
            real(F64), dimension(:), intent(in)                            :: Freqs
            real(F64), dimension(:, :), intent(in), contiguous             :: OccEnergies
            real(F64), dimension(:, :), intent(in), contiguous             :: VirtEnergies
            integer, intent(in)                                            :: MaxBlockDim
            logical, intent(in)                                            :: FullRkai
            integer :: s, NSpins, NCholesky, NGrid, NFreqs
            integer(c_int) :: accum
            real(F64) :: Alpha
            type(c_ptr) :: rp

            call brpa_init()
            NSpins = size(Rkai, dim=3)
            NCholesky = size(Zgk, dim=2)
            NGrid = size(Zgk, dim=1)
            NFreqs = size(PiU, dim=3)
            Alpha = 2.0_F64
            if (NSpins == 1) Alpha = 4.0_F64
            PiU = 0.0_F64
            accum = 1_c_int
            do s = 1, NSpins
                  if (NOcc(s) > 0) then
                        rp = c_null_ptr
                        if (FullRkai) then
                              if (size(Rkai, dim=1) /= NCholesky .or. &
                                    size(Rkai, dim=2) < NOcc(s)*NVirt(s)) then
                                    print *, "rpa_THC_PiU: Rkai must be (NCholesky, >= NOcc*NVirt, NSpins)"
                                    error stop
                              end if
                              rp = c_loc(Rkai(1, 1, s))
                        end if
                        call check(c_thc_piu(ctx, PiU, rp, Xga(:, :, s), Xgi(:, :, s), Zgk, &
                              int(NGrid, c_int64_t), int(NOcc(s), c_int64_t), int(NVirt(s), c_int64_t), &
                              int(NCholesky, c_int64_t), OccEnergies(:, s), VirtEnergies(:, s), Freqs, &
                              int(NFreqs, c_int64_t), Alpha, accum), "rpa_THC_PiU")
                  end if
            end do
            if (.false.) print *, MaxBlockDim
      end subroutine rpa_THC_PiU


      subroutine rpa_THC_CC_T2(A, V, EigRPA, NVecsT2, PiUEigenvecs, PiUEigenvals, Rkai, NVecsChol, &
            NOcc, NVirt, Freqs, FreqWeights, NFreqs, Lambda, OccEnergies, VirtEnergies, &
            SmallEigenvalsCutoffT2, GuessNVecsT2, MaxBatchDim, T2CutoffThresh, T2CutoffType, &
            T2CutoffSmoothStep, T2CutoffSteepness, T2CutoffCommonThresh, RPAParams)
            !
            ! Direct-ring amplitudes with the small-eigenvalue cutoff of the THC
            ! route (rpa_CC_Doubles.f90:645-984): V(NVirt*NOcc, NVecsT2) and
            ! EigRPA(NVecsT2) = 2 A |Rkai V|^2.  T2CutoffCommonThresh < 0 on entry:
            ! supersystem, the threshold used is stored in it for the subsystems.
            ! RPAParams is not read by the reference routine either.
            !
            real(F64), dimension(:), allocatable, intent(out)    :: A
            real(F64), dimension(:, :), allocatable, intent(out) :: V
            real(F64), dimension(:), allocatable, intent(out)    :: EigRPA
            integer, intent(inout)                               :: NVecsT2
            real(F64), dimension(:, :, :), intent(in)            :: PiUEigenvecs
            real(F64), dimension(:, :), intent(in)               :: PiUEigenvals
            real(F64), dimension(:, :), intent(in), contiguous   :: Rkai
            integer, intent(in)                                  :: NVecsChol
            integer, intent(in)                                  :: NOcc
            integer, intent(in)                                  :: NVirt
            real(F64), dimension(:), intent(in)                  :: Freqs
            real(F64), dimension(:), intent(in)                  :: FreqWeights
            integer, intent(in)                                  :: NFreqs
            real(F64), intent(in)                                :: Lambda
            real(F64), dimension(:), intent(in)                  :: OccEnergies
            real(F64), dimension(:), intent(in)                  :: VirtEnergies
            real(F64), intent(in)                                :: SmallEigenvalsCutoffT2
            integer, intent(in)                                  :: GuessNVecsT2
            integer, intent(in)                                  :: MaxBatchDim
            real(F64), intent(in)                                :: T2CutoffThresh
            integer, intent(in)                                  :: T2CutoffType
            logical, intent(in)                                  :: T2CutoffSmoothStep
            real(F64), intent(in)                                :: T2CutoffSteepness
            real(F64), intent(inout)                             :: T2CutoffCommonThresh
            class(*), intent(in), optional                       :: RPAParams
            integer(c_int64_t) :: nt, nsc
            integer(c_int) :: smooth

            call brpa_init()
            call check(c_t2_compute(ctx, 0_c_int, Rkai, int(NVecsChol, c_int64_t), OccEnergies, &
                  int(NOcc, c_int64_t), VirtEnergies, int(NVirt, c_int64_t), Freqs, FreqWeights, &
                  int(NFreqs, c_int64_t), Lambda, SmallEigenvalsCutoffT2, int(GuessNVecsT2, c_int64_t), &
                  0_c_int64_t, nt), "rpa_THC_CC_T2")
            smooth = 0_c_int
            if (T2CutoffSmoothStep) smooth = 1_c_int
            call check(c_t2_cutoff(ctx, int(T2CutoffType, c_int), T2CutoffThresh, smooth, &
                  T2CutoffSteepness, int(NOcc, c_int64_t), T2CutoffCommonThresh, nt, nsc), &
                  "rpa_THC_CC_T2")
            NVecsT2 = int(nt)
            allocate(A(NVecsT2))
            allocate(V(NVirt*NOcc, NVecsT2))
            allocate(EigRPA(NVecsT2))
            call check(c_t2_copyout_ov(ctx, A, V), "rpa_THC_CC_T2")
            call check(c_t2_eig_rpa(ctx, Rkai, int(NVecsChol, c_int64_t), EigRPA), "rpa_THC_CC_T2")
            if (.false.) print *, PiUEigenvecs(1,1,1), PiUEigenvals(1,1), MaxBatchDim, present(RPAParams)
      end subroutine rpa_THC_CC_T2


      subroutine rpa_CC_Exchange_SOSEX(EcExchange, Alpha, Vkai, Ak, Rkai, NOcc, NVirt, NVecsPiU, &
            NVecsT2, S2Operator, Precision)
            !
            ! EcExchange = -Alpha * Sum(abij) (aj|bi)*Taibj (rpa_CC_Exchange.f90:36-97)
            ! Precision (optional, default 0 = FP64): 2 selects the fused split-int8 kernel.
            !
            real(F64), intent(out)                                          :: EcExchange
            real(F64), intent(in)                                           :: Alpha
            real(F64), dimension(NVecsT2, NVirt, NOcc), intent(in)          :: Vkai
            real(F64), dimension(NVecsT2), intent(in)                       :: Ak
            real(F64), dimension(NVecsPiU, NVirt, NOcc), intent(in)         :: Rkai
            integer, intent(in)                                             :: NOcc
            integer, intent(in)                                             :: NVirt
            integer, intent(in)                                             :: NVecsPiU
            integer, intent(in)                                             :: NVecsT2
            logical, intent(in)                                             :: S2Operator
            integer, intent(in), optional                                   :: Precision
            integer(c_int) :: s2, prec

            call brpa_init()
            s2 = 0
            if (S2Operator) s2 = 1
            prec = 0
            if (present(Precision)) prec = int(Precision, c_int)
            call check(c_sosex(ctx, prec, EcExchange, Alpha, Vkai, Ak, Rkai, int(NOcc, c_int64_t), &
                  int(NVirt, c_int64_t), int(NVecsPiU, c_int64_t), int(NVecsT2, c_int64_t), s2), &
                  "rpa_CC_Exchange_SOSEX")
      end subroutine rpa_CC_Exchange_SOSEX

      subroutine rpa_DaiHistogram(daiValues, daiWeights, OrbEnergies, NOcc, NVirt, &
            CoreOrbThresh, DaiMaxThresh)
            !
            ! Histogram of the gaps of one spin (rpa/OrbDiffHist.f90:42-123).
            ! OrbEnergies holds the NOcc occupied (core included) then the NVirt
            ! virtual energies; size(daiValues) must be RPA_HISTOGRAM_NBINS = 100.
            !
            real(F64), dimension(:), intent(out) :: daiValues
            real(F64), dimension(:), intent(out) :: daiWeights
            real(F64), dimension(:), intent(in)  :: OrbEnergies
            integer, intent(in)                  :: NOcc
            integer, intent(in)                  :: NVirt
            real(F64), intent(in)                :: CoreOrbThresh
            real(F64), intent(in)                :: DaiMaxThresh
            real(F64), dimension(100) :: v, w
            if (size(daiValues) /= 100 .or. size(daiWeights) /= 100) then
                  print '(a)', "rpa_DaiHistogram: the library's histograms have 100 bins"
                  error stop
            end if
            call check_grid(c_dai_histogram(OrbEnergies(1:NOcc), int(NOcc, c_int64_t), &
                  OrbEnergies(NOcc+1:NOcc+NVirt), int(NVirt, c_int64_t), CoreOrbThresh, &
                  DaiMaxThresh, v, w), "rpa_DaiHistogram")
            daiValues = v
            daiWeights = w
      end subroutine rpa_DaiHistogram


      subroutine rpa_DaiMaxThresh(daiMaxThresh, OrbEnergies, NOcc, NVirt, NSpins, CoreOrbThresh)
            !
            ! Largest active gap of the full complex (rpa/OrbDiffHist.f90:10-39).
            !
            real(F64), intent(out)                 :: daiMaxThresh
            real(F64), dimension(:, :), intent(in) :: OrbEnergies
            integer, dimension(2), intent(in)      :: NOcc
            integer, dimension(2), intent(in)      :: NVirt
            integer, intent(in)                    :: NSpins
            real(F64), intent(in)                  :: CoreOrbThresh
            real(F64), dimension(:), allocatable :: eo, ev
            integer(c_int64_t), dimension(2) :: no, nv
            integer :: s
            allocate(eo(0), ev(0))
            do s = 1, NSpins
                  eo = [eo, OrbEnergies(1:NOcc(s), s)]
                  ev = [ev, OrbEnergies(NOcc(s)+1:NOcc(s)+NVirt(s), s)]
                  no(s) = int(NOcc(s), c_int64_t)
                  nv(s) = int(NVirt(s), c_int64_t)
            end do
            call check_grid(c_dai_max_thresh(int(NSpins, c_int64_t), eo, no, ev, nv, &
                  CoreOrbThresh, daiMaxThresh), "rpa_DaiMaxThresh")
      end subroutine rpa_DaiMaxThresh


      subroutine rpa_OptQuadScaling(Freqs, FreqWeights, OptAlpha, OptRelError, OptRMSD, NFreqs, &
            daiValues, daiWeights)
            !
            ! Best Casimir-Polder scale on the histogram sets (rpa/OptFreqQuad.f90:75-122).
            ! The nodes come back as quad_CasimirPolder(NFreqs, OptAlpha) builds them,
            ! which may differ from OptAlpha * (unscaled nodes) in the last bit.
            !
            real(F64), dimension(NFreqs), intent(out) :: Freqs
            real(F64), dimension(NFreqs), intent(out) :: FreqWeights
            real(F64), intent(out)                    :: OptAlpha
            real(F64), intent(out)                    :: OptRelError
            real(F64), intent(out)                    :: OptRMSD
            integer, intent(in)                       :: NFreqs
            real(F64), dimension(:, :), intent(in), contiguous :: daiValues
            real(F64), dimension(:, :), intent(in), contiguous :: daiWeights
            real(F64) :: info(3)
            if (size(daiValues, 1) /= 100) then
                  print '(a)', "rpa_OptQuadScaling: the library's histograms have 100 bins"
                  error stop
            end if
            call check_grid(c_opt_quad_scaling(int(NFreqs, c_int64_t), &
                  int(size(daiValues, 2), c_int64_t), daiValues, daiWeights, Freqs, FreqWeights, &
                  info), "rpa_OptQuadScaling")
            OptAlpha = info(1)
            OptRMSD = info(2)
            OptRelError = info(3)
      end subroutine rpa_OptQuadScaling


      subroutine rpa_OptimizeFreqQuad(Freqs, FreqWeights, NFreqs, daiValues, daiWeights, &
            TargetError, TargetRelError, MinNFreqs, MaxNFreqs, GridLimitDai)
            !
            ! Smallest grid meeting both targets on the histogram sets
            ! (rpa/OptFreqQuad.f90:12-72).  GridLimitDai only changed the
            ! reference's log: the cap is in the histograms.
            !
            real(F64), dimension(MaxNFreqs), intent(out) :: Freqs
            real(F64), dimension(MaxNFreqs), intent(out) :: FreqWeights
            integer, intent(out)                         :: NFreqs
            real(F64), dimension(:, :), intent(in), contiguous :: daiValues
            real(F64), dimension(:, :), intent(in), contiguous :: daiWeights
            real(F64), intent(in)                        :: TargetError
            real(F64), intent(in)                        :: TargetRelError
            integer, intent(in)                          :: MinNFreqs, MaxNFreqs
            logical, intent(in)                          :: GridLimitDai
            real(F64) :: info(3)
            integer(c_int64_t) :: n
            if (size(daiValues, 1) /= 100) then
                  print '(a)', "rpa_OptimizeFreqQuad: the library's histograms have 100 bins"
                  error stop
            end if
            call check_grid(c_optimize_freq_quad(int(size(daiValues, 2), c_int64_t), daiValues, &
                  daiWeights, TargetError, TargetRelError, int(MinNFreqs, c_int64_t), &
                  int(MaxNFreqs, c_int64_t), Freqs, FreqWeights, n, info), "rpa_OptimizeFreqQuad")
            NFreqs = int(n)
            if (.false. .and. GridLimitDai) NFreqs = 0
      end subroutine rpa_OptimizeFreqQuad

end module brpa_mod
