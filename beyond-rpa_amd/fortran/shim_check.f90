!
! shim_check -- calls the Fortran drop-in routines of brpa_mod the way the
! reference's rpa_CC_Energy_2 does (rpa/rpa_CC.f90:519-527, 785, 884-892,
! 965-971) and writes the results.  Used by the GPU test of the Fortran
! boundary.
!
!   shim_check input.bin output.txt
!
! input.bin (stream, 64-bit integers and doubles):
!   NOcc, NVirt, NVecs, NFreqs, Seed
!   Rkai(NVecs, NVirt*NOcc), OccEnergies(NOcc), VirtEnergies(NVirt),
!   Freqs(NFreqs), FreqWeights(NFreqs)
!
program shim_check
      use iso_c_binding
      use brpa_mod
      implicit none
      integer, parameter :: F64 = c_double
      integer :: NOcc, NVirt, NVecs, NFreqs, Seed, u, k, NVecsT2, io
      real(F64), allocatable :: Rkai(:, :), Rk3(:, :, :), Eo(:), Ev(:), Freqs(:), Weights(:)
      real(F64), allocatable :: PiU(:, :, :), DCRG(:, :), OccE(:, :), VirtE(:, :)
      real(F64), allocatable :: A(:), V(:, :), Q(:, :, :), W(:, :)
      integer :: NOccs(2), NVirts(2), OccBounds(2, 2), VirtBounds(2, 2)
      real(F64) :: EcRPA, EcFused, EcSplit, Ex
      character(len=512) :: fin, fout

      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
      open(newunit=io, file=trim(fin), access='stream', form='unformatted', status='old')
      read(io) NOcc, NVirt, NVecs, NFreqs, Seed
      allocate(Rkai(NVecs, NVirt*NOcc), Eo(NOcc), Ev(NVirt), Freqs(NFreqs), Weights(NFreqs))
      read(io) Rkai, Eo, Ev, Freqs, Weights
      close(io)

      ! Pi(u) through rpa_GPiUG_MO, one closed-shell spin block per frequency
      allocate(PiU(NVecs, NVecs, NFreqs), DCRG(NVecs, NVirt*NOcc))
      allocate(Rk3(NVecs, NVirt*NOcc, 1), OccE(NOcc, 1), VirtE(NVirt, 1))
      Rk3(:, :, 1) = Rkai
      OccE(:, 1) = Eo
      VirtE(:, 1) = Ev
      NOccs = [NOcc, 0]
      NVirts = [NVirt, 0]
      OccBounds = reshape([1, NOcc, 1, 0], [2, 2])
      VirtBounds = reshape([1, NVirt, 1, 0], [2, 2])
      PiU = 0.0_F64
      do u = 1, NFreqs
            call rpa_GPiUG_MO(PiU(:, :, u), Rk3, DCRG, OccE, VirtE, NOccs, NVirts, &
                  OccBounds, VirtBounds, 1, Freqs(u))
      end do
      call rpa_CC_EcRPA_Analytic(EcRPA, PiU, Freqs, Weights, NFreqs, NVecs)
      call brpa_ecrpa_fused(EcFused, Rkai, NVecs, Eo, Ev, NOcc, NVirt, Freqs, Weights, NFreqs, .false.)
      call brpa_ecrpa_fused(EcSplit, Rkai, NVecs, Eo, Ev, NOcc, NVirt, Freqs, Weights, NFreqs, .true.)

      allocate(Q(NVecs, NVecs, NFreqs), W(NVecs, NFreqs))
      Q = 0.0_F64
      W = 0.0_F64
      NVecsT2 = 0
      call rpa_CC_T2(A, V, NVecsT2, Q, W, Rkai, NVecs, NOcc, NVirt, [1, NOcc], [1, NVirt], &
            Freqs, Weights, NFreqs, 1.0_F64, Eo, Ev, 1.0E-10_F64, 0, 100)
      call rpa_CC_Exchange_SOSEX(Ex, 1.0_F64, V, A, Rkai, NOcc, NVirt, NVecs, NVecsT2, .false.)

      open(newunit=io, file=trim(fout), status='replace')
      write(io, '(es25.17)') EcRPA
      write(io, '(es25.17)') EcFused
      write(io, '(es25.17)') EcSplit
      write(io, '(i0)') NVecsT2
      write(io, '(es25.17)') Ex
      do k = 1, NVecsT2
            write(io, '(es25.17)') A(k)
      end do
      close(io)
      call brpa_finalize()
      if (.false.) print *, Seed
end program shim_check
