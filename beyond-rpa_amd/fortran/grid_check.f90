!
! grid_check -- calls the frequency-grid routines of brpa_mod the way the
! reference's driver assembles a grid for a complex and one of its subsystems
! (rpa/rpa_driver.f90:118-172, GridLimitDai on; rpa/rpa.f90:445) and writes
! the results.  Host code only: no GPU context is created.
!
!   grid_check input.bin output.txt
!
! input.bin (stream, 64-bit integers and doubles):
!   NOccC, NVirtC, NOccA, NVirtA, NFreqs,
!   OrbEnergiesC(NOccC+NVirtC), OrbEnergiesA(NOccA+NVirtA)   (occupied first,
!   core included, each block ascending)
! output.txt: DaiMaxThresh, daiValues(100, 2), daiWeights(100, 2), OptAlpha,
!   OptRMSD, OptRelError, Freqs(NFreqs), FreqWeights(NFreqs) of
!   rpa_OptQuadScaling, then NFreqs2, Freqs(1:NFreqs2), FreqWeights(1:NFreqs2)
!   of rpa_OptimizeFreqQuad (targets 1e-6 and 1e-3, 8 to 64 points)
!
program grid_check
      use iso_c_binding
      use brpa_mod
      implicit none
      integer, parameter :: F64 = c_double
      integer :: NOccC, NVirtC, NOccA, NVirtA, NFreqs, NFreqs2, io, k
      real(F64), allocatable :: OrbC(:), OrbA(:), OrbC2(:, :)
      real(F64), allocatable :: Freqs(:), Weights(:), Freqs2(:), Weights2(:)
      real(F64) :: daiValues(100, 2), daiWeights(100, 2)
      real(F64) :: Thresh, Alpha, RMSD, RelError
      integer :: NOccs(2), NVirts(2)
      character(len=512) :: fin, fout

      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
      open(newunit=io, file=trim(fin), access='stream', form='unformatted', status='old')
      read(io) NOccC, NVirtC, NOccA, NVirtA, NFreqs
      allocate(OrbC(NOccC+NVirtC), OrbA(NOccA+NVirtA))
      read(io) OrbC, OrbA
      close(io)

      allocate(OrbC2(NOccC+NVirtC, 1))
      OrbC2(:, 1) = OrbC
      NOccs = [NOccC, 0]
      NVirts = [NVirtC, 0]
      call rpa_DaiMaxThresh(Thresh, OrbC2, NOccs, NVirts, 1, -3.0_F64)
      call rpa_DaiHistogram(daiValues(:, 1), daiWeights(:, 1), OrbC, NOccC, NVirtC, -3.0_F64, Thresh)
      call rpa_DaiHistogram(daiValues(:, 2), daiWeights(:, 2), OrbA, NOccA, NVirtA, -3.0_F64, Thresh)
      allocate(Freqs(NFreqs), Weights(NFreqs), Freqs2(64), Weights2(64))
      call rpa_OptQuadScaling(Freqs, Weights, Alpha, RelError, RMSD, NFreqs, daiValues, daiWeights)
      call rpa_OptimizeFreqQuad(Freqs2, Weights2, NFreqs2, daiValues, daiWeights, 1.0E-6_F64, &
            1.0E-3_F64, 8, 64, .true.)

      open(newunit=io, file=trim(fout), status='replace')
      write(io, '(es26.17e3)') Thresh
      write(io, '(es26.17e3)') daiValues
      write(io, '(es26.17e3)') daiWeights
      write(io, '(es26.17e3)') Alpha, RMSD, RelError
      write(io, '(es26.17e3)') (Freqs(k), k = 1, NFreqs)
      write(io, '(es26.17e3)') (Weights(k), k = 1, NFreqs)
      write(io, '(i0)') NFreqs2
      write(io, '(es26.17e3)') (Freqs2(k), k = 1, NFreqs2)
      write(io, '(es26.17e3)') (Weights2(k), k = 1, NFreqs2)
      close(io)
end program grid_check
