!
! shim_check -- calls the Fortran drop-in routines of brpa_mod the way the
! reference's rpa_CC_Energy_2 does (rpa/rpa_CC.f90:519-527, 785, 884-892,
! 965-971) and the THC route's amplitude call (rpa/rpa_THC.f90:115-121), and
! writes the results.  Used by the GPU test of the Fortran
! boundary.
!
!   shim_check input.bin output.txt [thc_input.bin thc_output.txt [aomo_input.bin aomo_output.txt]]
!
! aomo_input.bin (stream): NAO, NShells, NVecs, NOcc, NVirt, ShellSize(NShells),
!   RG(NVecs, Npq) (all shell pairs A >= B in order, rpa_C packing),
!   OccCoeffs(NAO, NOcc), VirtCoeffs(NAO, NVirt)
! aomo_output.txt: CRG(NVecs, NVirt*NOcc) of rpa_CRG with the shell pairs split
! into two subsets
!
! thc_input.bin (stream): NOcc, NVirt, NCholesky, NGrid, NFreqs,
!   Xga(NGrid, NVirt), Xgi(NGrid, NOcc), Zgk(NGrid, NCholesky),
!   OccEnergies(NOcc), VirtEnergies(NVirt), Freqs(NFreqs)
! thc_output.txt: PiU(NCholesky, NCholesky, NFreqs) then Rkai(NCholesky, NOcc*NVirt)
! of a closed-shell rpa_THC_PiU call with FullRkai = .true.
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
      real(F64), allocatable :: Ath(:), Vth(:, :), EigRPA(:)
      real(F64) :: CommonThresh
      integer :: NVecsTHC
      integer :: NOccs(2), NVirts(2), OccBounds(2, 2), VirtBounds(2, 2)
      real(F64) :: EcRPA, EcFused, EcSplit, Ex, E31, E32
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
      call rpa_CC_Exchange_MBPT3_1(E31, V, A, NVecsT2, Rkai, NVecs, Eo, Ev, NOcc, NVirt)
      call rpa_CC_Exchange_MBPT3_2(E32, Rkai, V, A, NOcc, NVirt, NVecs, NVecsT2)

      ! THC route: supersystem call with the EIG/MAXEIG rule and the smooth step
      CommonThresh = -1.0_F64
      NVecsTHC = 0
      call rpa_THC_CC_T2(Ath, Vth, EigRPA, NVecsTHC, Q, W, Rkai, NVecs, NOcc, NVirt, Freqs, &
            Weights, NFreqs, 1.0_F64, Eo, Ev, 1.0E-10_F64, 0, 100, 1.0E-3_F64, &
            RPA_T2_CUTOFF_EIG_DIV_MAXEIG, .true., 0.1_F64, CommonThresh)

      open(newunit=io, file=trim(fout), status='replace')
      write(io, '(es25.17)') EcRPA
      write(io, '(es25.17)') EcFused
      write(io, '(es25.17)') EcSplit
      write(io, '(i0)') NVecsT2
      write(io, '(es25.17)') Ex
      do k = 1, NVecsT2
            write(io, '(es25.17)') A(k)
      end do
      write(io, '(es25.17)') E31
      write(io, '(es25.17)') E32
      write(io, '(i0)') NVecsTHC
      write(io, '(es25.17)') CommonThresh
      write(io, '(es25.17)') sum(EigRPA)
      do k = 1, NVecsTHC
            write(io, '(es25.17)') Ath(k)
      end do
      ! layout-sensitive, sign-free: |V(:, 1)|^2 and T2(ai=1, bj=1) = sum_mu A_mu V(1, mu)^2
      write(io, '(es25.17)') sum(Vth(:, 1)**2)
      write(io, '(es25.17)') sum(Ath * Vth(1, :)**2)
      close(io)
      if (command_argument_count() >= 4) call thc_check()
      if (command_argument_count() >= 6) call aomo_check()
      call brpa_finalize()
      if (.false.) print *, Seed

contains

      subroutine thc_check()
            integer :: to, tv, tn, tg, tf, i
            integer :: NO2(2), NV2(2)
            real(F64), allocatable :: Xa(:, :, :), Xi(:, :, :), Zg(:, :), Eo2(:, :), Ev2(:, :), Fr(:)
            real(F64), allocatable :: P3(:, :, :), R3(:, :, :)
            character(len=512) :: tin, tout

            call get_command_argument(3, tin)
            call get_command_argument(4, tout)
            open(newunit=io, file=trim(tin), access='stream', form='unformatted', status='old')
            read(io) to, tv, tn, tg, tf
            allocate(Xa(tg, tv, 1), Xi(tg, to, 1), Zg(tg, tn), Eo2(to, 1), Ev2(tv, 1), Fr(tf))
            read(io) Xa, Xi, Zg, Eo2, Ev2, Fr
            close(io)
            allocate(P3(tn, tn, tf), R3(tn, to*tv, 1))
            NO2 = [to, 0]
            NV2 = [tv, 0]
            call rpa_THC_PiU(P3, R3, Xa, Xi, Zg, NO2, NV2, Fr, Eo2, Ev2, 100, .true.)
            open(newunit=io, file=trim(tout), status='replace')
            do i = 1, size(P3)
                  write(io, '(es25.17)') P3(mod(i-1, tn) + 1, mod((i-1)/tn, tn) + 1, (i-1)/(tn*tn) + 1)
            end do
            do i = 1, size(R3)
                  write(io, '(es25.17)') R3(mod(i-1, tn) + 1, (i-1)/tn + 1, 1)
            end do
            close(io)
      end subroutine thc_check

      subroutine aomo_check()
            integer :: nao_, nsh, nv, no_, nvr, a, b, k, h, npq, npq1, i
            integer, allocatable :: sz(:), sloc(:), sp(:, :), sploc(:, :), spdim(:), spidx(:)
            integer :: sb(2, 2), sdim(2)
            real(F64), allocatable :: RGfull(:, :), RG3(:, :, :), Oc(:, :, :), Vc(:, :, :), CR(:, :, :)
            character(len=512) :: ain, aout

            call get_command_argument(5, ain)
            call get_command_argument(6, aout)
            open(newunit=io, file=trim(ain), access='stream', form='unformatted', status='old')
            read(io) nao_, nsh, nv, no_, nvr
            allocate(sz(nsh), sloc(nsh), spidx(nsh))
            read(io) sz
            npq = 0
            k = 0
            do a = 1, nsh
                  do b = 1, a
                        k = k + 1
                        if (a /= b) then
                              npq = npq + sz(a) * sz(b)
                        else
                              npq = npq + sz(a) * (sz(a) + 1) / 2
                        end if
                  end do
            end do
            allocate(RGfull(nv, npq), Oc(nao_, no_, 1), Vc(nao_, nvr, 1))
            read(io) RGfull, Oc, Vc
            close(io)
            sloc(1) = 1
            do a = 2, nsh
                  sloc(a) = sloc(a - 1) + sz(a - 1)
            end do
            do a = 1, nsh
                  spidx(a) = a
            end do
            allocate(sp(2, k), sploc(2, k), spdim(k))
            h = k / 2
            k = 0
            npq1 = 0
            do a = 1, nsh
                  do b = 1, a
                        k = k + 1
                        sp(:, k) = [a, b]
                        if (a /= b) then
                              spdim(k) = sz(a) * sz(b)
                        else
                              spdim(k) = sz(a) * (sz(a) + 1) / 2
                        end if
                        if (k == 1 .or. k == h + 1) then
                              sploc(2, k) = 1
                        else
                              sploc(2, k) = sploc(2, k - 1) + spdim(k - 1)
                        end if
                        if (k <= h) npq1 = npq1 + spdim(k)
                  end do
            end do
            sploc(1, :) = 0
            sb(:, 1) = [1, h]
            sb(:, 2) = [h + 1, k]
            sdim = [npq1, npq - npq1]
            allocate(RG3(nv, max(sdim(1), sdim(2)), 2))
            RG3 = 0.0_F64
            RG3(:, 1:sdim(1), 1) = RGfull(:, 1:npq1)
            RG3(:, 1:sdim(2), 2) = RGfull(:, npq1 + 1:npq)
            allocate(CR(nv, no_ * nvr, 1))
            call rpa_CRG(CR, RG3, Oc, Vc, [no_, 0], [nvr, 0], reshape([1, no_, 1, 0], [2, 2]), &
                  reshape([1, nvr, 1, 0], [2, 2]), 1, sp, sploc, spdim, sloc, spidx, sb, sdim, [2, 1], &
                  sz, nao_, nv)
            open(newunit=io, file=trim(aout), status='replace')
            do i = 1, size(CR)
                  write(io, '(es25.17)') CR(mod(i - 1, nv) + 1, (i - 1) / nv + 1, 1)
            end do
            close(io)
      end subroutine aomo_check
end program shim_check
