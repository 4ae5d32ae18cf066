! Fortran caller of the C ABI through the stda_hip module: it stands where
! std2's `stda` stands at stda.f:452 and does what the replaced block does for
! the default sTDA run (screening, PT2, TDA in the energy window, velocity-
! corrected properties) or for sTD-DFT, then writes tda.dat in the layout of
! apbtrafo.f:115-128,265 (header, then (i4,F10.4,4f13.6) rows: state, energy in
! eV with 27.21139, fL, fV, RL, RV / Rv(corr)).
!
! Usage: tda_harness <input> <tda.dat>     solve and write tda.dat
!        tda_harness --index-size          print the index sizes of library and caller
!
! <input> is a stream file holding what `stda` has at stda.f:452:
!   int32   magic 1396982849 ('ADTS'), version 1
!   int32   ncent, nao, no, nv, nroots (<= 0: energy window), velcorr, rpa
!   float64 thr_ev, thrp, fthr_ev (<= 0: unlimited), ax, alphak, betaj, xmolw
!   float64 clow(nao, no+nv), epsi(no+nv)
!   int32   iaoat(nao)
!   float64 xyz(4, ncent)
!   float64 xl, yl, zl, xm, ym, zm, xv, yv, zv, each (no+nv)(no+nv+1)/2
program tda_harness
   use iso_c_binding
   use iso_fortran_env, only: int32, real64
   use stda_hip
   implicit none
   character(len=4096) :: inp, outp
   integer(int32) :: hdr(2), dims(7)
   real(real64) :: par(7)
   integer(int32), allocatable :: iao4(:)
   integer(stda_ik), allocatable, target :: iaoat(:)
   integer(stda_ik), allocatable :: iconf(:, :)
   real(c_double), allocatable, target :: clow(:, :), epsi(:), xyz(:, :)
   real(c_double), allocatable :: mom(:, :), ed(:), eci(:), props(:, :)
   real(c_double) :: alpha_sos(6)
   integer :: ncent, nao, no, nv, moci, nroots, nci, nex, nexpt, nroot, st, i, u, ios
   integer(c_int64_t) :: npk
   logical :: velcorr, rpa
   type(c_ptr) :: ctx

   call get_command_argument(1, inp)
   if (trim(inp) == '--index-size') then
      write (*, '(a,i0,a,i0,a,i0)') 'library ', stda_hip_index_size(), ' module ', &
         storage_size(0_stda_ik)/8, ' default integer ', storage_size(0)/8
      stop
   end if
   call get_command_argument(2, outp)
   open (newunit=u, file=trim(inp), access='stream', form='unformatted', status='old', &
         iostat=ios)
   if (ios /= 0) error stop 'tda_harness: cannot open the input file'
   read (u) hdr
   if (hdr(1) /= 1396982849 .or. hdr(2) /= 1) error stop 'tda_harness: not a version-1 input'
   read (u) dims, par
   ncent = dims(1); nao = dims(2); no = dims(3); nv = dims(4); nroots = dims(5)
   velcorr = dims(6) /= 0
   rpa = dims(7) /= 0
   moci = no + nv
   npk = int(moci, c_int64_t)*int(moci + 1, c_int64_t)/2
   allocate (clow(nao, moci), epsi(moci), iao4(nao), iaoat(nao), xyz(4, ncent), mom(npk, 9))
   read (u) clow, epsi, iao4, xyz, mom
   close (u)
   iaoat = int(iao4, stda_ik)      ! as the driver's default integers hold it

   call stda_hip_open(ctx)
   st = stda_hip_setup_arrays(ctx, ncent, nao, no, nv, clow, epsi, iaoat, xyz, &
                              mom(:, 1), mom(:, 2), mom(:, 3), mom(:, 4), mom(:, 5), &
                              mom(:, 6), mom(:, 7), mom(:, 8), mom(:, 9), par(1), par(2), &
                              par(3), par(4), par(5), par(6), .false., .false., &
                              0.5d0, 0.1d0, 4.0d0, nci, nex, nexpt)
   call stda_hip_stop(st, ctx)
   write (*, *) nex, 'CSF included by energy.'
   write (*, *) nexpt, 'considered in PT2.'
   write (*, *) nci - nex, 'CSF included by PT.'
   allocate (iconf(nci, 2), ed(nci), eci(nci), props(5, nci))
   st = stda_hip_csf_list(ctx, iconf, ed)
   call stda_hip_stop(st, ctx)
   if (rpa) then
      st = stda_hip_rpa(ctx, nroots, 1.0d-8, nroot, eci, props, alpha_sos)
   else
      st = stda_hip_tda(ctx, nroots, velcorr, 1.0d-8, nroot, eci, props, alpha_sos)
   end if
   call stda_hip_stop(st, ctx)
   write (*, *) nroot, 'roots'

   open (newunit=u, file=trim(outp), status='replace')
   write (u, *) 'NM'
   write (u, *) 'VELO'
   write (u, *) 'MMASS'
   write (u, *) par(7)
   write (u, *) 'LFAKTOR'
   write (u, *) ' 0.5'
   write (u, *) 'RFAKTOR'
   write (u, *) ' 1.0'
   write (u, *) 'WIDTH'
   write (u, *) ' 0.20'
   write (u, *) 'SHIFT'
   write (u, *) ' 0.00'
   write (u, *) 'DATXY'
   do i = 1, nroot
      write (u, '(i4,F10.4,4f13.6)') i, eci(i)*27.21139d0, props(1:4, i)
   end do
   close (u)
   st = stda_hip_destroy(ctx)
end program tda_harness
