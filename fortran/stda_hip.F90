! ISO_C_BINDING module over the C ABI of the MI355X excited-state core
! (include/stda_hip.h).  A std2 build adds this file and calls it from the RKS
! driver `stda` after the Loewdin MO norm check (stda.f:452), in place of the
! block it replaces (stda.f:454-1471: gamma, charges, screening, PT2, A or A+-B,
! the eigensolve, the property loop; srpapack.f:32-197; linear_response.f:23-369).
!
! Two integer models, selected by the preprocessor as std2's own build selects
! them (meson_options.txt:37-43):
!   default            index arrays (iaoat, iconf, -dual atoms) are 4-byte
!                      integers; link libstda_hip.so;
!   -DSTDA_HIP_ILP64   with -fdefault-integer-8 (-i8): 8-byte index arrays;
!                      link libstda_hip_ilp64.so.
! Scalars cross the boundary as the C types of the header in both models; the
! wrapper procedures below take default integers and convert.
module stda_hip
   use iso_c_binding
   implicit none
   private

#ifdef STDA_HIP_ILP64
   integer, parameter, public :: stda_ik = c_int64_t
#else
   integer, parameter, public :: stda_ik = c_int32_t
#endif

   ! status codes (include/stda_hip.h)
   integer, parameter, public :: STDA_HIP_OK = 0, STDA_HIP_ERR_ARG = 1, STDA_HIP_ERR_NO_CSF = 2, &
      STDA_HIP_ERR_DEVICE = 3, STDA_HIP_ERR_NOMEM = 4, STDA_HIP_ERR_STATE = 5, &
      STDA_HIP_ERR_NOT_CONVERGED = 6, STDA_HIP_ERR_AMB_NOT_PD = 7, STDA_HIP_ERR_NO_ROOTS = 8, &
      STDA_HIP_ERR_TOO_LARGE = 9

   type, bind(C), public :: stda_hip_system
      integer(c_int32_t) :: ncent, nao, no, nv
      type(c_ptr) :: clow, epsi, iaoat, xyz, moments
   end type

   type, bind(C), public :: stda_hip_params
      real(c_double) :: thr_ev, thrp, fthr_ev, ax, alphak, betaj
      integer(c_int32_t) :: triplet, kshift
      real(c_double) :: kshift_max_ev, kshift_width_ev, kshift_steep
   end type

   public :: stda_hip_index_size, stda_hip_create, stda_hip_destroy, stda_hip_setup, &
      stda_hip_get_csf, stda_hip_solve_tda, stda_hip_solve_rpa, stda_hip_lresp, &
      stda_hip_set_precision, stda_hip_set_selection, stda_hip_build_a, &
      stda_hip_build_apb_amb, stda_hip_build_bhalf, stda_hip_matvec
   public :: stda_hip_message, stda_hip_stop, stda_hip_open, stda_hip_setup_arrays, &
      stda_hip_csf_list, stda_hip_tda, stda_hip_rpa, stda_hip_alpha

   interface
      integer(c_int) function stda_hip_index_size() bind(C)
         import
      end function
      integer(c_int) function stda_hip_create(ctx, device) bind(C)
         import
         type(c_ptr) :: ctx
         integer(c_int), value :: device
      end function
      integer(c_int) function stda_hip_destroy(ctx) bind(C)
         import
         type(c_ptr), value :: ctx
      end function
      type(c_ptr) function stda_hip_error_string(st) bind(C)
         import
         integer(c_int), value :: st
      end function
      type(c_ptr) function stda_hip_last_error(ctx) bind(C)
         import
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function stda_hip_setup(ctx, sys, par, nci, nprim, nsec) bind(C)
         import
         type(c_ptr), value :: ctx
         type(stda_hip_system), intent(in) :: sys
         type(stda_hip_params), intent(in) :: par
         integer(c_int64_t) :: nci, nprim, nsec
      end function
      integer(c_int) function stda_hip_get_csf(ctx, iconf, ed) bind(C)
         import
         type(c_ptr), value :: ctx
         integer(stda_ik) :: iconf(*)
         real(c_double) :: ed(*)
      end function
      integer(c_int) function stda_hip_solve_tda(ctx, nroots, capacity, velcorr, tol, nfound, &
                                                 omega, x, props, alpha_sos) bind(C)
         import
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nroots, capacity, velcorr
         real(c_double), value :: tol
         integer(c_int32_t) :: nfound
         real(c_double) :: omega(*), props(*), alpha_sos(6)
         type(c_ptr), value :: x
      end function
      integer(c_int) function stda_hip_solve_rpa(ctx, nroots, capacity, tol, nfound, omega, x, y, &
                                                 props, alpha_sos) bind(C)
         import
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nroots, capacity
         real(c_double), value :: tol
         integer(c_int32_t) :: nfound
         real(c_double) :: omega(*), props(*), alpha_sos(6)
         type(c_ptr), value :: x, y
      end function
      integer(c_int) function stda_hip_lresp(ctx, nfreq, omega, tol, alpha, xpy) bind(C)
         import
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: nfreq
         real(c_double) :: omega(*), alpha(*)
         real(c_double), value :: tol
         type(c_ptr), value :: xpy
      end function
      integer(c_int) function stda_hip_set_precision(ctx, precision) bind(C)
         import
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: precision
      end function
      integer(c_int) function stda_hip_set_selection(ctx, thr2_ev, natoms, atoms, pt2_off) bind(C)
         import
         type(c_ptr), value :: ctx
         real(c_double), value :: thr2_ev
         integer(c_int32_t), value :: natoms, pt2_off
         integer(stda_ik) :: atoms(*)
      end function
      integer(c_int) function stda_hip_build_a(ctx, a) bind(C)
         import
         type(c_ptr), value :: ctx
         real(c_double) :: a(*)
      end function
      integer(c_int) function stda_hip_build_apb_amb(ctx, apb, amb) bind(C)
         import
         type(c_ptr), value :: ctx
         real(c_double) :: apb(*), amb(*)
      end function
      integer(c_int) function stda_hip_build_bhalf(ctx, bhalf) bind(C)
         import
         type(c_ptr), value :: ctx
         real(c_double) :: bhalf(*)
      end function
      integer(c_int) function stda_hip_matvec(ctx, which, nvec, v, av) bind(C)
         import
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: which, nvec
         real(c_double) :: v(*), av(*)
      end function
      integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
         import
         type(c_ptr), value :: s
      end function
   end interface

contains

   ! Fortran string of a C string owned by the library
   function c_to_f(p) result(s)
      type(c_ptr), intent(in) :: p
      character(len=:), allocatable :: s
      character(kind=c_char), pointer :: a(:)
      integer :: n, i
      if (.not. c_associated(p)) then
         s = ''
         return
      end if
      n = int(c_strlen(p))
      call c_f_pointer(p, a, [n])
      allocate (character(len=n) :: s)
      do i = 1, n
         s(i:i) = a(i)
      end do
   end function

   ! The message the reference stops with for a status, plus the context's detail.
   function stda_hip_message(st, ctx) result(msg)
      integer, intent(in) :: st
      type(c_ptr), intent(in) :: ctx
      character(len=:), allocatable :: msg, detail
      msg = c_to_f(stda_hip_error_string(int(st, c_int)))
      if (c_associated(ctx)) then
         detail = c_to_f(stda_hip_last_error(ctx))
         if (len(detail) > 0) msg = msg//': '//detail
      end if
   end function

   ! The reference stops the program on these errors (e.g. stda.f:644,
   ! srpapack.f:92,229); so does the binding.
   subroutine stda_hip_stop(st, ctx)
      integer, intent(in) :: st
      type(c_ptr), intent(in) :: ctx
      if (st == STDA_HIP_OK) return
      write (*, '(a)') stda_hip_message(st, ctx)
      error stop 1
   end subroutine

   ! Create a context on the current device after checking the index model.
   subroutine stda_hip_open(ctx)
      type(c_ptr), intent(out) :: ctx
      integer :: st
      ctx = c_null_ptr
      if (stda_hip_index_size() /= storage_size(0_stda_ik)/8) then
         write (*, '(a)') 'stda_hip: library and caller differ in the index integer size'
         error stop 1
      end if
      st = stda_hip_create(ctx, -1_c_int)
      call stda_hip_stop(st, ctx)
   end subroutine

   ! Setup (stda.f:454-790 and ptselect) from the arrays as `stda` holds them at
   ! stda.f:452: clow(nao,moci), epsi(moci), iaoat(nao) (default integers),
   ! xyz(4,ncent) and the nine packed moment arrays xl..zv (moci(moci+1)/2 each,
   ! stda.f:346-412); thr in eV as std2's -e, fthr <= 0 for "unlimited".
   function stda_hip_setup_arrays(ctx, ncent, nao, no, nv, clow, epsi, iaoat, xyz, &
                                  xl, yl, zl, xm, ym, zm, xv, yv, zv, thr_ev, thrp, fthr_ev, &
                                  ax, alphak, betaj, triplet, dokshift, shftmax, shftwidth, &
                                  shftsteep, nci, nex, nexpt) result(st)
      type(c_ptr), intent(in) :: ctx
      integer, intent(in) :: ncent, nao, no, nv
      real(c_double), intent(in), target :: clow(nao, no + nv), epsi(no + nv), xyz(4, ncent)
      integer(stda_ik), intent(in), target :: iaoat(nao)
      real(c_double), intent(in) :: xl(*), yl(*), zl(*), xm(*), ym(*), zm(*), xv(*), yv(*), zv(*)
      real(c_double), intent(in) :: thr_ev, thrp, fthr_ev, ax, alphak, betaj
      logical, intent(in) :: triplet, dokshift
      real(c_double), intent(in) :: shftmax, shftwidth, shftsteep
      integer, intent(out) :: nci, nex, nexpt
      integer :: st
      integer(c_int64_t) :: n8, p8, s8
      integer(c_int64_t) :: npk
      real(c_double), allocatable, target :: mom(:)
      type(stda_hip_system) :: sy
      type(stda_hip_params) :: pa
      npk = int(no + nv, c_int64_t)*int(no + nv + 1, c_int64_t)/2
      allocate (mom(9*npk))
      mom(1:npk) = xl(1:npk)
      mom(npk + 1:2*npk) = yl(1:npk)
      mom(2*npk + 1:3*npk) = zl(1:npk)
      mom(3*npk + 1:4*npk) = xm(1:npk)
      mom(4*npk + 1:5*npk) = ym(1:npk)
      mom(5*npk + 1:6*npk) = zm(1:npk)
      mom(6*npk + 1:7*npk) = xv(1:npk)
      mom(7*npk + 1:8*npk) = yv(1:npk)
      mom(8*npk + 1:9*npk) = zv(1:npk)
      sy%ncent = int(ncent, c_int32_t)
      sy%nao = int(nao, c_int32_t)
      sy%no = int(no, c_int32_t)
      sy%nv = int(nv, c_int32_t)
      sy%clow = c_loc(clow)
      sy%epsi = c_loc(epsi)
      sy%iaoat = c_loc(iaoat)
      sy%xyz = c_loc(xyz)
      sy%moments = c_loc(mom)
      pa%thr_ev = thr_ev
      pa%thrp = thrp
      pa%fthr_ev = fthr_ev
      pa%ax = ax
      pa%alphak = alphak
      pa%betaj = betaj
      pa%triplet = merge(1_c_int32_t, 0_c_int32_t, triplet)
      pa%kshift = merge(1_c_int32_t, 0_c_int32_t, dokshift)
      pa%kshift_max_ev = shftmax
      pa%kshift_width_ev = shftwidth
      pa%kshift_steep = shftsteep
      st = int(stda_hip_setup(ctx, sy, pa, n8, p8, s8))
      nci = int(n8)
      nex = int(p8)
      nexpt = int(s8)
   end function

   ! iconf(nci,2) (occupied MO, absolute virtual MO; 1-based) and ed(nci).
   function stda_hip_csf_list(ctx, iconf, ed) result(st)
      type(c_ptr), intent(in) :: ctx
      integer(stda_ik), intent(out) :: iconf(:, :)
      real(c_double), intent(out) :: ed(:)
      integer :: st
      st = int(stda_hip_get_csf(ctx, iconf, ed))
   end function

   ! TDA roots: nroots <= 0 gives every root in (0, Ethr] as ssyevr does
   ! (stda.f:1134-1150).  props(5, :) = fL, fV, RL, RV or Rv(corr), 0.
   ! x(nci, capacity) receives the eigenvectors when present.
   function stda_hip_tda(ctx, nroots, velcorr, tol, nroot, eci, props, alpha_sos, x) result(st)
      type(c_ptr), intent(in) :: ctx
      integer, intent(in) :: nroots
      logical, intent(in) :: velcorr
      real(c_double), intent(in) :: tol
      integer, intent(out) :: nroot
      real(c_double), intent(out) :: eci(:), props(:, :), alpha_sos(6)
      real(c_double), intent(out), optional, target :: x(:, :)
      integer :: st
      integer(c_int32_t) :: nf
      type(c_ptr) :: px
      px = c_null_ptr
      if (present(x)) px = c_loc(x)
      st = int(stda_hip_solve_tda(ctx, int(nroots, c_int32_t), int(size(eci), c_int32_t), &
                                  merge(1_c_int32_t, 0_c_int32_t, velcorr), tol, nf, eci, px, &
                                  props, alpha_sos))
      nroot = int(nf)
   end function

   ! sTD-DFT roots (srpapack), X and Y optional (nci, capacity).
   function stda_hip_rpa(ctx, nroots, tol, nroot, eci, props, alpha_sos, x, y) result(st)
      type(c_ptr), intent(in) :: ctx
      integer, intent(in) :: nroots
      real(c_double), intent(in) :: tol
      integer, intent(out) :: nroot
      real(c_double), intent(out) :: eci(:), props(:, :), alpha_sos(6)
      real(c_double), intent(out), optional, target :: x(:, :), y(:, :)
      integer :: st
      integer(c_int32_t) :: nf
      type(c_ptr) :: px, py
      px = c_null_ptr
      py = c_null_ptr
      if (present(x)) px = c_loc(x)
      if (present(y)) py = c_loc(y)
      st = int(stda_hip_solve_rpa(ctx, int(nroots, c_int32_t), int(size(eci), c_int32_t), tol, &
                                  nf, eci, px, py, props, alpha_sos))
      nroot = int(nf)
   end function

   ! alpha(3,3,nfreq) at the frequencies w (Hartree) (lresp, linear_response.f:23-192).
   function stda_hip_alpha(ctx, w, tol, alpha) result(st)
      type(c_ptr), intent(in) :: ctx
      real(c_double), intent(in) :: w(:), tol
      real(c_double), intent(out) :: alpha(:, :, :)
      integer :: st
      real(c_double), allocatable :: wc(:)
      wc = w
      st = int(stda_hip_lresp(ctx, int(size(w), c_int32_t), wc, tol, alpha, c_null_ptr))
   end function

end module stda_hip
