! best_hip -- iso_c_binding interface to libbesthip.so (include/best_hip.h) for Fortran hosts.
!
! Part 1 of the C header is the H-matrix ABI that BEST's module besolve already binds
! (best/besolve.f90:48-89); this module binds Part 2, the handle API: one solver per handle, the
! mesh and its factors resident on the GPU, and the whole solve of best/best.f:261-498 (without
! file I/O) in one call.  Arrays follow BEST: V1, V2, V3 are (NAR,3) REAL(8), column-major.
module best_hip
   use, intrinsic :: iso_c_binding
   implicit none
   private
   public :: bhg_result, bhg_stats, bhg_be_params, bhg_create, bhg_destroy, bhg_set_mesh, bhg_solve_mesh, &
             bhg_get_stats, bhg_get_solution, bhg_besolv, bhg_error, bhg_write_be

   ! struct bhg_result: the values of BEST's .be file and the tensors behind them
   type, bind(C) :: bhg_result
      real(c_double) :: dtt(9), dtr(9), drr(9)
      real(c_double) :: urc(3), udr(3)
      real(c_double) :: ktt(9), krt(9), krr(9)
      real(c_double) :: drr_eig(3), dtt_eig(3)
      real(c_double) :: dr, dtrans, chiprime, chi
      real(c_double) :: dvisc(3)
      real(c_double) :: afac, totarea, centroid(3), diagm
      integer(c_int) :: ncomp, nopen, ndef
   end type bhg_result

   ! struct bhg_stats: phase timings (ms) and diagnostics of the last call
   type, bind(C) :: bhg_stats
      real(c_double) :: t_prep, t_assemble, t_factor, t_solve, t_reduce, t_total
      integer(c_int) :: ir_iterations
      real(c_double) :: ir_correction
      integer(c_int) :: pivoted, npad
      real(c_double) :: t_adjoint_solve, t_adjoint_pairs, t_adjoint_tri
      integer(c_int) :: ir_above_tol
   end type bhg_stats

   ! struct bhg_be_params: what OUTPUT (best/output.f:1-3) takes besides the result
   type, bind(C) :: bhg_be_params
      type(c_ptr) :: ver
      integer(c_int) :: iarea
      real(c_double) :: truearea, truevol
      integer(c_int) :: mw
      real(c_double) :: wm
      integer(c_int) :: nar, ivisco, icov
      real(c_double) :: cg
   end type bhg_be_params

   interface
      type(c_ptr) function bhg_create(device) bind(C, name='bhg_create')
         import :: c_ptr, c_int
         integer(c_int), value :: device
      end function bhg_create
      subroutine bhg_destroy(s) bind(C, name='bhg_destroy')
         import :: c_ptr
         type(c_ptr), value :: s
      end subroutine bhg_destroy
      type(c_ptr) function bhg_last_error(s) bind(C, name='bhg_last_error')
         import :: c_ptr
         type(c_ptr), value :: s
      end function bhg_last_error
      integer(c_int) function bhg_set_mesh(s, n, v1, v2, v3, truearea, truevol, icomp) bind(C, name='bhg_set_mesh')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: s, icomp
         integer(c_int), value :: n
         real(c_double), intent(in) :: v1(*), v2(*), v3(*)
         real(c_double), value :: truearea, truevol
      end function bhg_set_mesh
      integer(c_int) function bhg_solve_mesh(s, iarea, cg, ivisco, icov, precision, res) bind(C, name='bhg_solve_mesh')
         import :: c_ptr, c_int, c_double, bhg_result
         type(c_ptr), value :: s
         integer(c_int), value :: iarea, ivisco, icov, precision
         real(c_double), value :: cg
         type(bhg_result), intent(out) :: res
      end function bhg_solve_mesh
      integer(c_int) function bhg_get_stats(s, st) bind(C, name='bhg_get_stats')
         import :: c_ptr, c_int, bhg_stats
         type(c_ptr), value :: s
         type(bhg_stats), intent(out) :: st
      end function bhg_get_stats
      integer(c_int) function bhg_get_solution(s, nrhs, out, ldo) bind(C, name='bhg_get_solution')
         import :: c_ptr, c_int, c_long, c_double
         type(c_ptr), value :: s
         integer(c_int), value :: nrhs
         real(c_double), intent(out) :: out(*)
         integer(c_long), value :: ldo
      end function bhg_get_solution
      integer(c_int) function bhg_besolv(s, nrhs, b, ldb) bind(C, name='bhg_besolv')
         import :: c_ptr, c_int, c_long, c_double
         type(c_ptr), value :: s
         integer(c_int), value :: nrhs
         real(c_double), intent(inout) :: b(*)
         integer(c_long), value :: ldb
      end function bhg_besolv
      integer(c_int) function bhg_write_be(path, res, par, err, errlen) bind(C, name='bhg_write_be')
         import :: c_int, c_char, bhg_result, bhg_be_params
         character(kind=c_char), intent(in) :: path(*)
         type(bhg_result), intent(in) :: res
         type(bhg_be_params), intent(in) :: par
         character(kind=c_char), intent(out) :: err(*)
         integer(c_int), value :: errlen
      end function bhg_write_be
      integer(c_size_t) function c_strlen(p) bind(C, name='strlen')
         import :: c_ptr, c_size_t
         type(c_ptr), value :: p
      end function c_strlen
   end interface

contains

   ! The solver's last error message as a Fortran string.
   function bhg_error(s) result(msg)
      type(c_ptr), intent(in) :: s
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      p = bhg_last_error(s)
      if (.not. c_associated(p)) then
         msg = ''
         return
      end if
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function bhg_error

end module best_hip
