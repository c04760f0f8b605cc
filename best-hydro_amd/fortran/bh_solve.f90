! bh_solve -- one BEST mesh solved on the GPU from Fortran, through the best_hip module.
!
! Usage: bh_solve MESHFILE [PRECISION [IAREA [IVISCO [ICOV]]]]
!   MESHFILE   BEST's input: NAR TRUEAREA TRUEVOL, then 9 vertex coordinates per triangle
!              (best/best.f:189-219)
!   PRECISION  32 (FP32 LU on the matrix cores + FP64 refinement, default) or 64 (FP64 LU)
!   IAREA      1 area correction (BEST's default), 0 as -na;  IVISCO 1 as -v;  ICOV 1 as -vc
! Prints D_t (cm^2/s), D_r (1/s), CHI, CHI*V and the centre vectors in full precision.
program bh_solve
   use, intrinsic :: iso_c_binding
   use best_hip
   implicit none
   character(len=512) :: fname, arg
   integer :: nar, n, prec, iarea, ivisco, icov, rc
   real(c_double) :: truearea, truevol
   real(c_double), allocatable :: v1(:,:), v2(:,:), v3(:,:)
   type(c_ptr) :: s
   type(bhg_result) :: r
   type(bhg_stats) :: st

   if (command_argument_count() < 1) then
      print '(a)', 'usage: bh_solve MESHFILE [32|64 [IAREA [IVISCO [ICOV]]]]'
      stop 2
   end if
   call get_command_argument(1, fname)
   prec = 32
   iarea = 1
   ivisco = 1
   icov = 0
   if (command_argument_count() >= 2) then
      call get_command_argument(2, arg)
      read(arg, *) prec
   end if
   if (command_argument_count() >= 3) then
      call get_command_argument(3, arg)
      read(arg, *) iarea
   end if
   if (command_argument_count() >= 4) then
      call get_command_argument(4, arg)
      read(arg, *) ivisco
   end if
   if (command_argument_count() >= 5) then
      call get_command_argument(5, arg)
      read(arg, *) icov
   end if

   open(10, file=trim(fname), status='old', action='read')
   read(10, *) nar, truearea, truevol
   allocate(v1(nar,3), v2(nar,3), v3(nar,3))
   do n = 1, nar
      read(10, *) v1(n,1), v1(n,2), v1(n,3), v2(n,1), v2(n,2), v2(n,3), v3(n,1), v3(n,2), v3(n,3)
   end do
   close(10)

   s = bhg_create(0_c_int)
   if (.not. c_associated(s)) then
      print '(a)', 'bh_solve: no GPU solver could be created'
      stop 1
   end if
   rc = bhg_set_mesh(s, int(nar, c_int), v1, v2, v3, truearea, truevol, c_null_ptr)
   if (rc /= 0) then
      print '(2a)', 'bh_solve: ', bhg_error(s)
      stop 1
   end if
   rc = bhg_solve_mesh(s, int(iarea, c_int), 1.0_c_double, int(ivisco, c_int), int(icov, c_int), int(prec, c_int), r)
   if (rc /= 0) then
      print '(2a)', 'bh_solve: ', bhg_error(s)
      stop 1
   end if
   rc = bhg_get_stats(s, st)
   print '(a,i0,a,i0,a,i0)', 'triangles ', nar, '  surfaces ', r%ncomp, '  precision ', prec
   print '(a,es24.16)', 'DTRANS (cm^2/s) ', r%dtrans
   print '(a,es24.16)', 'DR (1/s)        ', r%dr
   print '(a,es24.16)', 'CHI             ', r%chi
   print '(a,es24.16)', 'CHI*V (A^3)     ', r%chiprime
   print '(a,3es24.16)', 'URC (A)         ', r%urc
   print '(a,3es24.16)', 'DVISC (A)       ', r%dvisc
   print '(a,f10.3,a,i0)', 'total ms ', st%t_total, '  refinement steps ', st%ir_iterations
   call bhg_destroy(s)
end program bh_solve
