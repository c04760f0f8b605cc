! bh_solve -- one BEST mesh solved on the GPU from Fortran, through the best_hip module.
!
! Usage: bh_solve MESHFILE [PRECISION [IAREA [IVISCO [ICOV [WM]]]]]
!   MESHFILE   BEST's input: NAR TRUEAREA TRUEVOL, then 9 vertex coordinates per triangle
!              (best/best.f:189-219)
!   PRECISION  32 (FP32 LU on the matrix cores + FP64 refinement, default) or 64 (FP64 LU)
!   IAREA      1 area correction (BEST's default), 0 as -na;  IVISCO 1 as -v;  ICOV 1 as -vc
!   WM         molecular weight as -mw (g/mol; 0 or absent: none)
! Prints D_t (cm^2/s), D_r (1/s), CHI, CHI*V and the centre vectors in full precision, and writes
! BEST's .be file beside the mesh, named as BEST names it (best/best.f:224-249: the mesh file name,
! then v, c, m, n for -v, -vc, -mw, -na, then .be).
program bh_solve
   use, intrinsic :: iso_c_binding
   use best_hip
   implicit none
   character(len=512) :: fname, arg
   integer :: nar, n, prec, iarea, ivisco, icov, rc
   real(c_double) :: truearea, truevol, wm
   character(len=600) :: befile
   character(kind=c_char, len=256), target :: err
   character(kind=c_char, len=17), target :: ver
   type(bhg_be_params) :: par
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
   wm = 0.0_c_double
   if (command_argument_count() >= 6) then
      call get_command_argument(6, arg)
      read(arg, *) wm
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

   befile = trim(fname)
   if (ivisco == 1) befile = trim(befile)//'v'
   if (icov == 1) befile = trim(befile)//'c'
   if (wm > 0.0_c_double) befile = trim(befile)//'m'
   if (iarea == 0) befile = trim(befile)//'n'
   befile = trim(befile)//'.be'
   ver = '4.0 SFSU/ARAGON'//c_null_char
   par%ver = c_loc(ver)
   par%iarea = int(iarea, c_int)
   par%truearea = truearea
   par%truevol = truevol
   par%mw = merge(1_c_int, 0_c_int, wm > 0.0_c_double)
   par%wm = wm
   par%nar = int(nar, c_int)
   par%ivisco = int(ivisco, c_int)
   par%icov = int(icov, c_int)
   par%cg = 1.0_c_double
   rc = bhg_write_be(trim(befile)//c_null_char, r, par, err, int(len(err), c_int))
   if (rc /= 0) then
      print '(2a)', 'bh_solve: ', err(1:index(err, c_null_char) - 1)
      stop 1
   end if
   print '(2a)', 'wrote ', trim(befile)
   call bhg_destroy(s)
end program bh_solve
