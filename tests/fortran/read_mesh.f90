! The READ statements of best/best.f:189-220 in a program of their own, for the tests of the
! library's mesh reader (tests/test_files.py): reads the mesh file named on the command line the
! way BEST does (list-directed, one READ for the header and one per triangle) and prints NAR,
! TRUEAREA, TRUEVOL and every coordinate in hexadecimal bit patterns, one per line.
program read_mesh
   implicit none
   character(len=512) :: fname, arg
   integer :: iarea, nar, n, k
   real(8) :: truearea, truevol
   real(8), allocatable :: v1(:,:), v2(:,:), v3(:,:)
   call get_command_argument(1, fname)
   call get_command_argument(2, arg)
   read(arg, *) iarea
   truearea = 0.0d0
   truevol = 0.0d0
   open(unit=10, file=trim(fname), status='old', action='read')
   if (iarea == 1) then
      read(10, *) nar, truearea, truevol
   else
      read(10, *) nar
   end if
   allocate(v1(nar,3), v2(nar,3), v3(nar,3))
   do n = 1, nar
      read(10, *) v1(n,1), v1(n,2), v1(n,3), v2(n,1), v2(n,2), v2(n,3), v3(n,1), v3(n,2), v3(n,3)
   end do
   close(10)
   print '(i0)', nar
   print '(z16.16)', truearea
   print '(z16.16)', truevol
   do n = 1, nar
      do k = 1, 3
         print '(z16.16)', v1(n,k)
      end do
      do k = 1, 3
         print '(z16.16)', v2(n,k)
      end do
      do k = 1, 3
         print '(z16.16)', v3(n,k)
      end do
   end do
end program read_mesh
