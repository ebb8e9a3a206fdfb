! Test helper: writes or reads the flow-field restart file with Fortran's own
! formatted I/O, in the record layout of the PLES restart (header nscal / ti /
! bbx dbbx list-directed, then format (10e18.10) records: u v w fi rho per
! cell, then uc, vc, wc one per record).  Values travel to and from the test
! as a raw binary stream of doubles.
!   restart_io write <bin-in> <restart-out> jx jy jz nscal
!   restart_io read  <restart-in> <bin-out> jx jy jz nscal
program restart_io
   implicit none
   character(len=256) :: mode, fin, fout, arg
   integer :: jx, jy, jz, nscal, i, j, k, isc, nval
   double precision :: ti, bbx, dbbx
   double precision, allocatable :: cellv(:, :, :, :), uc(:, :, :), vc(:, :, :), wc(:, :, :)
   call get_command_argument(1, mode)
   call get_command_argument(2, fin)
   call get_command_argument(3, fout)
   call get_command_argument(4, arg); read (arg, *) jx
   call get_command_argument(5, arg); read (arg, *) jy
   call get_command_argument(6, arg); read (arg, *) jz
   call get_command_argument(7, arg); read (arg, *) nscal
   nval = 4 + nscal
   allocate (cellv(nval, 0:jx + 1, 0:jy + 1, 0:jz + 1), uc(0:jx, 1:jy, 1:jz), vc(1:jx, 0:jy, 1:jz), &
             wc(1:jx, 1:jy, 0:jz))
   if (trim(mode) == 'write') then
      open (20, file=fin, access='stream', form='unformatted', status='old')
      read (20) ti, bbx, dbbx, cellv, uc, vc, wc
      close (20)
      open (10, file=fout, status='replace')
      write (10, *) nscal
      write (10, *) ti
      write (10, *) bbx, dbbx
      do k = 0, jz + 1
         do j = 0, jy + 1
            do i = 0, jx + 1
               write (10, 1000) (cellv(isc, i, j, k), isc=1, nval)
            end do
         end do
      end do
      do k = 1, jz
         do j = 1, jy
            do i = 0, jx
               write (10, 1000) uc(i, j, k)
            end do
         end do
      end do
      do k = 1, jz
         do j = 0, jy
            do i = 1, jx
               write (10, 1000) vc(i, j, k)
            end do
         end do
      end do
      do k = 0, jz
         do j = 1, jy
            do i = 1, jx
               write (10, 1000) wc(i, j, k)
            end do
         end do
      end do
      close (10)
   else
      open (10, file=fin, status='old')
      read (10, *) i
      read (10, *) ti
      read (10, *) bbx, dbbx
      do k = 0, jz + 1
         do j = 0, jy + 1
            do i = 0, jx + 1
               read (10, 1000) (cellv(isc, i, j, k), isc=1, nval)
            end do
         end do
      end do
      do k = 1, jz
         do j = 1, jy
            do i = 0, jx
               read (10, 1000) uc(i, j, k)
            end do
         end do
      end do
      do k = 1, jz
         do j = 0, jy
            do i = 1, jx
               read (10, 1000) vc(i, j, k)
            end do
         end do
      end do
      do k = 0, jz
         do j = 1, jy
            do i = 1, jx
               read (10, 1000) wc(i, j, k)
            end do
         end do
      end do
      close (10)
      open (20, file=fout, access='stream', form='unformatted', status='replace')
      write (20) ti, bbx, dbbx, cellv, uc, vc, wc
      close (20)
   end if
1000 format(10e18.10)
end program restart_io
