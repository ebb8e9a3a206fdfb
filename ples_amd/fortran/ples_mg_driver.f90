! Fortran driver of the drop-in solve (used by the tests): a warped 32x16x32
! grid, the seeded divergence rhs generated by the library, one ples_multi
! solve from zero, the residual recomputed through ples_mg_residual.
program ples_mg_driver
   use iso_c_binding
   use ples_mg_f
   implicit none
   integer, parameter :: jx = 32, jy = 16, jz = 32, nlev = 3
   real(c_double), parameter :: pi = 3.14159265358979323846d0
   real(c_double), parameter :: lx = 2.d0 * pi, ly = 2.d0, lz = pi, gam = 1.5d0, amp = 0.2d0
   real(c_double), target :: x(0:jx, 0:jy, 0:jz), y(0:jx, 0:jy, 0:jz), z(0:jx, 0:jy, 0:jz)
   real(c_double), allocatable, target :: fi(:, :, :), rhs(:, :, :), r(:, :, :)
   real(c_double), target :: cs1(jy, jz), cs2(jy, jz), cs3(jx, jz), cs4(jx, jz), cs5(jx, jy), cs6(jx, jy)
   integer :: i, j, k, st, cycles, prec, nargs
   real(c_double) :: xi, zk, eta, s, zb, resmax
   character(len=16) :: arg

   prec = PLES_MG_FP64
   nargs = command_argument_count()
   if (nargs >= 1) then
      call get_command_argument(1, arg)
      if (trim(arg) == 'mixed') prec = PLES_MG_MIXED
   end if
   do k = 0, jz
      do j = 0, jy
         do i = 0, jx
            xi = i * lx / jx
            zk = k * lz / jz
            eta = dble(j) / jy
            s = 0.5d0 * (1.d0 + tanh(gam * (2.d0 * eta - 1.d0)) / tanh(gam))
            zb = amp * sin(2.d0 * pi * xi / (lx / 4.d0)) * cos(2.d0 * pi * zk / (lz / 4.d0))
            x(i, j, k) = xi
            y(i, j, k) = zb + (ly - zb) * s
            z(i, j, k) = zk
         end do
      end do
   end do
   cs1 = 0; cs2 = 0; cs3 = 0; cs4 = 0; cs5 = 0; cs6 = 0
   allocate (fi(0:jx + 1, 0:jy + 1, 0:jz + 1), rhs(jx, jy, jz), r(jx, jy, jz))

   call ples_multi_init(jx, jy, jz, 0, 1, 0, nlev, x, y, z, prec, st)
   if (st /= PLES_MG_OK) then
      print '(a,i0)', 'init failed ', st
      stop 1
   end if
   st = ples_mg_case_rhs(ples_multi_handle(), 1002_c_int64_t, 1.d0)
   st = ples_mg_get_rhs(ples_multi_handle(), rhs)
   fi = 0.d0
   call ples_multi(1.d-11, 200, jx, jy, jz, 1, 1, fi, rhs, cs1, cs2, cs3, cs4, cs5, cs6, 1.d0, cycles, resmax, st)
   st = ples_mg_residual(ples_multi_handle(), fi, 1_c_int, 1_c_int, rhs, c_loc(cs1), c_loc(cs2), c_loc(cs3), &
                         c_loc(cs4), c_loc(cs5), c_loc(cs6), 1.d0, r)
   print '(a,i0,a,i0,a,es12.4,a,es12.4,a,es22.14,a,es22.14)', 'status=', st, ' cycles=', cycles, ' resmax=', &
      resmax, ' check=', maxval(abs(r)), ' fi_mean=', sum(fi(1:jx, 1:jy, 1:jz)) / (jx * jy * jz), &
      ' fi_rms=', sqrt(sum((fi(1:jx, 1:jy, 1:jz) - sum(fi(1:jx, 1:jy, 1:jz)) / (jx * jy * jz))**2) / (jx * jy * jz))
   call ples_multi_finalize()
end program ples_mg_driver
