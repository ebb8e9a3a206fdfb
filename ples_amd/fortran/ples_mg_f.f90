! ples_mg_f -- Fortran 2003 binding of the MI355X pressure multigrid
! (include/ples_mg.h) for the PLES time loop.
!
! Drop-in for the reference pressure solve
!     call multi(eps, ficycle, nlevel, jxc, jyc, jzc, kparasta, kparaend, ...)
! (fsrc/stratiParticle.f03:2881-2887, fsrc/multigrid.f03:42-653): the caller
! creates one solver handle at init, after the grid nodes are known (where the
! reference runs metrica -> mul_met -> wall, fsrc/stratiParticle.f03:767-769),
! and calls ples_multi each time step with the same pressure, rhs, boundary
! flux and time step arrays that multi reads from its modules.  All arrays are
! passed in place (first index fastest, the reference's own lower bounds), no
! copies on the Fortran side.
module ples_mg_f
   use iso_c_binding
   implicit none
   private

   integer(c_int), parameter, public :: PLES_MG_OK = 0
   integer(c_int), parameter, public :: PLES_MG_NOT_CONVERGED = 1
   integer(c_int), parameter, public :: PLES_MG_FP64 = 0
   integer(c_int), parameter, public :: PLES_MG_MIXED = 1
   integer(c_int), parameter, public :: PLES_MG_STOP_REL_L2 = 0
   integer(c_int), parameter, public :: PLES_MG_STOP_ABS_MAX = 1

   type, bind(C), public :: ples_mg_grid
      integer(c_int) :: jx, jy, jz
      integer(c_int) :: ip, jp, kp
      integer(c_int) :: nlevel
      type(c_ptr) :: x, y, z
      type(c_ptr) :: g(9)
   end type ples_mg_grid

   type, bind(C), public :: ples_mg_opts
      integer(c_int) :: precision
      integer(c_int) :: nu_pre, nu_post, nu_coarse
      real(c_double) :: omega
      integer(c_int) :: krylov, stop_mode, project_rhs, halo_passes, device, graphs, inner_cycles
      integer(c_int) :: autotune, fp32_krylov, smoother_fp16, coarse_exact, sweep_schedule, semicoarsen
   end type ples_mg_opts

   type, bind(C), public :: ples_mg_result
      integer(c_int) :: status, cycles
      real(c_double) :: rel_res, abs_res, rhs_mean_removed, seconds
   end type ples_mg_result

   public :: ples_mg_default_opts, ples_mg_create, ples_mg_destroy, ples_mg_solve
   public :: ples_mg_residual, ples_mg_case_rhs, ples_mg_get_rhs
   public :: ples_mg_diver, ples_mg_gradie, ples_mg_update, ples_mg_contra, ples_mg_vel_up
   public :: ples_mg_check_divergence, ples_mg_projection_step
   public :: ples_multi_init, ples_multi, ples_multi_finalize, ples_multi_handle

   interface
      subroutine ples_mg_default_opts(opts) bind(C, name="ples_mg_default_opts")
         import :: ples_mg_opts
         type(ples_mg_opts), intent(out) :: opts
      end subroutine

      integer(c_int) function ples_mg_create(grid, opts, h) bind(C, name="ples_mg_create")
         import :: c_int, c_ptr, ples_mg_grid, ples_mg_opts
         type(ples_mg_grid), intent(in) :: grid
         type(ples_mg_opts), intent(in) :: opts
         type(c_ptr), intent(out) :: h
      end function

      integer(c_int) function ples_mg_destroy(h) bind(C, name="ples_mg_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function

      ! launch-shape table for the handles created afterwards; path is a
      ! NUL-terminated string: trim(path)//c_null_char
      integer(c_int) function ples_mg_tuning_load(path) bind(C, name="ples_mg_tuning_load")
         import :: c_int, c_char
         character(kind=c_char), dimension(*), intent(in) :: path
      end function

      integer(c_int) function ples_mg_solve(h, fi, deepl, deepr, rhs, cs1, cs2, cs3, cs4, cs5, cs6, &
                                            dt, eps, ficycle, res) bind(C, name="ples_mg_solve")
         import :: c_int, c_ptr, c_double, ples_mg_result
         type(c_ptr), value :: h
         real(c_double), intent(inout) :: fi(*)
         integer(c_int), value :: deepl, deepr
         real(c_double), intent(in) :: rhs(*)
         type(c_ptr), value :: cs1, cs2, cs3, cs4, cs5, cs6
         real(c_double), value :: dt, eps
         integer(c_int), value :: ficycle
         type(ples_mg_result), intent(out) :: res
      end function

      integer(c_int) function ples_mg_residual(h, fi, deepl, deepr, rhs, cs1, cs2, cs3, cs4, cs5, cs6, &
                                               dt, r) bind(C, name="ples_mg_residual")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: fi(*)
         integer(c_int), value :: deepl, deepr
         real(c_double), intent(in) :: rhs(*)
         type(c_ptr), value :: cs1, cs2, cs3, cs4, cs5, cs6
         real(c_double), value :: dt
         real(c_double), intent(out) :: r(*)
      end function

      integer(c_int) function ples_mg_case_rhs(h, seed, dt) bind(C, name="ples_mg_case_rhs")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: h
         integer(c_int64_t), value :: seed
         real(c_double), value :: dt
      end function

      integer(c_int) function ples_mg_get_rhs(h, rhs) bind(C, name="ples_mg_get_rhs")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: rhs(*)
      end function

      ! diver (fsrc/diver.f03:2-46): rhs from uc(0:jx,1:jy,1:jz), vc(1:jx,0:jy,1:jz), wc(1:jx,1:jy,0:jz)
      integer(c_int) function ples_mg_diver(h, uc, vc, wc, dt, rhs) bind(C, name="ples_mg_diver")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: uc(*), vc(*), wc(*)
         real(c_double), value :: dt
         real(c_double), intent(out) :: rhs(*)
      end function

      ! gradie (fsrc/gradie.f03:2-120); outputs may be c_null_ptr
      integer(c_int) function ples_mg_gradie(h, fi, deepl, deepr, dt, gra1, gra2, gra3, cgra1, cgra2, cgra3) &
         bind(C, name="ples_mg_gradie")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: fi(*)
         integer(c_int), value :: deepl, deepr
         real(c_double), value :: dt
         type(c_ptr), value :: gra1, gra2, gra3, cgra1, cgra2, cgra3
      end function

      ! update (fsrc/update.f03:2-289): u,v,w shaped like fi; delu.. (0:jx+1,0:jy+1,0:jz+1)
      integer(c_int) function ples_mg_update(h, u, v, w, deepl, deepr, delu, delv, delw) bind(C, name="ples_mg_update")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(inout) :: u(*), v(*), w(*)
         integer(c_int), value :: deepl, deepr
         real(c_double), intent(in) :: delu(*), delv(*), delw(*)
      end function

      ! contra (fsrc/contra.f03:2-241): cs(1:6) = c_loc(cs1) .. c_loc(cs6) (or c_null_ptr)
      integer(c_int) function ples_mg_contra(h, u, v, w, deepl, deepr, uc, vc, wc, cs, potenziale) &
         bind(C, name="ples_mg_contra")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: u(*), v(*), w(*)
         integer(c_int), value :: deepl, deepr
         real(c_double), intent(inout) :: uc(*), vc(*), wc(*)
         type(c_ptr), intent(in) :: cs(6)
         integer(c_int), value :: potenziale
      end function

      ! vel_up (fsrc/vel_up.f03:2-313): gra(1:3), cgra(1:3) = c_loc of the gradient arrays
      integer(c_int) function ples_mg_vel_up(h, u, v, w, deepl, deepr, uc, vc, wc, gra, cgra, dt, potenziale) &
         bind(C, name="ples_mg_vel_up")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(inout) :: u(*), v(*), w(*)
         integer(c_int), value :: deepl, deepr
         real(c_double), intent(inout) :: uc(*), vc(*), wc(*)
         type(c_ptr), intent(in) :: gra(3), cgra(3)
         real(c_double), value :: dt
         integer(c_int), value :: potenziale
      end function

      ! check_divergence (fsrc/check_divergence.f03:131-190)
      integer(c_int) function ples_mg_check_divergence(h, uc, vc, wc, out) bind(C, name="ples_mg_check_divergence")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: uc(*), vc(*), wc(*)
         real(c_double), intent(out) :: out(4)
      end function

      ! the projection step update -> contra -> diver -> multi -> gradie -> vel_up -> check_divergence
      integer(c_int) function ples_mg_projection_step(h, u, v, w, deepl, deepr, delu, delv, delw, uc, vc, wc, cs, &
                                                      fi, dt, eps, ficycle, potenziale, res, div) &
         bind(C, name="ples_mg_projection_step")
         import :: c_int, c_ptr, c_double, ples_mg_result
         type(c_ptr), value :: h
         real(c_double), intent(inout) :: u(*), v(*), w(*)
         integer(c_int), value :: deepl, deepr
         real(c_double), intent(in) :: delu(*), delv(*), delw(*)
         real(c_double), intent(inout) :: uc(*), vc(*), wc(*)
         type(c_ptr), intent(in) :: cs(6)
         real(c_double), intent(inout) :: fi(*)
         real(c_double), value :: dt, eps
         integer(c_int), value :: ficycle, potenziale
         type(ples_mg_result), intent(out) :: res
         real(c_double), intent(out) :: div(4)
      end function
   end interface

   type(c_ptr), save :: handle = c_null_ptr

contains

   ! Create the solver at init.  x, y, z: grid nodes (0:jx, 0:jy, 0:jz)
   ! (the interior part of the reference's x(-8:n1+8, -8:n2+8, ...) arrays).
   ! precision: PLES_MG_FP64 or PLES_MG_MIXED.  status /= 0 on failure.
   subroutine ples_multi_init(jx, jy, jz, ip, jp, kp, nlevel, x, y, z, precision, status)
      integer, intent(in) :: jx, jy, jz, ip, jp, kp, nlevel, precision
      real(c_double), intent(in), target :: x(0:jx, 0:jy, 0:jz), y(0:jx, 0:jy, 0:jz), z(0:jx, 0:jy, 0:jz)
      integer, intent(out) :: status
      type(ples_mg_grid) :: g
      type(ples_mg_opts) :: o
      integer :: m
      g%jx = jx; g%jy = jy; g%jz = jz
      g%ip = ip; g%jp = jp; g%kp = kp
      g%nlevel = nlevel
      g%x = c_loc(x); g%y = c_loc(y); g%z = c_loc(z)
      do m = 1, 9
         g%g(m) = c_null_ptr
      end do
      call ples_mg_default_opts(o)
      o%precision = precision
      o%stop_mode = PLES_MG_STOP_ABS_MAX     ! the reference's max-norm test
      status = ples_mg_create(g, o, handle)
   end subroutine

   ! The drop-in for `call multi(...)`: same inputs as multi reads from its
   ! modules (fi with its ghost layers, rhs, cs1..cs6, dt) and the same
   ! eps / ficycle.  fi is overwritten with the solution incl. ghosts.
   ! deepl, deepr: fi's k ghost depth as allocated (fsrc/iniz.f03:98).
   subroutine ples_multi(eps, ficycle, jx, jy, jz, deepl, deepr, fi, rhs, cs1, cs2, cs3, cs4, cs5, cs6, &
                         dt, cycles, resmax, status)
      real(c_double), intent(in) :: eps, dt
      integer, intent(in) :: ficycle, jx, jy, jz, deepl, deepr
      real(c_double), intent(inout) :: fi(0:jx + 1, 0:jy + 1, 1 - deepl:jz + deepr)
      real(c_double), intent(in) :: rhs(1:jx, 1:jy, 1:jz)
      real(c_double), intent(in), target :: cs1(jy, jz), cs2(jy, jz), cs3(jx, jz), cs4(jx, jz)
      real(c_double), intent(in), target :: cs5(jx, jy), cs6(jx, jy)
      integer, intent(out) :: cycles, status
      real(c_double), intent(out) :: resmax
      type(ples_mg_result) :: res
      status = ples_mg_solve(handle, fi, int(deepl, c_int), int(deepr, c_int), rhs, c_loc(cs1), c_loc(cs2), &
                             c_loc(cs3), c_loc(cs4), c_loc(cs5), c_loc(cs6), dt, eps, int(ficycle, c_int), res)
      cycles = res%cycles
      resmax = res%abs_res
   end subroutine

   type(c_ptr) function ples_multi_handle()
      ples_multi_handle = handle
   end function

   subroutine ples_multi_finalize()
      integer(c_int) :: st
      if (c_associated(handle)) st = ples_mg_destroy(handle)
      handle = c_null_ptr
   end subroutine

end module ples_mg_f
