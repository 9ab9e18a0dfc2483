! ISO_C_BINDING interface to libeup.so (include/eup.h) for the sponge-flow Fortran host.
!
! The reference calls its time step as argument-less subroutines over module state
! (CODE/SRC/bgk3d.F90:128-138 with `use storage`); with this module the same host keeps its deck,
! obstacle and restore formats and calls the GPU library through a handle instead.  Errors come
! back as integer codes; eup_check maps them to the reference's "write(6,*) ERROR; stop" style.
module eup
   use iso_c_binding
   implicit none

   integer(c_int32_t), parameter :: EUP_STORE_F64 = 0, EUP_STORE_F32 = 1
   integer(c_int32_t), parameter :: EUP_STORE_F16 = 2, EUP_STORE_BF16 = 3, EUP_STORE_E3M12 = 4
   integer(c_int32_t), parameter :: EUP_MODE_PARITY = 0, EUP_MODE_PHYSICS = 1, EUP_MODE_GREY = 2
   integer(c_int32_t), parameter :: EUP_FACE_PERIODIC = 0, EUP_FACE_INFLOW = 4
   integer(c_int32_t), parameter :: EUP_FACE_OUTFLOW = 5
   integer(c_int32_t), parameter :: EUP_FLAG_NO_TILE_SKIP = 1, EUP_FLAG_SHIFTED_DEVIATION = 2
   integer(c_int32_t), parameter :: EUP_VTK_YZ = 0, EUP_VTK_XZ = 1, EUP_VTK_XY = 2, EUP_VTK_3D = 3

   ! mirrors struct eup_config (flags is uint32_t in C; same size and alignment)
   type, bind(c) :: eup_config
      integer(c_int32_t) :: l, m, n
      integer(c_int32_t) :: mode, storage
      integer(c_int32_t) :: front, rear, up, down, left, right
      integer(c_int32_t) :: device
      integer(c_int32_t) :: flags
      real(c_double) :: omega, fgrad, u_inflow
   end type eup_config

   ! mirrors struct eup_info
   type, bind(c) :: eup_info
      integer(c_int64_t) :: sites, nsolid
      integer(c_int64_t) :: tiles, tiles_fluid, tiles_mixed, tiles_solid
      integer(c_int32_t) :: bbox(6)
      integer(c_int32_t) :: bytes_per_site
      integer(c_int64_t) :: device_bytes
   end type eup_info

   interface
      integer(c_int) function eup_create(cfg, h) bind(c, name='eup_create')
         import :: c_int, c_ptr, eup_config
         type(eup_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: h
      end function eup_create

      subroutine eup_destroy(h) bind(c, name='eup_destroy')
         import :: c_ptr
         type(c_ptr), value :: h
      end subroutine eup_destroy

      type(c_ptr) function eup_last_error(h) bind(c, name='eup_last_error')
         import :: c_ptr
         type(c_ptr), value :: h
      end function eup_last_error

      integer(c_int) function eup_set_mask(h, obs) bind(c, name='eup_set_mask')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), intent(in) :: obs(*)
      end function eup_set_mask

      integer(c_int) function eup_add_obstacles(h, ijk, count) bind(c, name='eup_add_obstacles')
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value :: h
         integer(c_int32_t), intent(in) :: ijk(*)
         integer(c_int64_t), value :: count
      end function eup_add_obstacles

      integer(c_int) function eup_commit_obstacles(h) bind(c, name='eup_commit_obstacles')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function eup_commit_obstacles

      integer(c_int) function eup_set_pops(h, q, a) bind(c, name='eup_set_pops')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: q
         real(c_double), intent(in) :: a(*)
      end function eup_set_pops

      integer(c_int) function eup_get_pops(h, q, a) bind(c, name='eup_get_pops')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: q
         real(c_double), intent(out) :: a(*)
      end function eup_get_pops

      integer(c_int) function eup_step(h, nsteps) bind(c, name='eup_step')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h
         integer(c_int32_t), value :: nsteps
      end function eup_step

      integer(c_int) function eup_step_timed(h, nsteps, ms) bind(c, name='eup_step_timed')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: nsteps
         real(c_double), intent(out) :: ms
      end function eup_step_timed

      integer(c_int) function eup_drag_cv(h, box, fx, fy) bind(c, name='eup_drag_cv')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), intent(in) :: box(6)
         real(c_double), intent(out) :: fx, fy
      end function eup_drag_cv

      integer(c_int) function eup_drag_cv_noalloc(h, box, fx, fy) bind(c, name='eup_drag_cv_noalloc')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), intent(in) :: box(6)
         real(c_double), intent(out) :: fx, fy
      end function eup_drag_cv_noalloc

      integer(c_int) function eup_drag_mem(h, f) bind(c, name='eup_drag_mem')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: f(3)
      end function eup_drag_mem

      integer(c_int) function eup_probe(h, i, j, k, out) bind(c, name='eup_probe')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: i, j, k
         real(c_double), intent(out) :: out(4)
      end function eup_probe

      integer(c_int) function eup_diagno(h, out) bind(c, name='eup_diagno')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: out(5)
      end function eup_diagno

      integer(c_int) function eup_vtk_velocity(h, plane, index, out) bind(c, name='eup_vtk_velocity')
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: plane, index
         real(c_float), intent(out) :: out(*)
      end function eup_vtk_velocity

      integer(c_int) function eup_vtk_vorticity(h, k0, out) bind(c, name='eup_vtk_vorticity')
         import :: c_int, c_ptr, c_int32_t, c_float
         type(c_ptr), value :: h
         integer(c_int32_t), value :: k0
         real(c_float), intent(out) :: out(*)
      end function eup_vtk_vorticity

      integer(c_int32_t) function eup_abi_version() bind(c, name='eup_abi_version')
         import :: c_int32_t
      end function eup_abi_version

      integer(c_int) function eup_set_state(h, a) bind(c, name='eup_set_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: a(*)
      end function eup_set_state

      integer(c_int) function eup_get_state(h, a) bind(c, name='eup_get_state')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: a(*)
      end function eup_get_state

      integer(c_int) function eup_init_equilibrium(h, ux, uy, uz, noise, seed) &
            bind(c, name='eup_init_equilibrium')
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value :: h
         real(c_double), value :: ux, uy, uz, noise
         integer(c_int64_t), value :: seed
      end function eup_init_equilibrium

      integer(c_int) function eup_set_omega(h, omega) bind(c, name='eup_set_omega')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), value :: omega
      end function eup_set_omega

      integer(c_int) function eup_sync(h) bind(c, name='eup_sync')
         import :: c_int, c_ptr
         type(c_ptr), value :: h
      end function eup_sync

      integer(c_int) function eup_get_info(h, info) bind(c, name='eup_get_info')
         import :: c_int, c_ptr, eup_info
         type(c_ptr), value :: h
         type(eup_info), intent(out) :: info
      end function eup_get_info

      type(c_ptr) function eup_stream(h) bind(c, name='eup_stream')
         import :: c_ptr
         type(c_ptr), value :: h
      end function eup_stream

      integer(c_int) function eup_membw(device, bytes, reps, out) bind(c, name='eup_membw')
         import :: c_int, c_int32_t, c_int64_t, c_double
         integer(c_int32_t), value :: device
         integer(c_int64_t), value :: bytes
         integer(c_int32_t), value :: reps
         real(c_double), intent(out) :: out(3)
      end function eup_membw

      ! grey lattice and adjoint (EUP_MODE_GREY)
      integer(c_int) function eup_set_grey(h, s) bind(c, name='eup_set_grey')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(in) :: s(*)
      end function eup_set_grey

      integer(c_int) function eup_set_grey_device(h, s, dtype) bind(c, name='eup_set_grey_device')
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value :: h, s
         integer(c_int32_t), value :: dtype
      end function eup_set_grey_device

      integer(c_int) function eup_grey_force(h, f) bind(c, name='eup_grey_force')
         import :: c_int, c_ptr, c_double
         type(c_ptr), value :: h
         real(c_double), intent(out) :: f(3)
      end function eup_grey_force

      ! grad may be c_null_ptr (objective only) or c_loc of an l*m*n real(c_double) array
      integer(c_int) function eup_grey_objective(h, t0, t1, every, norm, j, grad) &
            bind(c, name='eup_grey_objective')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: t0, t1, every
         real(c_double), value :: norm
         real(c_double), intent(out) :: j
         type(c_ptr), value :: grad
      end function eup_grey_objective

      integer(c_int) function eup_grey_objective_device(h, t0, t1, every, norm, j, grad, dtype) &
            bind(c, name='eup_grey_objective_device')
         import :: c_int, c_ptr, c_int32_t, c_double
         type(c_ptr), value :: h
         integer(c_int32_t), value :: t0, t1, every
         real(c_double), value :: norm
         real(c_double), intent(out) :: j
         type(c_ptr), value :: grad
         integer(c_int32_t), value :: dtype
      end function eup_grey_objective_device
   end interface

contains

   ! Stop with the library's message when rc /= 0 (the reference reports errors with
   ! write(6,*) and stop, e.g. CODE/SRC/input.F90:88-101).
   subroutine eup_check(rc, h, what)
      integer(c_int), intent(in) :: rc
      type(c_ptr), intent(in) :: h
      character(len=*), intent(in) :: what
      character(kind=c_char), pointer :: msg(:)
      type(c_ptr) :: p
      integer :: i
      character(len=512) :: text
      if (rc == 0) return
      p = eup_last_error(h)
      text = ' '
      if (c_associated(p)) then
         call c_f_pointer(p, msg, [512])
         do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
         end do
      end if
      write(6,*) "ERROR: ", what, " failed (", rc, "): ", trim(text)
      stop 1
   end subroutine eup_check

end module eup
