! Fortran host for the GPU time step: the reference's serial run flow with libeup underneath.
!
! Keeps the reference's files and formats unchanged:
!   bgk.input               namelist &parameters (CODE/SRC/input.F90:46-62)
!   obs.0000.00000000.bin   obstacle list, flag1 = 1 (build.obs.F90:147-186)
!   restore.obs.bin         with EUP_OBS=dense and flag1 = 1: the dense obstacle field in place of
!                           the list, the HDF5-free form of the -DHDF5 build's restore.obs.h5
!                           (restore.obs.hdf5.serial.F90): records itime, (l, m, n), then one
!                           record of l*m FP64 values per z plane; truncated to integer, 1 = solid
!   restore.00.00.00.bin    raw restart, irestart = 1 (restore_raw.F90:43-63)
!   save.00.00.00.T.bin     raw dump every itsave steps (save_raw.F90:54-74)
!   drag.0000.dat           control-volume drag every icheck steps (drag.F90:48-67), written when
!                           EUP_DRAG=1 (the reference's -DDRAG build)
!   probe_g.0000.dat        velocity probe every icheck steps (probe_global.F90:77,93;
!                           site from diagnostic.F90:145)
!   diagno.dat              global means at itime 0 for a fresh start (diagno.F90:157)
!   tec_*.0000.T.vtk        binary VTK files every ivtim steps (diagnostic.F90:75-90): the default
!                           branch (xy plane n/2-1, xz plane m/2-1, yz plane l/2-1, vorticity of
!                           the xy plane n/2+1), or with EUP_VTK=hpc the -DHPC branch (3-D volume,
!                           xy plane n/2+1); EUP_VTK=none writes none
!   bgk.perf                MLUPS line (finalize.F90:107-108,222)
! Case: CYLINDER_INF + OBSTACLES (inflow at x = 0, outflow at x = l+1, periodic y and z;
! build.bcond.F90:319-339).  The step semantics and storage come from the environment:
!   EUP_MODE    = parity (default; the reference's serial FP64 step bit for bit) | physics
!   EUP_STORAGE = f64 (default) | f32 | f16 | bf16   (physics mode)
!   EUP_DEVICE  = HIP device ordinal (default 0)
program bgk3d_eup
   use iso_c_binding
   use eup
   implicit none

   real(c_double) :: svisc, u0, u_inflow, omega, fgrad
   integer :: itfin, ivtim, isignal, itsave, icheck, irestart, init_v
   integer :: lx, ly, lz, proc_x, proc_y, proc_z, tstep, flag1, flag2, flag3, ipad, jpad, kpad
   namelist /parameters/ svisc, u0, itfin, ivtim, isignal, itsave, icheck, irestart, init_v, &
      lx, ly, lz, proc_x, proc_y, proc_z, tstep, flag1, flag2, flag3, ipad, jpad, kpad

   type(eup_config) :: cfg
   type(c_ptr) :: h
   integer(c_int32_t), allocatable :: chunk(:, :)
   real(c_double), allocatable :: buf(:)
   real(c_double) :: pr(4), dg(5), fx, fy, tloop
   integer(c_int32_t) :: box(6)
   integer :: l, m, n, q, itime, itstart, nobs, idx, i, j, k, nxt, delta, cnt, r
   integer(8) :: c0, c1, crate
   integer :: do_drag, vtk_branch
   integer(c_int) :: rc
   character(len=32) :: env
   character(len=27) :: fname
   integer :: dims(3), itobs
   real(c_double), allocatable :: plane(:)

   flag1 = 2; flag2 = 2; flag3 = 1; ipad = 0; jpad = 0; kpad = 0
   open(15, file='bgk.input', status='old')
   read(15, parameters)
   close(15)
   if (proc_x * proc_y * proc_z /= 1) then
      write(6,*) "ERROR: this host runs one task (proc_x = proc_y = proc_z = 1)"
      stop 1
   end if
   l = lx; m = ly; n = lz
   ! build_bcond for CYLINDER_INF: the deck's u0 becomes the inflow velocity, no volume force
   u_inflow = u0
   u0 = 0
   ! hencol
   omega = 2.d0 / (6.d0 * svisc + 1.d0)
   fgrad = 4.0 * u0 * svisc / (5.0 * real(n, c_double) * real(n, c_double))

   cfg%l = l; cfg%m = m; cfg%n = n
   cfg%mode = EUP_MODE_PARITY
   cfg%storage = EUP_STORE_F64
   call get_environment_variable('EUP_MODE', env)
   if (trim(env) == 'physics') cfg%mode = EUP_MODE_PHYSICS
   call get_environment_variable('EUP_STORAGE', env)
   select case (trim(env))
   case ('f32'); cfg%storage = EUP_STORE_F32
   case ('f16'); cfg%storage = EUP_STORE_F16
   case ('bf16'); cfg%storage = EUP_STORE_BF16
   end select
   cfg%device = 0
   call get_environment_variable('EUP_DEVICE', env)
   if (len_trim(env) > 0) read(env, *) cfg%device
   do_drag = 0
   call get_environment_variable('EUP_DRAG', env)
   if (trim(env) == '1') do_drag = 1
   vtk_branch = 1
   call get_environment_variable('EUP_VTK', env)
   if (trim(env) == 'hpc') vtk_branch = 2
   if (trim(env) == 'none') vtk_branch = 0
   cfg%front = EUP_FACE_OUTFLOW
   cfg%rear = EUP_FACE_INFLOW
   cfg%up = EUP_FACE_PERIODIC; cfg%down = EUP_FACE_PERIODIC
   cfg%left = EUP_FACE_PERIODIC; cfg%right = EUP_FACE_PERIODIC
   cfg%flags = 0
   cfg%omega = omega
   cfg%fgrad = fgrad
   cfg%u_inflow = u_inflow
   rc = eup_create(cfg, h)
   call eup_check(rc, c_null_ptr, 'eup_create')

   ! obstacles (build_obs, flag1 = 1 reads the task's list; 0 = no obstacle): the list is streamed
   ! to the device in chunks of at most 2^20 records, no dense obs array on the host
   call get_environment_variable('EUP_OBS', env)
   if (flag1 == 1 .and. trim(env) == 'dense') then
      ! the dense field, one z plane at a time; its solid sites go through the same list path
      open(54, file='restore.obs.bin', form='unformatted', status='old')
      read(54) itobs
      read(54) dims
      if (dims(1) /= l .or. dims(2) /= m .or. dims(3) /= n) then
         write(6,*) "ERROR: size doesn't match", l, m, n, dims
         stop 1
      end if
      allocate(plane(l * m), chunk(3, 1048576))
      cnt = 0
      do k = 1, n
         read(54) plane
         do j = 1, m
            do i = 1, l
               if (int(plane(i + l * (j - 1))) == 1) then
                  cnt = cnt + 1
                  chunk(1, cnt) = i; chunk(2, cnt) = j; chunk(3, cnt) = k
                  if (cnt == size(chunk, 2)) then
                     rc = eup_add_obstacles(h, chunk, int(cnt, c_int64_t))
                     call eup_check(rc, h, 'eup_add_obstacles')
                     cnt = 0
                  end if
               end if
            end do
         end do
      end do
      if (cnt > 0) then
         rc = eup_add_obstacles(h, chunk, int(cnt, c_int64_t))
         call eup_check(rc, h, 'eup_add_obstacles')
      end if
      close(54)
      deallocate(plane, chunk)
   else if (flag1 == 1) then
      open(53, file='obs.0000.00000000.bin', form='unformatted', status='old')
      read(53) nobs
      allocate(chunk(3, max(1, min(nobs, 1048576))))
      idx = 0
      do while (idx < nobs)
         cnt = min(size(chunk, 2), nobs - idx)
         do r = 1, cnt
            read(53) chunk(1, r), chunk(2, r), chunk(3, r)
         end do
         rc = eup_add_obstacles(h, chunk, int(cnt, c_int64_t))
         call eup_check(rc, h, 'eup_add_obstacles')
         idx = idx + cnt
      end do
      close(53)
      deallocate(chunk)
   else if (flag1 /= 0) then
      write(6,*) "ERROR: flag1 must be 0 or 1 (obstacle list file)", flag1
      stop 1
   end if
   rc = eup_commit_obstacles(h)
   call eup_check(rc, h, 'eup_commit_obstacles')

   ! populations: restore_raw or the rest state of init opt 0
   allocate(buf((l + 2) * (m + 2) * (n + 2)))
   if (irestart == 1) then
      open(30, file='restore.00.00.00.bin', form='unformatted', status='old')
      read(30) itstart
      do q = 1, 19
         read(30) buf
         rc = eup_set_pops(h, q, buf)
         call eup_check(rc, h, 'eup_set_pops')
      end do
      close(30)
   else
      if (init_v /= 0) then
         write(6,*) "ERROR: this host starts from rest (init_v = 0) or from a restore file"
         stop 1
      end if
      itstart = 0
      do q = 1, 19
         buf = weight(q)
         rc = eup_set_pops(h, q, buf)
         call eup_check(rc, h, 'eup_set_pops')
      end do
      rc = eup_diagno(h, dg)
      call eup_check(rc, h, 'eup_diagno')
      open(63, file='diagno.dat', status='unknown')
      write(63, 1004) 0, dg(2), dg(3), dg(4), dg(1), dg(5)
      close(63)
   end if

   open(66, file='drag.0000.dat', status='unknown')
   open(67, file='probe_g.0000.dat', status='unknown')

   call system_clock(c0, crate)
   itime = itstart
   do while (itime < itfin)
      nxt = min(itfin, (itime / icheck + 1) * icheck, (itime / itsave + 1) * itsave, &
                (itime / ivtim + 1) * ivtim)
      rc = eup_step(h, nxt - itime)
      call eup_check(rc, h, 'eup_step')
      itime = nxt
      if (mod(itime, ivtim) == 0) then
         if (vtk_branch == 1) then
            call vtk_velocity('xy', EUP_VTK_XY, n / 2 - 1)
            call vtk_velocity('xz', EUP_VTK_XZ, m / 2 - 1)
            call vtk_velocity('yz', EUP_VTK_YZ, l / 2 - 1)
            call vtk_vorticity(n / 2 + 1)
         else if (vtk_branch == 2) then
            call vtk_velocity('3d', EUP_VTK_3D, 0)
            call vtk_velocity('xy', EUP_VTK_XY, n / 2 + 1)
         end if
      end if
      if (mod(itime, icheck) == 0) then
         if (do_drag == 1) then
            do delta = 52, 67, 5
               box = [lx / 4 - 100 - delta, lx / 4 + 100 + delta, ly / 2 - 100 - delta, &
                      ly / 2 + 100 + delta, 1, lz]
               rc = eup_drag_cv(h, box, fx, fy)
               call eup_check(rc, h, 'eup_drag_cv')
               write(66, *) delta, itime, fx, fy, 0
            end do
         end if
         i = 8 * lx / 20; j = 9 * ly / 16; k = lz / 2 - 5
         if (i >= 1 .and. i <= l .and. j >= 1 .and. j <= m .and. k >= 1 .and. k <= n) then
            rc = eup_probe(h, i, j, k, pr)
            call eup_check(rc, h, 'eup_probe')
            write(67, 1002) itime, pr(1), pr(2), pr(3), pr(4)
         end if
      end if
      if (mod(itime, itsave) == 0) then
         fname = 'save.00.00.00.xxxxxxxx.bin'
         write(fname(15:22), '(i8.8)') itime
         open(20, file=fname, form='unformatted', status='unknown')
         write(20) itime
         do q = 1, 19
            rc = eup_get_pops(h, q, buf)
            call eup_check(rc, h, 'eup_get_pops')
            write(20) buf
         end do
         close(20)
      end if
   end do
   call system_clock(c1)
   tloop = real(c1 - c0, c_double) / real(crate, c_double)
   close(66)
   close(67)

   open(69, file='bgk.perf', status='unknown')
   write(69, *) "# Run info (libeup)"
   write(69, *) lx, ly, lz
   write(69, *) "# Derived (global) metrics "
   if (tloop > 0) then
      write(69, 1106) float(lx) * float(ly) * float(lz) * (itfin - itstart) / tloop / 1000.0 / 1000.0
   end if
   close(69)
   call eup_destroy(h)
   write(6,*) "That's all folks!!!!"

1002 format(i8, 4(e14.6, 1x))
1004 format(i8, 5(e14.6, 1x))
1106 format(" # Mlups      ", 1(f14.6, 1x))

contains

   ! ASCII header of the binary VTK files, in the writers' edit descriptors (vtk_xy.binary.F90:51-64);
   ! the file is then reopened as a big-endian stream for the data
   subroutine vtk_header(fn, dims, origin, scalar)
      character(len=*), intent(in) :: fn
      integer, intent(in) :: dims(3), origin(3)
      logical, intent(in) :: scalar
      open(52, file=fn, status='unknown')
      write(52, '(A26)') '# vtk DataFile Version 2.0'
      write(52, '(A5)') 'Campo'
      write(52, '(A6)') 'BINARY'
      write(52, '(A25)') 'DATASET STRUCTURED_POINTS'
      write(52, '(A11,I10,A1,I10,A1,I10)') 'DIMENSIONS ', dims(1), ' ', dims(2), ' ', dims(3)
      write(52, '(A7,I10,A1,I10,A1,I10)') 'ORIGIN ', origin(1), ' ', origin(2), ' ', origin(3)
      write(52, '(A8,I10,A1,I10,A1,I10)') 'SPACING ', 1, ' ', 1, ' ', 1
      write(52, '(A10,I10)') 'POINT_DATA', dims(1) * dims(2) * dims(3)
      if (scalar) then
         write(52, '(A24)') 'SCALARS vorticity float'
         write(52, '(A20)') 'LOOKUP_TABLE default'
      else
         write(52, '(A23)') 'VECTORS velocity float'
      end if
      close(52)
      open(52, file=fn, status='old', position='append', form='unformatted', access='STREAM', &
           convert='BIG_ENDIAN')
   end subroutine vtk_header

   ! velocity file of one plane (index = plane position) or of the volume (vtk_*_bin)
   subroutine vtk_velocity(tag, plane, index)
      character(len=2), intent(in) :: tag
      integer(c_int32_t), intent(in) :: plane
      integer, intent(in) :: index
      character(len=24) :: fn
      integer :: dims(3), origin(3)
      real(c_float), allocatable :: vel(:)
      select case (plane)
      case (EUP_VTK_YZ); dims = [1, m, n]; origin = [1 + index, 1, 1]
      case (EUP_VTK_XZ); dims = [l, 1, n]; origin = [1, 1 + index, 1]
      case (EUP_VTK_XY); dims = [l, m, 1]; origin = [1, 1, 1 + index]
      case default; dims = [l, m, n]; origin = [1, 1, 1]
      end select
      allocate(vel(3 * dims(1) * dims(2) * dims(3)))
      rc = eup_vtk_velocity(h, plane, int(index, c_int32_t), vel)
      call eup_check(rc, h, 'eup_vtk_velocity')
      fn = 'tec_' // tag // '.0000.xxxxxxxx.vtk'
      write(fn(13:20), '(i8.8)') itime
      call vtk_header(fn, dims, origin, .false.)
      write(52) vel
      close(52)
      deallocate(vel)
   end subroutine vtk_velocity

   ! vorticity file of the xy plane k0 (vtk_om_bin): (l-2) x (m-2) points, origin (2, 2, 1)
   subroutine vtk_vorticity(k0)
      integer, intent(in) :: k0
      character(len=24) :: fn
      real(c_float), allocatable :: vor(:)
      allocate(vor((l - 2) * (m - 2)))
      rc = eup_vtk_vorticity(h, int(k0, c_int32_t), vor)
      call eup_check(rc, h, 'eup_vtk_vorticity')
      fn = 'tec_om.0000.xxxxxxxx.vtk'
      write(fn(13:20), '(i8.8)') itime
      call vtk_header(fn, [l - 2, m - 2, 1], [2, 2, 1], .true.)
      write(52) vor
      close(52)
      deallocate(vor)
   end subroutine vtk_vorticity

   ! D3Q19 weight of population q (storage.F90:77-81; rest state of init.F90 opt 0)
   real(c_double) function weight(q)
      integer, intent(in) :: q
      select case (q)
      case (19); weight = 1.d0 / 3.d0
      case (5, 6, 8, 14, 15, 17); weight = 1.d0 / 18.d0
      case default; weight = 1.d0 / 36.d0
      end select
   end function weight

end program bgk3d_eup
