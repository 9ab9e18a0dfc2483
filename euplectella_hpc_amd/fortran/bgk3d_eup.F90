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
!   bgk.perf                run info, time per section, ratios, derived metrics and memory
!                           (finalize.F90:57-125): "coll time" is the whole GPU step (boundaries,
!                           propagation and collision run fused on the GPU; HIP-event time), so
!                           the move, bc and Obst times are 0 and the move bandwidth is taken over
!                           the step time; memory is the peak resident set of the host process
!   fort.100, stdout        every isignal steps the mean time per step and the per-section time
!                           increments (profile.F90)
!   tec_ob.0000.00000000.vtk, obs.0000.00000000.bin
!                           obstacle dumps of flag3 = 1 (ASCII VTK) and 2 (list), build.obs.F90:450-531
! Obstacles (build_obs): flag1 = 0 none, 1 the list file above, 2 created in the code
! (build.obs.F90:201-439): flag2 = 1 a spherical shell of radius 20 around (lx/4, ly/2, lz/2),
! 2 the CYLINDER_INF column of radius 100 around (lx/4, ly/2) through the whole height, 3 the half
! column of radius min(20, lz/8) below z = 60; the reference's obstacle count starts at -100 in
! that path, which shows in the dumps' counts as it does there.
! Case: CYLINDER_INF + OBSTACLES (inflow at x = 0, outflow at x = l+1, periodic y and z;
! build.bcond.F90:319-339).  The serial build runs one task whatever proc_x, proc_y, proc_z say
! (input.F90:64-72).  The step semantics and storage come from the environment:
!   EUP_MODE    = parity (default; the reference's serial FP64 step bit for bit) | physics
!   EUP_STORAGE = f64 (default) | f32 | f16 | bf16 | e3m12   (physics mode)
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
   integer(1), allocatable :: obsd(:, :, :)   ! dense obstacle field, only for the dumps of flag3
   logical :: dense_obs                      ! EUP_OBS=dense: the obstacles come from restore.obs.bin
   ! section times as the reference keeps them (mod_tools.F90: single precision; the second column
   ! from the time-of-day clock, time() of mod_tools.F90:44-52)
   real(4) :: time_loop, time_loop1, time_coll, time_coll1, time_move, time_move1, time_bc, time_bc1
   real(4) :: time_io, time_io1, time_dg, time_dg1, time_obs, time_obs1, time_mp, time_mp1
   real(4) :: tl0, tl1, ts0, ts1, tinn, old_mp, old_coll, old_move
   real(c_double) :: gpu_ms, mem_start, mem_stop
   integer(8) :: cs0, cs1, cd0

   flag1 = 2; flag2 = 2; flag3 = 1; ipad = 0; jpad = 0; kpad = 0
   open(15, file='bgk.input', status='old')
   read(15, parameters)
   close(15)
   ! the serial build: one task (input.F90:64-72)
   proc_x = 1; proc_y = 1; proc_z = 1
   l = lx; m = ly; n = lz
   if (flag1 > 2) then
      write(6,*) "ERROR: flag1 not supported -->", flag1
      stop
   end if
   if (flag2 > 3) then
      write(6,*) "ERROR: flag2 not supported -->", flag2
      stop
   end if
   if (flag3 > 3) then
      write(6,*) "ERROR: flag3 not supported -->", flag3
      stop
   end if
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
   case ('e3m12'); cfg%storage = EUP_STORE_E3M12
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
   dense_obs = flag1 == 1 .and. trim(env) == 'dense'
   if (dense_obs) then
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
      nobs = 0   ! the dense field's solid sites (build.obs.F90:112-132)
      do k = 1, n
         read(54) plane
         do j = 1, m
            do i = 1, l
               if (int(plane(i + l * (j - 1))) == 1) then
                  cnt = cnt + 1
                  nobs = nobs + 1
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
   else if (flag1 == 2) then
      call create_obstacles()
   else
      write(6,*) "INFO: no obstacle"
      nobs = -100
   end if
   rc = eup_commit_obstacles(h)
   call eup_check(rc, h, 'eup_commit_obstacles')
   if (flag3 == 1 .or. flag3 == 2) call dump_obstacles()
   if (flag3 == 3) then
      write(6,*) "ERROR: HDF5 preproc flag not activated..."
      stop
   end if

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
   mem_start = peak_rss_mb()

   time_coll = 0; time_coll1 = 0; time_move = 0; time_move1 = 0; time_bc = 0; time_bc1 = 0
   time_io = 0; time_io1 = 0; time_dg = 0; time_dg1 = 0; time_obs = 0; time_obs1 = 0
   time_mp = 0; time_mp1 = 0; old_mp = 0; old_coll = 0; old_move = 0
   call system_clock(c0, crate)
   tl0 = clock_of_day()
   cd0 = c0
   itime = itstart
   do while (itime < itfin)
      nxt = min(itfin, (itime / icheck + 1) * icheck, (itime / itsave + 1) * itsave, &
                (itime / ivtim + 1) * ivtim, (itime / isignal + 1) * isignal)
      ! boundaries, propagation and collision run fused on the GPU: their time is the step's
      ts0 = clock_of_day()
      rc = eup_step_timed(h, nxt - itime, gpu_ms)
      call eup_check(rc, h, 'eup_step_timed')
      ts1 = clock_of_day()
      time_coll = time_coll + real(gpu_ms / 1000.d0, 4)
      time_coll1 = time_coll1 + (ts1 - ts0)
      itime = nxt
      call system_clock(cs0)
      ts0 = clock_of_day()
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
      call system_clock(cs1)
      ts1 = clock_of_day()
      time_io = time_io + real(cs1 - cs0) / real(crate)
      time_io1 = time_io1 + (ts1 - ts0)
      call system_clock(cs0)
      ts0 = clock_of_day()
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
      call system_clock(cs1)
      ts1 = clock_of_day()
      time_dg = time_dg + real(cs1 - cs0) / real(crate)
      time_dg1 = time_dg1 + (ts1 - ts0)
      call system_clock(cs0)
      ts0 = clock_of_day()
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
      call system_clock(cs1)
      ts1 = clock_of_day()
      time_io = time_io + real(cs1 - cs0) / real(crate)
      time_io1 = time_io1 + (ts1 - ts0)
      if (mod(itime, isignal) == 0) then
         ! profile.F90: mean time per step since the last report, and the section increments
         call system_clock(cs1)
         tinn = real(cs1 - cd0) / real(crate)
         write(6, 1001) tinn / isignal, itime, itfin
         write(100, 2001) itime, (time_mp - old_mp), (time_coll - old_coll), (time_move - old_move)
         old_mp = time_mp; old_coll = time_coll; old_move = time_move
         call system_clock(cd0)
      end if
   end do
   call system_clock(c1)
   tl1 = clock_of_day()
   tloop = real(c1 - c0, c_double) / real(crate, c_double)
   time_loop = real(c1 - c0) / real(crate)
   time_loop1 = tl1 - tl0
   close(66)
   close(67)
   mem_stop = peak_rss_mb()
   call write_perf()
   call eup_destroy(h)
   write(6,*) "That's all folks!!!!"

1001 format(" Mean   time", 1(e14.6, 1x), i8, "/", i8)
1002 format(i8, 4(e14.6, 1x))
1004 format(i8, 5(e14.6, 1x))
2001 format(i8, 3(e14.6, 1x))

contains

   ! bgk.perf in the reference's layout (finalize.F90:57-125, serial build with ORIGINAL and
   ! OBSTACLES): the fused GPU step is the "coll" section, so move, bc and Obst are 0 and the move
   ! bandwidth is taken over the step time
   subroutine write_perf()
      real(4) :: tmove_bw
      tmove_bw = time_move
      if (tmove_bw <= 0) tmove_bw = time_coll
      open(69, file='bgk.perf', status='unknown')
      write(69, 9999)
      write(69, *) "# Run info "
      write(69, *) lx, ly, lz
      write(69, *) l, m, n
      write(69, 9999)
      write(69, *) "# Time for section "
      write(69, 1101) time_loop, time_loop1
      write(69, 1102) time_coll, time_coll1
      write(69, 1103) time_move, time_move1
      write(69, 1104) time_bc, time_bc1
      write(69, 1105) time_io, time_io1
      write(69, 1114) time_dg, time_dg1
      write(69, 1116) time_obs, time_obs1
      write(69, 1117) time_loop - (time_coll + time_move + time_bc + time_io + time_dg + time_mp + time_obs)
      write(69, 9999)
      write(69, *) "# Ratio per section "
      write(69, 1202) time_io / time_loop, time_io1 / time_loop1
      write(69, 1203) time_bc / time_loop, time_bc1 / time_loop1
      write(69, 1204) time_coll / time_loop, time_coll1 / time_loop1
      write(69, 1205) time_dg / time_loop, time_dg1 / time_loop1
      write(69, 1206) time_move / time_loop, time_move / time_loop1
      write(69, 1207) time_obs / time_loop, time_obs / time_loop1
      write(69, 9999)
      write(69, *) "# Derived (global) metrics "
      write(69, 1106) float(lx) * float(ly) * float(lz) * (itfin - itstart) / (time_loop) / 1000.0 / 1000.0
      write(69, 1107) float(215) * float(lx) * float(ly) * float(lz) * (itfin - itstart) / (time_coll) / 1000.0 / 1000.0
      write(69, 1108) float(19 * 8) * float(lx) * float(ly) * float(lz) * (itfin - itstart) / (time_coll) / 1000.0 / 1000.0
      write(69, 1109) float(18 * 8) * float(lx) * float(ly) * float(lz) * (itfin - itstart) / (tmove_bw) / 1000.0 / 1000.0
      write(69, 9999)
      write(69, *) " # Memory (task 0) metrics "
      write(69, 1110) mem_start, mem_stop
      write(69, *) (l / 1024.0) * (m / 1024.0) * (n / 1024.0) * 19 * 2 * 4
      write(69, 9999)
      close(69)
9999  format(" #--------------------------------")
1101  format(" # loop   time", 2(e14.6, 1x))
1102  format(" # coll   time", 2(e14.6, 1x))
1103  format(" # move   time", 2(e14.6, 1x))
1104  format(" # bc     time", 2(e14.6, 1x))
1105  format(" # I/O    time", 2(e14.6, 1x))
1114  format(" # diagno time", 2(e14.6, 1x))
1116  format(" # Obst   time", 2(e14.6, 1x))
1117  format(" # Check      ", 1(e14.6, 1x))
1202  format(" # Ratio I/O  ", 2(f7.3, 1x))
1203  format(" # Ratio BC   ", 2(f7.3, 1x))
1204  format(" # Ratio Coll ", 2(f7.3, 1x))
1205  format(" # Ratio Diag.", 2(f7.3, 1x))
1206  format(" # Ratio Move ", 2(f7.3, 1x))
1207  format(" # Ratio Obs  ", 2(f7.3, 1x))
1106  format(" # Mlups      ", 1(f14.6, 1x))
1107  format(" # coll   Flop", 1(e14.6, 1x), "MFlops")
1108  format(" # coll   BW  ", 1(e14.6, 1x), "GB/s")
1109  format(" # move   BW  ", 1(e14.6, 1x), "GB/s")
1110  format(" # Memory (start,stop)", 2(f14.6, 1x), "MB")
   end subroutine write_perf

   ! seconds of the day from the date clock, as the reference's time() (mod_tools.F90:44-52)
   real(4) function clock_of_day()
      integer :: v(8)
      call date_and_time(values=v)
      clock_of_day = 3600. * v(5) + 60. * v(6) + v(7) + v(8) / 1000.
   end function clock_of_day

   ! peak resident set of this process in MB (what the reference's get_mem reports, from the
   ! kernel's VmHWM line)
   real(c_double) function peak_rss_mb()
      character(len=256) :: line
      integer :: u, ios, kb
      peak_rss_mb = 0
      open(newunit=u, file='/proc/self/status', status='old', action='read', iostat=ios)
      if (ios /= 0) return
      do
         read(u, '(A)', iostat=ios) line
         if (ios /= 0) exit
         if (line(1:6) == 'VmHWM:') then
            read(line(7:), *, iostat=ios) kb
            if (ios == 0) peak_rss_mb = kb / 1024.d0
            exit
         end if
      end do
      close(u)
   end function peak_rss_mb

   ! build_obs with flag1 = 2 (build.obs.F90:201-439): the obstacle created in the code, streamed
   ! to the device in chunks; nobs counts from -100 as the reference's does
   subroutine create_obstacles()
      integer :: ii, jj, kk, cn
      integer(c_int32_t), allocatable :: ch(:, :)
      write(6,*) "INFO: creating obstacle"
      nobs = -100
      allocate(ch(3, 1048576))
      cn = 0
      do kk = 1, n
         do jj = 1, m
            do ii = 1, l
               if (created_solid(ii, jj, kk)) then
                  nobs = nobs + 1
                  cn = cn + 1
                  ch(1, cn) = ii; ch(2, cn) = jj; ch(3, cn) = kk
                  if (cn == size(ch, 2)) then
                     rc = eup_add_obstacles(h, ch, int(cn, c_int64_t))
                     call eup_check(rc, h, 'eup_add_obstacles')
                     cn = 0
                  end if
               end if
            end do
         end do
      end do
      if (cn > 0) then
         rc = eup_add_obstacles(h, ch, int(cn, c_int64_t))
         call eup_check(rc, h, 'eup_add_obstacles')
      end if
      deallocate(ch)
      write(6,*) "INFO: num. obs            -->", nobs
   end subroutine create_obstacles

   ! whether site (ii, jj, kk) is solid in the obstacle of flag2 (serial: no task offsets).  The
   ! squared distance is an integer, compared in double with the real radii.
   logical function created_solid(ii, jj, kk)
      integer, intent(in) :: ii, jj, kk
      integer :: ic, jc, kc, rad, sz
      real(c_double) :: d2, r, r2a, r2b
      created_solid = .false.
      select case (flag2)
      case (1)   ! spherical shell (d2 strictly between (R-2)^2 and (R+2)^2)
         ic = lx / 4; jc = ly / 2; kc = lz / 2; rad = 20
         r = rad * 1.0d0
         r2a = (r - 2) * (r - 2); r2b = (r + 2) * (r + 2)
         d2 = (ic - ii) * (ic - ii) + (jc - jj) * (jc - jj) + (kc - kk) * (kc - kk)
         created_solid = d2 > r2a .and. d2 < r2b
      case (2)   ! CYLINDER_INF column: radius 100 through the whole height
         ic = lx / 4; jc = ly / 2; rad = 100
         r = rad * 1.0d0
         r2b = (r + 2) * (r + 2)
         d2 = (ic - ii) * (ic - ii) + (jc - jj) * (jc - jj)
         created_solid = d2 >= 0 .and. d2 < r2b
      case (3)   ! half column: radius min(20, lz/8), z from 1 below 60
         ic = lx / 4; jc = ly / 2; rad = min(20, lz / 8); sz = 60
         r = rad * 1.0d0
         r2b = (r + 2) * (r + 2)
         d2 = (ic - ii) * (ic - ii) + (jc - jj) * (jc - jj)
         created_solid = kk >= 0 .and. kk < sz .and. d2 >= 0 .and. d2 < r2b
      end select
   end function created_solid

   ! whether site (ii, jj, kk) is solid, for the dumps: the created obstacle, or the dense field
   ! built from the obstacle file
   logical function site_solid(ii, jj, kk)
      integer, intent(in) :: ii, jj, kk
      if (flag1 == 2) then
         site_solid = created_solid(ii, jj, kk)
      else if (allocated(obsd)) then
         site_solid = obsd(ii, jj, kk) == 1
      else
         site_solid = .false.
      end if
   end function site_solid

   ! obstacle dumps of flag3 (build.obs.F90:450-531): 1 = ASCII VTK of the whole field when the
   ! count is positive, 2 = the list in the obstacle-file format, the count record first
   subroutine dump_obstacles()
      integer :: ii, jj, kk, check, idx2, rr
      character(len=24) :: fvtk
      character(len=21) :: fbin
      if (flag1 == 1) then
         ! the dense field of the obstacles just read, for the site loops below
         allocate(obsd(l, m, n))
         obsd = 0
         if (dense_obs) then
            allocate(plane(l * m))
            open(54, file='restore.obs.bin', form='unformatted', status='old')
            read(54) itobs
            read(54) dims
            do kk = 1, n
               read(54) plane
               do jj = 1, m
                  do ii = 1, l
                     if (int(plane(ii + l * (jj - 1))) == 1) obsd(ii, jj, kk) = 1
                  end do
               end do
            end do
            close(54)
            deallocate(plane)
         else
            open(53, file='obs.0000.00000000.bin', form='unformatted', status='old')
            read(53) idx2
            do rr = 1, idx2
               read(53) ii, jj, kk
               obsd(ii, jj, kk) = 1
            end do
            close(53)
         end if
      end if
      if (flag3 == 1) then
         if (nobs <= 0) then
            write(6,*) "task", 0, "has no obstacles"
         else
            write(6,*) "task", 0, "has ", nobs, " obstacles"
            fvtk = 'tec_ob.0000.00000000.vtk'
            open(52, file=fvtk, status='unknown')
            write(52, '(A26)') '# vtk DataFile Version 2.0'
            write(52, '(A5)') 'Campo'
            write(52, '(A5)') 'ASCII'
            write(52, '(A24)') 'DATASET RECTILINEAR_GRID'
            write(52, '(A11,I10,A1,I10,A1,I10)') 'DIMENSIONS ', l, ' ', m, ' ', n
            write(52, '(A14,I10,A7)') 'X_COORDINATES ', l, ' double'
            do ii = 1, l
               write(52, *) ii
            end do
            write(52, '(A14,I10,A7)') 'Y_COORDINATES ', m, ' double'
            do jj = 1, m
               write(52, *) jj
            end do
            write(52, '(A14,I10,A7)') 'Z_COORDINATES ', n, ' double'
            do kk = 1, n
               write(52, *) kk
            end do
            write(52, '(A10,I10)') 'POINT_DATA ', l * m * n
            write(52, '(A21)') 'SCALARS obs double'
            write(52, '(A20)') 'LOOKUP_TABLE default'
            do kk = 1, n
               do jj = 1, m
                  do ii = 1, l
                     if (site_solid(ii, jj, kk)) then
                        write(52, *) 1 * 1.0
                     else
                        write(52, *) 0 * 1.0
                     end if
                  end do
               end do
            end do
            write(6,*) "INFO: obs vtk dump done"
            close(52)
         end if
      end if
      if (flag3 == 2) then
         fbin = 'obs.0000.00000000.bin'
         write(6,*) "INFO: obstacle binary dump ", fbin
         open(53, file=fbin, form='unformatted', status='unknown')
         check = 0
         write(53) nobs
         do kk = 1, n
            do jj = 1, m
               do ii = 1, l
                  if (site_solid(ii, jj, kk)) then
                     write(53) ii, jj, kk
                     check = check + 1
                  end if
               end do
            end do
         end do
         if (check /= nobs) then
            write(6,*) "ERROR: obstacles not correct->", check, nobs
         else
            write(6,*) "INFO: obs binary dump done"
         end if
         close(53)
      end if
      if (allocated(obsd)) deallocate(obsd)
   end subroutine dump_obstacles

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
