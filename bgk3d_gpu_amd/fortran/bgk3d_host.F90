!=====================================================================
! bgk3d_host -- a Fortran host that drives the gfx950 library through the
! reference-compatible C-ABI the way the reference's program does
! (SRC/bgk3d.F90:75-246): read the `&parameters` namelist of a bgk.input
! file, allocate the 19 + 18 population arrays (here with hipMalloc through
! bgk_dev_malloc instead of OpenMP `target data` maps), build f = f_eq on
! the device (bgk_init_gpu), then per step call the boundary routine and
! col_mc_gpu and swap the a/b pointers (SRC/col_MC.F90:539-596); every icheck steps write the
! diagno.dat line (SRC/diagno.F90:145) from the device reduction.
!
! Usage: bgk3d_host.x BGK_INPUT CASE OUTDIR
!   CASE: ldc | periodic | channel   (the reference's LDC / TGV3D / POF builds)
! Writes OUTDIR/diagno.dat, OUTDIR/out.bin (a01..a19 incl. halo, raw) and
! a "# Mlups" line on stdout (SRC/finalize.F90:106-107).
!=====================================================================
      program bgk3d_host
      use iso_c_binding
      use bgk3d_gpu_mod
      implicit none
      character(len=512) :: inp, outdir
      character(len=16)  :: case
      real(c_real) :: svisc, u0, cteS
      integer :: itfin, ivtim, isignal, itsave, icheck, irestart, init_v
      integer :: lx, ly, lz, flag1, flag2, flag3, ipad, jpad, radius
      namelist /parameters/ svisc, u0, itfin, ivtim, isignal, itsave, icheck, &
                            irestart, init_v, lx, ly, lz, flag1, flag2, flag3, &
                            ipad, jpad, radius, cteS
      integer :: nx, ny, nz, d, itime, rc, mode
      integer(c_size_t) :: count, bytes
      real(c_real), allocatable, target :: h(:,:,:,:)
      type(c_ptr) :: pa(19), pb(19), ph(19), tmp
      real(c_real) :: omega, fgrad, force, u00, uf
      real(c_real), parameter :: uno = 1, zero = 0
      real(c_real), parameter :: p0 = real(1.d0/3.d0, c_real)
      real(c_real), parameter :: p1 = real(1.d0/18.d0, c_real)
      real(c_real), parameter :: p2 = real(1.d0/36.d0, c_real)
      real(c_double) :: dg(5)
      integer(8) :: c0, c1, crate
      real(8) :: tloop

      call get_command_argument(1, inp)
      call get_command_argument(2, case)
      call get_command_argument(3, outdir)
      ipad = 0; jpad = 0; init_v = 0; icheck = 1000000; cteS = 0.1
      open(15, file=trim(inp), status='old')
      read(15, parameters)
      close(15)

      nx = lx + 2 + ipad; ny = ly + 2 + jpad; nz = lz + 2
      count = int(nx, c_size_t) * ny * nz
      bytes = count * c_sizeof(svisc)

! build_bcond / hencol (SRC/build.bcond.F90:29-35, SRC/hencol.F90:31,52)
      if (trim(case) == 'channel') then
         u00 = zero; uf = u0
      else
         u00 = u0;   uf = zero
      endif
      omega = (uno+uno)/((6.0*svisc)+1.0)
      fgrad = 4.0*uf*svisc/(5.0*real(ly,c_real)*real(ly,c_real))
      force = u00/(6.0)
      mode = 0
      if (trim(case) == 'periodic') mode = 1

! f = f_eq written straight into both lattices on the device (bit-identical
! to the reference's host init; sin/cos per coordinate on the host)
      allocate(h(0:nx-1, 0:ny-1, 0:nz-1, 19))
      do d = 1, 19
         ph(d) = c_loc(h(0,0,0,d))
      enddo
      do d = 1, 19
         rc = bgk_dev_malloc(pa(d), bytes)
         if (rc /= 0) stop 'hipMalloc failed'
      enddo
      do d = 1, 18
         rc = bgk_dev_malloc(pb(d), bytes)
         if (rc /= 0) stop 'hipMalloc failed'
      enddo
      pb(19) = pa(19)
! bgk_init_gpu writes the cells 0:l+1, 0:m+1, 0:n+1; the row and plane
! padding (ipad, jpad > 0) is zeroed first so that out.bin is defined there
      h = zero
      do d = 1, 19
         rc = bgk_memcpy_h2d(pa(d), ph(d), bytes)
         if (rc /= 0) stop 'h2d copy failed'
      enddo
      do d = 1, 18
         rc = bgk_memcpy_h2d(pb(d), ph(d), bytes)
         if (rc /= 0) stop 'h2d copy failed'
      enddo
      rc = bgk_init_gpu(pa, lx, ly, lz, nx, ny, mode, u00)
      if (rc /= 0) stop 'bgk_init_gpu failed'
      rc = bgk_init_gpu(pb, lx, ly, lz, nx, ny, mode, u00)
      if (rc /= 0) stop 'bgk_init_gpu failed'

      open(63, file=trim(outdir)//'/diagno.dat', status='unknown')
      call gpu_device_sync()
      call system_clock(c0, crate)
      do itime = 1, itfin
! boundaries (SRC/boundaries.F90:31-43)
         select case (trim(case))
         case ('periodic')
            call bcond_periodic_gpu(pa, lx, ly, lz, nx, ny)
         case ('channel')
            call bcond_channel_gpu(pa, lx, ly, lz, nx, ny)
         case default
            call bcond_driven_gpu(pa, lx, ly, lz, nx, ny, force)
         end select
! collision (SRC/collision.F90:42-46)
         if (trim(case) == 'channel') then
            call col_mc_force_gpu(pa, pb, lx, ly, lz, nx, ny, omega, fgrad, 1)
         else
            call col_mc_gpu(pa(1),pa(2),pa(3),pa(4),pa(5),pa(6),pa(7),    &
                 pa(8),pa(9),pa(10),pa(11),pa(12),pa(13),pa(14),pa(15),   &
                 pa(16),pa(17),pa(18),pa(19),                            &
                 pb(1),pb(2),pb(3),pb(4),pb(5),pb(6),pb(7),pb(8),pb(9),   &
                 pb(10),pb(11),pb(12),pb(13),pb(14),pb(15),pb(16),pb(17), &
                 pb(18), lx, ly, lz, nx, ny,                             &
                 omega, zero, uno, p0, p1, p2, 3.0_c_real, 1.5_c_real,   &
                 3.0_c_real, zero, zero, zero)
         endif
! pointer swap a <-> b for directions 1..18
         do d = 1, 18
            tmp = pa(d); pa(d) = pb(d); pb(d) = tmp
         enddo
         if (mod(itime, icheck) == 0) then
            rc = diagno_gpu(pa, lx, ly, lz, nx, ny, uno, dg)
            write(63, '(i8,5(e14.6,1x))') itime, dg(1), dg(2), dg(3), dg(4), dg(5)
         endif
      enddo
      call gpu_device_sync()
      call system_clock(c1, crate)
      tloop = real(c1 - c0, 8) / real(crate, 8)
      close(63)
      write(6, '(a,es16.8)') '# Mlups ', real(lx,8)*ly*lz*itfin/max(tloop,1d-9)/1d6

      do d = 1, 19
         rc = bgk_memcpy_d2h(ph(d), pa(d), bytes)
      enddo
      open(99, file=trim(outdir)//'/out.bin', access='stream', form='unformatted', &
           status='replace')
      write(99) h
      close(99)
      do d = 1, 19
         rc = bgk_dev_free(pa(d))
      enddo
      do d = 1, 18
         rc = bgk_dev_free(pb(d))
      enddo
      end program bgk3d_host
