!=====================================================================
! ref_harness -- TEST INFRASTRUCTURE ONLY.
!
! A small driver of our own that links the reference's Fortran CPU
! routines (compiled from their sources under the reference tree by
! oracle/Makefile into oracle/_ref/) and runs them the way the
! reference's time loop does (SRC/bgk3d.F90:169-196), so that the C
! oracle and the committed fixtures can be checked against the real
! reference arithmetic.  It never travels into the product library.
!
! Usage:
!   ref_harness.x OUTDIR L M N SVISC U0 ITFIN ICHECK INIT_V IPAD JPAD MODE [WRITE]
! MODE 0: full loop (boundaries -> propagation -> collision, as
!         SRC/bgk3d.F90:177-179; propagation is empty in the fused build)
!         for ITFIN steps, with the
!         reference's diagno/probe/prof_i/prof_j/prof_k (and draglift
!         under DRAG) every ICHECK, as SRC/diagnostic.F90:98-120 calls them;
! MODE 1: one call of `boundaries` only;
! MODE 2: one call of `col_MC` only (includes the a<->b pointer swap);
! MODE 3: like MODE 0 but timing only: no diagnostics, no output files.
! INIT_V >= 0 calls the reference's init(INIT_V); INIT_V = -1 reads the
! 19 populations (full halo arrays, storage kind, raw) from OUTDIR/in.bin;
! INIT_V = -2 calls the reference's restore_raw (file restore.bin in the
! working directory) and runs the steps after the restored time step.
! WRITE = 1 (MODE 0): after the loop call the reference's save_raw,
! save_vel, vtk_xy_bin(n/2), vtk_xz_bin(m/2), vtk_yz_bin(l/2) and
! vtk_om_xy_bin(n/2); they write into the working directory.
! Output: OUTDIR/out.bin (a01..a19 after the run, raw), the reference's
! own text files (diagno.dat, prof_k.dat, u_med.dat, drag.lift.dat,
! bgk.log) and one stdout line "# Mlups <value>" with the loop rate
! (SRC/finalize.F90:106-107 definition: l*m*n*steps/time/1e6).
!=====================================================================
      program ref_harness
      use storage
      use timing
      implicit none
      character(len=512) :: outdir, arg
      integer :: itfin, icheck, init_v, mode, itime, iwrite, itstart
      integer(8) :: c0, c1, crate, cmax
      real(8) :: tloop

      call get_command_argument(1, outdir)
      call get_command_argument(2, arg); read(arg,*) lx
      call get_command_argument(3, arg); read(arg,*) ly
      call get_command_argument(4, arg); read(arg,*) lz
      call get_command_argument(5, arg); read(arg,*) svisc
      call get_command_argument(6, arg); read(arg,*) u0
      call get_command_argument(7, arg); read(arg,*) itfin
      call get_command_argument(8, arg); read(arg,*) icheck
      call get_command_argument(9, arg); read(arg,*) init_v
      call get_command_argument(10, arg); read(arg,*) ipad
      call get_command_argument(11, arg); read(arg,*) jpad
      call get_command_argument(12, arg); read(arg,*) mode
      iwrite = 0
      if (command_argument_count() >= 13) then
         call get_command_argument(13, arg); read(arg,*) iwrite
      endif
      itstart = 0

! what SRC/input.F90:50-72 sets after the namelist read
      myrank = 0
      flag1 = 0; flag2 = 0; flag3 = 0
      kpad = 0
      cteS = 0.1
      l = lx; m = ly; n = lz
      l1 = l+1; m1 = m+1; n1 = n+1
      u0x = zero; u0y = zero; u0z = zero
      offset = 0

      open(16, file=trim(outdir)//'/bgk.log',      status='unknown')
      open(62, file=trim(outdir)//'/u_med.dat',    status='unknown')
      open(63, file=trim(outdir)//'/diagno.dat',   status='unknown')
      open(61, file=trim(outdir)//'/prof_i.dat',   status='unknown')
      open(64, file=trim(outdir)//'/prof_j.dat',   status='unknown')
      open(65, file=trim(outdir)//'/prof_k.dat',   status='unknown')
      open(68, file=trim(outdir)//'/probe.dat',    status='unknown')
      open(66, file=trim(outdir)//'/drag.lift.dat',status='unknown')

! SRC/setup.F90:140 and SRC/initialize.F90:70-139
      call alloca()
      call build_bcond
#ifdef OBSTACLE
      call build_obs
#endif
      if (init_v >= 0) then
         call init(init_v)
      else if (init_v == -2) then
         call restore_raw(itstart)
      else
         open(98, file=trim(outdir)//'/in.bin', access='stream', &
              form='unformatted', status='old')
         read(98) a01,a02,a03,a04,a05,a06,a07,a08,a09,a10, &
                  a11,a12,a13,a14,a15,a16,a17,a18,a19
         close(98)
      endif
      call hencol

      if (mode == 1) then
         call boundaries
      else if (mode == 2) then
         call col_MC(1)
      else
         if (mode == 0) then
            call diagno(itstart)
            call varm(itstart)
            call point_and_lines(itstart)
         endif
         call system_clock(c0, crate, cmax)
         do itime = itstart+1, itfin
            call boundaries
            call propagation
            call collision(itime)
            if (mode == 0 .and. mod(itime, icheck) == 0) then
               call diagno(itime)
               call point_and_lines(itime)
#ifdef DRAG
               call draglift(itime)
#endif
            endif
         enddo
         call system_clock(c1, crate, cmax)
         tloop = real(c1-c0,8)/real(crate,8)
         if (mode == 0) call varm(itfin)
         if (mode == 0 .and. iwrite == 1) then
            call save_raw(itfin)
            call save_vel(itfin)
            call vtk_xy_bin(itfin, n/2)
            call vtk_xz_bin(itfin, m/2)
            call vtk_yz_bin(itfin, l/2)
            call vtk_om_xy_bin(itfin, n/2)
         endif
         if (tloop > 0.0d0) then
            write(6,'(a,es16.8)') '# Mlups ', &
                 real(l,8)*real(m,8)*real(n,8)*real(itfin-itstart,8)/tloop/1.0d6
         endif
         write(6,'(a,es16.8)') '# seconds ', tloop
      endif

      if (mode == 3) stop
      open(99, file=trim(outdir)//'/out.bin', access='stream', &
           form='unformatted', status='replace')
      write(99) a01,a02,a03,a04,a05,a06,a07,a08,a09,a10, &
                a11,a12,a13,a14,a15,a16,a17,a18,a19
      close(99)
#ifdef OBSTACLE
      open(97, file=trim(outdir)//'/obs.bin', access='stream', &
           form='unformatted', status='replace')
      write(97) obs
      close(97)
      write(6,'(a,7i8)') '# obs box/radius ', imin, imax, jmin, jmax, &
           kmin, kmax, radius
#endif
      write(6,'(a,2es24.16)') '# omega fgrad ', real(omega,8), real(fgrad,8)
      close(16); close(61); close(62); close(63); close(64); close(65)
      close(66); close(68)

      contains

! the per-ICHECK probe and line profiles of SRC/diagnostic.F90:98-120
      subroutine point_and_lines(it)
      integer, intent(in) :: it
#ifdef TGV3D
      call probe(it, l/3, m/5, n/3)
#else
      call probe(it, l/2, m/2, n/2)
#endif
      call prof_i(it, m/2, n/2)
      call prof_j(it, l/2, n/2)
      call prof_k(it, l/2, m/2)
      end subroutine point_and_lines
      end program ref_harness
