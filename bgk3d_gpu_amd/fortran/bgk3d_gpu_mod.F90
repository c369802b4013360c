!=====================================================================
! bgk3d_gpu_mod -- Fortran iso_c_binding interfaces to libbgk3d_gpu_sp.so
! (default) or libbgk3d_gpu_dp.so (-DDOUBLE_P), the gfx950 drop-in for the
! reference's col_MC_hip.o + bcond_hip.o.
!
! The reference symbols keep exactly the interfaces the reference declares
! (SRC/col_MC.F90:101-123, SRC/bcond_gpu_mod.F90:30-77): device pointers
! as type(c_ptr), sizes and scalars by value, the 19 lattice pointers of a
! boundary call as a host array.  The additional entry points (device
! diagnostics, channel force, host init, memory plumbing) follow the same
! conventions.  Every function returning integer(c_int) returns 0 on
! success.
!=====================================================================
      module bgk3d_gpu_mod
      use iso_c_binding
      implicit none
#ifdef DOUBLE_P
      integer, parameter :: c_real = c_double
#else
      integer, parameter :: c_real = c_float
#endif

      interface
        subroutine col_mc_gpu(                                           &
            a01,a02,a03,a04,a05,a06,a07,a08,a09,a10,                     &
            a11,a12,a13,a14,a15,a16,a17,a18,a19,                         &
            b01,b02,b03,b04,b05,b06,b07,b08,b09,b10,                     &
            b11,b12,b13,b14,b15,b16,b17,b18,                             &
            l,m,n,nx,ny,                                                 &
            omega,cte0,cte1,p0,p1,p2,rf,qf,tre,                          &
            forcex,forcey,forcez) bind(C, name="col_mc_gpu")
          import :: c_real, c_int, c_ptr
          type(c_ptr), value :: a01,a02,a03,a04,a05,a06,a07,a08,a09,a10
          type(c_ptr), value :: a11,a12,a13,a14,a15,a16,a17,a18,a19
          type(c_ptr), value :: b01,b02,b03,b04,b05,b06,b07,b08,b09,b10
          type(c_ptr), value :: b11,b12,b13,b14,b15,b16,b17,b18
          integer(c_int), value :: l,m,n,nx,ny
          real(c_real), value :: omega,cte0,cte1,p0,p1,p2,rf,qf,tre
          real(c_real), value :: forcex,forcey,forcez
        end subroutine col_mc_gpu

        subroutine bcond_periodic_gpu(a,l,m,n,nx,ny)                     &
                   bind(C, name="bcond_periodic_gpu")
          import :: c_ptr, c_int
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny
        end subroutine bcond_periodic_gpu

        subroutine bcond_channel_gpu(a,l,m,n,nx,ny)                      &
                   bind(C, name="bcond_channel_gpu")
          import :: c_ptr, c_int
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny
        end subroutine bcond_channel_gpu

        subroutine bcond_driven_gpu(a,l,m,n,nx,ny,force)                 &
                   bind(C, name="bcond_driven_gpu")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny
          real(c_real),   value :: force
        end subroutine bcond_driven_gpu

        subroutine bcond_inflow_gpu(a,l,m,n,nx,ny,                       &
                   u_inflow,rf,qf,tre,p1,p2,cte1)                        &
                   bind(C, name="bcond_inflow_gpu")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny
          real(c_real),   value :: u_inflow,rf,qf,tre,p1,p2,cte1
        end subroutine bcond_inflow_gpu

        subroutine bcond_obs_gpu(a,obs,l,m,n,nx,ny,                      &
                   imin,imax,jmin,jmax,kmin,kmax,obs_nx,obs_ny)          &
                   bind(C, name="bcond_obs_gpu")
          import :: c_ptr, c_int
          type(c_ptr)           :: a(19)
          type(c_ptr),    value :: obs
          integer(c_int), value :: l,m,n,nx,ny
          integer(c_int), value :: imin,imax,jmin,jmax,kmin,kmax
          integer(c_int), value :: obs_nx,obs_ny
        end subroutine bcond_obs_gpu

        subroutine gpu_device_sync() bind(C, name="gpu_device_sync")
        end subroutine gpu_device_sync

        ! ---- additions in the same style ----------------------------
        subroutine col_mc_force_gpu(a,b,l,m,n,nx,ny,omega,fgrad,axis)    &
                   bind(C, name="col_mc_force_gpu")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: a(19), b(19)
          integer(c_int), value :: l,m,n,nx,ny,axis
          real(c_real),   value :: omega,fgrad
        end subroutine col_mc_force_gpu

        subroutine col_mc_les_gpu(a,b,l,m,n,nx,ny,omega1,cte_s,fgrad,   &
                   axis) bind(C, name="col_mc_les_gpu")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: a(19), b(19)
          integer(c_int), value :: l,m,n,nx,ny,axis
          real(c_real),   value :: omega1,cte_s,fgrad
        end subroutine col_mc_les_gpu

        integer(c_int) function diagno_gpu(a,l,m,n,nx,ny,cte1,out)       &
                   bind(C, name="diagno_gpu")
          import :: c_ptr, c_int, c_real, c_double
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny
          real(c_real),   value :: cte1
          real(c_double)        :: out(5)
        end function diagno_gpu

        integer(c_int) function varm_gpu(a,l,m,n,nx,ny,cte1,u,w,v,den)   &
                   bind(C, name="varm_gpu")
          import :: c_ptr, c_int, c_real, c_double
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny
          real(c_real),   value :: cte1
          real(c_double)        :: u(*), w(*), v(*), den(*)
        end function varm_gpu

        integer(c_int) function draglift_gpu(a,obs,l,m,n,nx,ny,          &
                   icoord,jcoord,kcoord,radius,u_inflow,out)             &
                   bind(C, name="draglift_gpu")
          import :: c_ptr, c_int, c_real, c_double
          type(c_ptr)           :: a(19)
          type(c_ptr),    value :: obs
          integer(c_int), value :: l,m,n,nx,ny,icoord,jcoord,kcoord,radius
          real(c_real),   value :: u_inflow
          real(c_double)        :: out(3)
        end function draglift_gpu

        integer(c_int) function prof_k_gpu(a,l,m,n,nx,ny,icoord,jcoord,  &
                   cte1,out) bind(C, name="prof_k_gpu")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny,icoord,jcoord
          real(c_real),   value :: cte1
          real(c_real)          :: out(*)
        end function prof_k_gpu

        integer(c_int) function prof_i_gpu(a,l,m,n,nx,ny,jcoord,kcoord,  &
                   cte1,out) bind(C, name="prof_i_gpu")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny,jcoord,kcoord
          real(c_real),   value :: cte1
          real(c_real)          :: out(*)
        end function prof_i_gpu

        integer(c_int) function prof_j_gpu(a,l,m,n,nx,ny,icoord,kcoord,  &
                   cte1,out) bind(C, name="prof_j_gpu")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny,icoord,kcoord
          real(c_real),   value :: cte1
          real(c_real)          :: out(*)
        end function prof_j_gpu

        integer(c_int) function probe_gpu(a,l,m,n,nx,ny,i0,j0,k0,cte1,   &
                   out) bind(C, name="probe_gpu")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny,i0,j0,k0
          real(c_real),   value :: cte1
          real(c_real)          :: out(4)
        end function probe_gpu

        integer(c_int) function bgk_init_host(h,l,m,n,nx,ny,mode,u00)    &
                   bind(C, name="bgk_init_host")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: h(19)
          integer(c_int), value :: l,m,n,nx,ny,mode
          real(c_real),   value :: u00
        end function bgk_init_host

        integer(c_int) function bgk_init_gpu(a,l,m,n,nx,ny,mode,u00)     &
                   bind(C, name="bgk_init_gpu")
          import :: c_ptr, c_int, c_real
          type(c_ptr)           :: a(19)
          integer(c_int), value :: l,m,n,nx,ny,mode
          real(c_real),   value :: u00
        end function bgk_init_gpu

        integer(c_int) function bgk_dev_malloc(ptr,bytes)                &
                   bind(C, name="bgk_dev_malloc")
          import :: c_ptr, c_size_t
          type(c_ptr)              :: ptr
          integer(c_size_t), value :: bytes
        end function bgk_dev_malloc

        integer(c_int) function bgk_dev_free(ptr) bind(C, name="bgk_dev_free")
          import :: c_ptr
          type(c_ptr), value :: ptr
        end function bgk_dev_free

        integer(c_int) function bgk_memcpy_h2d(dst,src,bytes)            &
                   bind(C, name="bgk_memcpy_h2d")
          import :: c_ptr, c_size_t
          type(c_ptr), value       :: dst, src
          integer(c_size_t), value :: bytes
        end function bgk_memcpy_h2d

        integer(c_int) function bgk_memcpy_d2h(dst,src,bytes)            &
                   bind(C, name="bgk_memcpy_d2h")
          import :: c_ptr, c_size_t
          type(c_ptr), value       :: dst, src
          integer(c_size_t), value :: bytes
        end function bgk_memcpy_d2h
      end interface
      end module bgk3d_gpu_mod
