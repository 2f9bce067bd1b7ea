!
!  horses_gpu_driver.f90 -- a Fortran host that drives libhorses2d_gpu.so through
!  horses_gpu_bindings, following TimeIntegrator_Integrate
!  (Approximation/DGTimeIntegrator.f90:159-255): estimate dt, clamp to the final time,
!  one RK3 TimeScheme call per iteration, residual monitor, dt re-estimate every 100
!  iterations.  The case is the Euler isentropic vortex on the periodic [-5,5]^2 square
!  (16x16 quads, N = 4, Roe, CFL 0.4), meshed like TestCases/Cases/VortexTransport.
!
!  Usage: horses_gpu_driver <iterations> <output file>   (writes the final Q, one per line)
!
program horses_gpu_driver
   use iso_c_binding
   use horses_gpu_bindings
   implicit none
   integer, parameter :: n = 16, N_POLY = 4
   real(c_double), parameter :: MACH = 0.5d0, BETA = 5.0d0, CFL = 0.4d0, PI = 3.141592653589793238462643d0
   real(c_double), allocatable, target :: points(:,:), cxy(:), bc_real(:,:), X(:,:), Q(:)
   integer(c_int32_t), allocatable, target :: quads(:,:), order(:), bedges(:,:), bmarker(:), bc_type(:), &
                                              bc_int(:,:), bc_rs(:), cedges(:)
   type(hg_mesh_desc) :: mesh
   type(hg_physics)   :: phys
   type(hg_options)   :: opt
   type(c_ptr)        :: h
   integer(c_int64_t) :: ndof, nnodes
   integer            :: i, j, k, e, it, niter, unit
   real(c_double)     :: dt, t, res(4), xx, yy, f, u, v, Tm, rho
   character(len=256) :: arg, outfile

   call get_command_argument(1, arg); read(arg, *) niter
   call get_command_argument(2, outfile)
!
!  HiOMesh arrays: nodes y-fastest, elements y-fastest, corners counter-clockwise
!
   allocate(points(2, (n+1)*(n+1)), quads(4, n*n), order(n*n), bedges(2, 4*n), bmarker(4*n))
   do i = 0, n ; do j = 0, n
      points(1, i*(n+1) + j + 1) = -5.0d0 + 10.0d0 * i / n
      points(2, i*(n+1) + j + 1) = -5.0d0 + 10.0d0 * j / n
   end do ; end do
   do i = 0, n-1 ; do j = 0, n-1
      e = i*n + j + 1
      k = i*(n+1) + j
      quads(:, e) = [k, k + (n+1), k + (n+1) + 1, k + 1]
   end do ; end do
   order = N_POLY
   do i = 0, n-1
      bedges(:, i+1)       = [i*(n+1), (i+1)*(n+1)]                  ; bmarker(i+1)       = 1
      bedges(:, n+i+1)     = [(i+1)*(n+1) + n, i*(n+1) + n]          ; bmarker(n+i+1)     = 2
      bedges(:, 2*n+i+1)   = [i+1, i]                                ; bmarker(2*n+i+1)   = 3
      bedges(:, 3*n+i+1)   = [n*(n+1) + i, n*(n+1) + i + 1]          ; bmarker(3*n+i+1)   = 4
   end do
   allocate(bc_type(4), bc_int(2,4), bc_real(4,4), bc_rs(4), cxy(1), cedges(1))
   bc_type = HG_BC_PERIODIC
   bc_int(:,1) = [2, 2] ; bc_int(:,2) = [2, 1] ; bc_int(:,3) = [1, 4] ; bc_int(:,4) = [1, 3]
   bc_real = 0.0d0 ; bc_rs = -1
   mesh%nnodes = size(points, 2)      ; mesh%points = c_loc(points)
   mesh%nelements = size(quads, 2)    ; mesh%quads = c_loc(quads) ; mesh%order = c_loc(order)
   mesh%nbedges = size(bedges, 2)     ; mesh%bedges = c_loc(bedges) ; mesh%bmarker = c_loc(bmarker)
   mesh%ncurved = 0 ; mesh%np1 = 0    ; mesh%cedges = c_null_ptr ; mesh%cxy = c_null_ptr
   mesh%nmarkers = 4
   mesh%bc_type = c_loc(bc_type) ; mesh%bc_int = c_loc(bc_int) ; mesh%bc_real = c_loc(bc_real)
   mesh%bc_riemann = c_loc(bc_rs)
   phys = hg_physics(MACH, 0.0d0, 0.72d0)
   opt = hg_options(0, HG_RIEMANN_ROE, HG_VISCOUS_NONE, HG_IP_SIPG, 1.0d0, HG_PRECISION_FP64, HG_SCHEME_RK3, &
                    HG_AD_NONE, 1.0d0)
   call hgf_check(hg_create(mesh, phys, opt, h), "hg_create")
!
!  Initial condition on the host (the libproblemfile hook): Shu's isentropic vortex
!
   ndof = hg_ndof(h) ; nnodes = ndof / 4
   allocate(X(2, nnodes), Q(ndof))
   call hgf_check(hg_get_node_coords(h, X, nnodes), "hg_get_node_coords")
   do k = 1, int(nnodes)
      xx = X(1, k) ; yy = X(2, k)
      f = exp(0.5d0 * (1.0d0 - (xx*xx + yy*yy)))
      u = MACH * sqrt(1.4d0) - BETA / (2.0d0 * PI) * yy * f
      v = BETA / (2.0d0 * PI) * xx * f
      Tm = 1.0d0 - 0.4d0 * BETA * BETA / (8.0d0 * 1.4d0 * PI * PI) * f * f
      rho = Tm ** (1.0d0 / 0.4d0)
      Q(4*k-3) = rho ; Q(4*k-2) = rho * u ; Q(4*k-1) = rho * v
      Q(4*k)   = rho * Tm / 0.4d0 + 0.5d0 * rho * (u*u + v*v)
   end do
   call hgf_check(hg_upload_state(h, Q, ndof), "hg_upload_state")
!
!  Time loop
!
   t = 0.0d0
   call hgf_check(hg_estimate_dt(h, 1.0d30, CFL, dt), "hg_estimate_dt")
   do it = 1, niter
      call hgf_check(hg_rk_step(h, t, dt), "hg_rk_step")
      t = t + dt
      call hgf_check(hg_residuals(h, res), "hg_residuals")
      if (mod(it, 100) == 0) call hgf_check(hg_estimate_dt(h, 1.0d30, CFL, dt), "hg_estimate_dt")
   end do
   call hgf_check(hg_download_state(h, Q, ndof), "hg_download_state")
   write(*, '(A,I0,A,ES24.16,A,4ES12.4)') "iterations ", niter, " t ", t, " residuals ", res
   open(newunit=unit, file=trim(outfile), status="replace", action="write")
   do k = 1, int(ndof)
      write(unit, '(ES25.17)') Q(k)
   end do
   close(unit)
   call hgf_check(hg_destroy(h), "hg_destroy")
end program horses_gpu_driver
