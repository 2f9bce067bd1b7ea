!
!  horses_gpu_driver.f90 -- a Fortran host that drives libhorses2d_gpu.so through
!  horses_gpu_bindings, following TimeIntegrator_Integrate
!  (Approximation/DGTimeIntegrator.f90:159-255): estimate dt, one TimeScheme call per
!  iteration, residual monitor, dt re-estimate every 100 iterations and, in steady mode, the
!  stop once the largest residual is below the residual target (:224-234).  The case is the
!  Euler isentropic vortex on the periodic [-5,5]^2 square (Roe, RK3, CFL 0.4), meshed like
!  TestCases/Cases/VortexTransport.
!
!  Usage: horses_gpu_driver <iterations> <output file> [residual target]
!             16x16 quads, N = 4; with a target > 0, steady mode; writes the final Q, one
!             value per line, and prints the last iteration
!         horses_gpu_driver batch <iterations> <output file>
!             the 16x16 N = 4 case and an 8x8 N = 6 case advanced together by the batched form
!             (hg_batch_rk_step), each with its own dt; writes both final states, in that order
!
program horses_gpu_driver
   use iso_c_binding
   use horses_gpu_bindings
   implicit none
   real(c_double), parameter :: CFL = 0.4d0
   type(c_ptr)               :: h(2), batch
   type(c_ptr), target       :: hs(2)
   type(hg_batch_options)    :: bopt
   real(c_double), allocatable :: Q(:)
   real(c_double)            :: dt(2), t(2), res(4), target_res
   integer(c_int64_t)        :: ndof
   integer                   :: it, niter, unit, c, k, last
   character(len=256)        :: arg, outfile

   call get_command_argument(1, arg)
   if (trim(arg) == "batch") then
      call get_command_argument(2, arg); read(arg, *) niter
      call get_command_argument(3, outfile)
      call vortex_case(16, 4, h(1))
      call vortex_case(8, 6, h(2))
      do c = 1, 2
         call hgf_check(hg_estimate_dt(h(c), 1.0d30, merge(CFL, 0.1d0, c == 1), dt(c)), "hg_estimate_dt")
      end do
      hs = h
      bopt = hg_batch_options(-1, -1, 1)
      call hgf_check(hg_batch_create(hs, 2_c_int32_t, bopt, batch), "hg_batch_create")
      t = 0.0d0
      do it = 1, niter
         call hgf_check(hg_batch_rk_step(batch, t, dt), "hg_batch_rk_step")
         t = t + dt
      end do
      call hgf_check(hg_batch_synchronize(batch), "hg_batch_synchronize")
      open(newunit=unit, file=trim(outfile), status="replace", action="write")
      do c = 1, 2
         ndof = hg_ndof(h(c))
         allocate(Q(ndof))
         call hgf_check(hg_download_state(h(c), Q, ndof), "hg_download_state")
         do k = 1, int(ndof)
            write(unit, '(ES25.17)') Q(k)
         end do
         deallocate(Q)
      end do
      close(unit)
      write(*, '(A,I0,A)') "batch: 2 cases, ", niter, " iterations"
      call hgf_check(hg_batch_destroy(batch), "hg_batch_destroy")
      do c = 1, 2
         call hgf_check(hg_destroy(h(c)), "hg_destroy")
      end do
      stop
   end if

   read(arg, *) niter
   call get_command_argument(2, outfile)
   target_res = 0.0d0
   if (command_argument_count() >= 3) then
      call get_command_argument(3, arg); read(arg, *) target_res
   end if
   call vortex_case(16, 4, h(1))
   t(1) = 0.0d0
   call hgf_check(hg_estimate_dt(h(1), 1.0d30, CFL, dt(1)), "hg_estimate_dt")
   last = 0
   do it = 1, niter
      last = it
      call hgf_check(hg_rk_step(h(1), t(1), dt(1)), "hg_rk_step")
      t(1) = t(1) + dt(1)
      call hgf_check(hg_residuals(h(1), res), "hg_residuals")
      if (mod(it, 100) == 0) call hgf_check(hg_estimate_dt(h(1), 1.0d30, CFL, dt(1)), "hg_estimate_dt")
      if (target_res > 0.0d0) then
         if (maxval(abs(res)) < target_res) exit
      end if
   end do
   ndof = hg_ndof(h(1))
   allocate(Q(ndof))
   call hgf_check(hg_download_state(h(1), Q, ndof), "hg_download_state")
   write(*, '(A,I0,A,ES24.16,A,4ES12.4)') "iterations ", last, " t ", t(1), " residuals ", res
   open(newunit=unit, file=trim(outfile), status="replace", action="write")
   do k = 1, int(ndof)
      write(unit, '(ES25.17)') Q(k)
   end do
   close(unit)
   call hgf_check(hg_destroy(h(1)), "hg_destroy")

contains

   ! The vortex case on an n x n periodic mesh of order N: HiOMesh arrays (nodes y-fastest,
   ! elements y-fastest, corners counter-clockwise), the four periodic zones, hg_create, and the
   ! initial condition evaluated on the host (the libproblemfile hook) and uploaded.
   subroutine vortex_case(n, N_POLY, handle)
      integer, intent(in)      :: n, N_POLY
      type(c_ptr), intent(out) :: handle
      real(c_double), parameter :: MACH = 0.5d0, BETA = 5.0d0, PI = 3.141592653589793238462643d0
      real(c_double), allocatable, target :: points(:,:), bc_real(:,:), X(:,:), Qc(:)
      integer(c_int32_t), allocatable, target :: quads(:,:), order(:), bedges(:,:), bmarker(:), bc_type(:), &
                                                 bc_int(:,:), bc_rs(:)
      type(hg_mesh_desc) :: mesh
      type(hg_physics)   :: phys
      type(hg_options)   :: opt
      integer(c_int64_t) :: nd, nnodes
      integer            :: i, j, kk, e
      real(c_double)     :: xx, yy, f, u, v, Tm, rho

      allocate(points(2, (n+1)*(n+1)), quads(4, n*n), order(n*n), bedges(2, 4*n), bmarker(4*n))
      do i = 0, n ; do j = 0, n
         points(1, i*(n+1) + j + 1) = -5.0d0 + 10.0d0 * i / n
         points(2, i*(n+1) + j + 1) = -5.0d0 + 10.0d0 * j / n
      end do ; end do
      do i = 0, n-1 ; do j = 0, n-1
         e = i*n + j + 1
         kk = i*(n+1) + j
         quads(:, e) = [kk, kk + (n+1), kk + (n+1) + 1, kk + 1]
      end do ; end do
      order = N_POLY
      do i = 0, n-1
         bedges(:, i+1)       = [i*(n+1), (i+1)*(n+1)]                  ; bmarker(i+1)       = 1
         bedges(:, n+i+1)     = [(i+1)*(n+1) + n, i*(n+1) + n]          ; bmarker(n+i+1)     = 2
         bedges(:, 2*n+i+1)   = [i+1, i]                                ; bmarker(2*n+i+1)   = 3
         bedges(:, 3*n+i+1)   = [n*(n+1) + i, n*(n+1) + i + 1]          ; bmarker(3*n+i+1)   = 4
      end do
      allocate(bc_type(4), bc_int(2,4), bc_real(4,4), bc_rs(4))
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
      call hgf_check(hg_create(mesh, phys, opt, handle), "hg_create")
      nd = hg_ndof(handle) ; nnodes = nd / 4
      allocate(X(2, nnodes), Qc(nd))
      call hgf_check(hg_get_node_coords(handle, X, nnodes), "hg_get_node_coords")
      do kk = 1, int(nnodes)
         xx = X(1, kk) ; yy = X(2, kk)
         f = exp(0.5d0 * (1.0d0 - (xx*xx + yy*yy)))
         u = MACH * sqrt(1.4d0) - BETA / (2.0d0 * PI) * yy * f
         v = BETA / (2.0d0 * PI) * xx * f
         Tm = 1.0d0 - 0.4d0 * BETA * BETA / (8.0d0 * 1.4d0 * PI * PI) * f * f
         rho = Tm ** (1.0d0 / 0.4d0)
         Qc(4*kk-3) = rho ; Qc(4*kk-2) = rho * u ; Qc(4*kk-1) = rho * v
         Qc(4*kk)   = rho * Tm / 0.4d0 + 0.5d0 * rho * (u*u + v*v)
      end do
      call hgf_check(hg_upload_state(handle, Qc, nd), "hg_upload_state")
   end subroutine vortex_case

end program horses_gpu_driver
