!
!  horses_gpu_bindings.f90 -- ISO_C_BINDING interface of libhorses2d_gpu.so for the
!  HORSES2D Fortran solver (declarations mirror include/horses2d_gpu.h).
!
!  The solver keeps its HiOCase, HiOMesh, libproblemfile and HORSES2D.NS/.Euler entry
!  points; a TimeScheme procedure built on hgf_rk_step replaces
!  TimeIntegrator_WilliamsonRK3 (Approximation/DGTimeIntegrator.f90:282-317); the
!  solver-side adapter is a subroutine of that interface calling hg_rk_step.
!
module horses_gpu_bindings
   use iso_c_binding
   implicit none
   private

   integer(c_int32_t), parameter, public :: HG_BC_PERIODIC = 1, HG_BC_DIRICHLET = 2, HG_BC_EULERWALL = 3, &
                                            HG_BC_VISCOUSWALL = 4, HG_BC_FARFIELD = 5, &
                                            HG_BC_PRESSUREOUTLET = 6, HG_BC_PRESSUREINLET = 7, &
                                            HG_BC_NEWDIRICHLET = 9
   integer(c_int32_t), parameter, public :: HG_OUTFLOW_PARTIALLY_NON_REFLECTIVE = 0, HG_OUTFLOW_REFLECTIVE = 1
   integer(c_int32_t), parameter, public :: HG_NEWDIRICHLET_SPECIFY_SPEED = 1, &
                                            HG_NEWDIRICHLET_SPECIFY_TOTAL_PRESSURE = 2, &
                                            HG_NEWDIRICHLET_STATIC_PRESSURE = 1, HG_NEWDIRICHLET_TOTAL_PRESSURE = 2
   integer(c_int32_t), parameter, public :: HG_AD_NONE = 0, HG_AD_PHYSICAL = 1, HG_AD_LAPLACIAN = 2
   integer(c_int32_t), parameter, public :: HG_PROBE_RHO = 0, HG_PROBE_RHOU = 1, HG_PROBE_RHOV = 2, &
                                            HG_PROBE_RHOE = 3, HG_PROBE_RHOT = 4, HG_PROBE_RHOUT = 5, &
                                            HG_PROBE_RHOVT = 6, HG_PROBE_RHOET = 7, HG_PROBE_U = 8, &
                                            HG_PROBE_V = 9, HG_PROBE_P = 10, HG_PROBE_MACH = 11
   integer(c_int32_t), parameter, public :: HG_RIEMANN_ROE = 0, HG_RIEMANN_HLL = 1, HG_RIEMANN_EXACT = 2, &
                                            HG_RIEMANN_LF = 3
   integer(c_int32_t), parameter, public :: HG_VISCOUS_NONE = 0, HG_VISCOUS_BR1 = 1, HG_VISCOUS_IP = 2
   integer(c_int32_t), parameter, public :: HG_IP_SIPG = 0, HG_IP_NIPG = 1, HG_IP_IIPG = 2
   integer(c_int32_t), parameter, public :: HG_PRECISION_FP64 = 0, HG_PRECISION_MIXED = 1
   integer(c_int32_t), parameter, public :: HG_SCHEME_EULER = 0, HG_SCHEME_RK3 = 1, HG_SCHEME_RK5 = 2

   type, bind(C), public :: hg_mesh_desc
      integer(c_int32_t) :: nnodes
      type(c_ptr)        :: points
      integer(c_int32_t) :: nelements
      type(c_ptr)        :: quads
      type(c_ptr)        :: order
      integer(c_int32_t) :: nbedges
      type(c_ptr)        :: bedges
      type(c_ptr)        :: bmarker
      integer(c_int32_t) :: ncurved
      integer(c_int32_t) :: np1
      type(c_ptr)        :: cedges
      type(c_ptr)        :: cxy
      integer(c_int32_t) :: nmarkers
      type(c_ptr)        :: bc_type
      type(c_ptr)        :: bc_int
      type(c_ptr)        :: bc_real
      type(c_ptr)        :: bc_riemann
   end type hg_mesh_desc

   type, bind(C), public :: hg_physics
      real(c_double) :: mach
      real(c_double) :: reynolds
      real(c_double) :: prandtl
   end type hg_physics

   type, bind(C), public :: hg_options
      integer(c_int32_t) :: navier_stokes
      integer(c_int32_t) :: riemann
      integer(c_int32_t) :: viscous
      integer(c_int32_t) :: ip_method
      real(c_double)     :: sigma0
      integer(c_int32_t) :: precision
      integer(c_int32_t) :: scheme
      integer(c_int32_t) :: artificial_dissipation
      real(c_double)     :: ad_intensity
   end type hg_options

   ! batched form: tile -1 autotune / 0 narrow / 1 wide; order -1 autotune / 0 case-major /
   ! 1 costliest tiles first / 2 cases interleaved / 3 one launch per polynomial order;
   ! use_graph 1 = replay a captured step
   type, bind(C), public :: hg_batch_options
      integer(c_int32_t) :: tile
      integer(c_int32_t) :: order
      integer(c_int32_t) :: use_graph
   end type hg_batch_options

   public :: hg_create, hg_destroy, hg_ndof, hg_get_node_coords, hg_upload_state, hg_download_state
   public :: hg_download_qdot, hg_time_derivative, hg_rk_step, hg_residuals, hg_estimate_dt
   public :: hg_reset_dt_state, hg_conservation_sums, hg_volume_integrals, hg_surface_integrals
   public :: hg_boundary_node_count, hg_boundary_nodes, hg_set_boundary_state
   public :: hg_probes_set, hg_probes_update
   public :: hg_synchronize, hg_last_error
   public :: hg_set_tiles, hg_get_tiles, hg_set_graph, hg_profile_enable, hg_profile_read
   public :: hg_batch_create, hg_batch_rk_step, hg_batch_synchronize, hg_batch_schedule
   public :: hg_batch_profile_enable, hg_batch_profile_read, hg_batch_destroy
   public :: hgf_check

   interface
      integer(c_int) function hg_create(mesh, phys, opt, handle) bind(C, name="hg_create")
         import :: c_int, c_ptr, hg_mesh_desc, hg_physics, hg_options
         type(hg_mesh_desc), intent(in) :: mesh
         type(hg_physics),   intent(in) :: phys
         type(hg_options),   intent(in) :: opt
         type(c_ptr),        intent(out) :: handle
      end function hg_create

      integer(c_int) function hg_destroy(handle) bind(C, name="hg_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function hg_destroy

      integer(c_int64_t) function hg_ndof(handle) bind(C, name="hg_ndof")
         import :: c_int64_t, c_ptr
         type(c_ptr), value :: handle
      end function hg_ndof

      integer(c_int) function hg_get_node_coords(handle, X, nnodes) bind(C, name="hg_get_node_coords")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value         :: handle
         real(c_double)             :: X(*)
         integer(c_int64_t), value  :: nnodes
      end function hg_get_node_coords

      integer(c_int) function hg_upload_state(handle, Q, ndof) bind(C, name="hg_upload_state")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value          :: handle
         real(c_double), intent(in)  :: Q(*)
         integer(c_int64_t), value   :: ndof
      end function hg_upload_state

      integer(c_int) function hg_download_state(handle, Q, ndof) bind(C, name="hg_download_state")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value          :: handle
         real(c_double), intent(out) :: Q(*)
         integer(c_int64_t), value   :: ndof
      end function hg_download_state

      integer(c_int) function hg_download_qdot(handle, QDot, ndof) bind(C, name="hg_download_qdot")
         import :: c_int, c_ptr, c_double, c_int64_t
         type(c_ptr), value          :: handle
         real(c_double), intent(out) :: QDot(*)
         integer(c_int64_t), value   :: ndof
      end function hg_download_qdot

      integer(c_int) function hg_time_derivative(handle, t) bind(C, name="hg_time_derivative")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value    :: handle
         real(c_double), value :: t
      end function hg_time_derivative

      integer(c_int) function hg_rk_step(handle, t, dt) bind(C, name="hg_rk_step")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value    :: handle
         real(c_double), value :: t, dt
      end function hg_rk_step

      integer(c_int) function hg_residuals(handle, res) bind(C, name="hg_residuals")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: handle
         real(c_double), intent(out) :: res(4)
      end function hg_residuals

      integer(c_int) function hg_estimate_dt(handle, dt_max, cfl, dt) bind(C, name="hg_estimate_dt")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: handle
         real(c_double), value       :: dt_max, cfl
         real(c_double), intent(out) :: dt
      end function hg_estimate_dt

      integer(c_int) function hg_reset_dt_state(handle) bind(C, name="hg_reset_dt_state")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function hg_reset_dt_state

      integer(c_int) function hg_conservation_sums(handle, sums) bind(C, name="hg_conservation_sums")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: handle
         real(c_double), intent(out) :: sums(4)
      end function hg_conservation_sums

      ! QuadMesh_Integrals.incf:22-65: [One, dSnorm]
      integer(c_int) function hg_volume_integrals(handle, vals) bind(C, name="hg_volume_integrals")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: handle
         real(c_double), intent(out) :: vals(2)
      end function hg_volume_integrals

      ! QuadMesh_Integrals.incf:67-346 over zone `marker`: Surface, pressure, density,
      ! pressure force (2), mass-flow, flow, viscous force (2)
      integer(c_int) function hg_surface_integrals(handle, marker, vals) bind(C, name="hg_surface_integrals")
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value           :: handle
         integer(c_int32_t), value    :: marker
         real(c_double), intent(out)  :: vals(9)
      end function hg_surface_integrals

      ! "Use function" zones: the caller evaluates its BoundaryState function at these nodes
      ! (edge % x of the zone's edges, ascending edge id) and hands the states over
      integer(c_int) function hg_boundary_node_count(handle, marker, n) bind(C, name="hg_boundary_node_count")
         import :: c_int, c_ptr, c_int32_t, c_int64_t
         type(c_ptr), value           :: handle
         integer(c_int32_t), value    :: marker
         integer(c_int64_t), intent(out) :: n
      end function hg_boundary_node_count

      integer(c_int) function hg_boundary_nodes(handle, marker, n, x) bind(C, name="hg_boundary_nodes")
         import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
         type(c_ptr), value           :: handle
         integer(c_int32_t), value    :: marker
         integer(c_int64_t), value    :: n
         real(c_double), intent(out)  :: x(2, n)
      end function hg_boundary_nodes

      integer(c_int) function hg_set_boundary_state(handle, marker, n, q) bind(C, name="hg_set_boundary_state")
         import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
         type(c_ptr), value           :: handle
         integer(c_int32_t), value    :: marker
         integer(c_int64_t), value    :: n
         real(c_double), intent(in)   :: q(4, n)
      end function hg_set_boundary_state

      ! Probe_Initialization (IO/Monitors.f90:647-737): elements 0-based (eID - 1), reference
      ! coordinates (xi, eta), variables HG_PROBE_* (rho .. Mach in Probe_Update's order, 0-based)
      integer(c_int) function hg_probes_set(handle, n, elem, xi, eta, var) bind(C, name="hg_probes_set")
         import :: c_int, c_ptr, c_double, c_int32_t
         type(c_ptr), value              :: handle
         integer(c_int32_t), value       :: n
         integer(c_int32_t), intent(in)  :: elem(*), var(*)
         real(c_double), intent(in)      :: xi(*), eta(*)
      end function hg_probes_set

      ! Probe_Update (IO/Monitors.f90:739-846): dimensionless values, one per probe
      integer(c_int) function hg_probes_update(handle, vals) bind(C, name="hg_probes_update")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value          :: handle
         real(c_double), intent(out) :: vals(*)
      end function hg_probes_update

      integer(c_int) function hg_synchronize(handle) bind(C, name="hg_synchronize")
         import :: c_int, c_ptr
         type(c_ptr), value :: handle
      end function hg_synchronize

      integer(c_int) function hg_last_error(buf, n) bind(C, name="hg_last_error")
         import :: c_int, c_char, c_int32_t
         character(kind=c_char) :: buf(*)
         integer(c_int32_t), value :: n
      end function hg_last_error

      ! performance controls (no reference counterpart): workgroup tiles, graph replay,
      ! kernel timing (which: 0 stage kernel, 1 gradient kernel)
      integer(c_int) function hg_set_tiles(handle, stage_tile, gradient_tile) bind(C, name="hg_set_tiles")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value        :: handle
         integer(c_int32_t), value :: stage_tile, gradient_tile
      end function hg_set_tiles

      integer(c_int) function hg_get_tiles(handle, stage_tile, gradient_tile) bind(C, name="hg_get_tiles")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value              :: handle
         integer(c_int32_t), intent(out) :: stage_tile, gradient_tile
      end function hg_get_tiles

      integer(c_int) function hg_set_graph(handle, on) bind(C, name="hg_set_graph")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value        :: handle
         integer(c_int32_t), value :: on
      end function hg_set_graph

      integer(c_int) function hg_profile_enable(handle, on) bind(C, name="hg_profile_enable")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value        :: handle
         integer(c_int32_t), value :: on
      end function hg_profile_enable

      integer(c_int) function hg_profile_read(handle, which, launches, ms) bind(C, name="hg_profile_read")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
         type(c_ptr), value              :: handle
         integer(c_int32_t), value       :: which
         integer(c_int64_t), intent(out) :: launches
         real(c_double), intent(out)     :: ms
      end function hg_profile_read

      ! batched form: many handles (own mesh, order, dt) advanced by one launch per RK stage;
      ! cases(:) holds the handles (they must outlive the batch)
      integer(c_int) function hg_batch_create(cases, ncases, opt, batch) bind(C, name="hg_batch_create")
         import :: c_int, c_ptr, c_int32_t, hg_batch_options
         type(c_ptr), intent(in)            :: cases(*)
         integer(c_int32_t), value          :: ncases
         type(hg_batch_options), intent(in) :: opt
         type(c_ptr), intent(out)           :: batch
      end function hg_batch_create

      ! one RK step of every case: t(c), dt(c) per case
      integer(c_int) function hg_batch_rk_step(batch, t, dt) bind(C, name="hg_batch_rk_step")
         import :: c_int, c_ptr, c_double
         type(c_ptr), value         :: batch
         real(c_double), intent(in) :: t(*), dt(*)
      end function hg_batch_rk_step

      integer(c_int) function hg_batch_synchronize(batch) bind(C, name="hg_batch_synchronize")
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
      end function hg_batch_synchronize

      integer(c_int) function hg_batch_schedule(batch, tile, order) bind(C, name="hg_batch_schedule")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value              :: batch
         integer(c_int32_t), intent(out) :: tile, order
      end function hg_batch_schedule

      integer(c_int) function hg_batch_profile_enable(batch, on) bind(C, name="hg_batch_profile_enable")
         import :: c_int, c_ptr, c_int32_t
         type(c_ptr), value        :: batch
         integer(c_int32_t), value :: on
      end function hg_batch_profile_enable

      integer(c_int) function hg_batch_profile_read(batch, which, launches, ms) bind(C, name="hg_batch_profile_read")
         import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
         type(c_ptr), value              :: batch
         integer(c_int32_t), value       :: which
         integer(c_int64_t), intent(out) :: launches
         real(c_double), intent(out)     :: ms
      end function hg_batch_profile_read

      integer(c_int) function hg_batch_destroy(batch) bind(C, name="hg_batch_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: batch
      end function hg_batch_destroy
   end interface

contains

   ! Maps a nonzero status to the solver's usual `stop "..."` (Generic/Defines.h:156 style),
   ! printing the library's message; the C side itself never exits.
   subroutine hgf_check(rc, what)
      integer(c_int), intent(in)   :: rc
      character(len=*), intent(in) :: what
      character(kind=c_char)       :: buf(512)
      character(len=512)           :: msg
      integer                      :: i, ierr
      if (rc == 0) return
      ierr = hg_last_error(buf, 512_c_int32_t)
      msg = ""
      do i = 1, 512
         if (buf(i) == c_null_char) exit
         msg(i:i) = buf(i)
      end do
      write(*, '(A,A,A,A)') "HORSES2D GPU: ", trim(what), " failed: ", trim(msg)
      error stop "Stopped."
   end subroutine hgf_check

end module horses_gpu_bindings
