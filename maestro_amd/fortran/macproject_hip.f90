! macproject_hip.f90 -- Fortran drop-in for MAESTRO's MAC projection on the
! MI355X library libmaestro_mac.so (C ABI: include/maestro_mac.h).
!
! The module keeps the reference's public name and argument list,
!   macproject(mla, umac, phi, rho, dx, the_bc_tower
!              [, macrhs][, beta0_1d, beta0_1d_edge][, beta0_cart_edge])
! (Source/macproject.f90:51-76), over the build's own per-box containers in
! mac_types_module (FBoxLib's multifab/ml_layout/bc_tower are not part of
! this build).  Each fab holds one box with its ghost cells, indexed by global
! cell indices exactly like an FBoxLib fab:
!   cell centred  p(lo1-ng:hi1+ng, lo2-ng:hi2+ng, lo3-ng:hi3+ng)
!   edge array    the same with hi(d)+1+ng in its normal direction d
! (in 2D the third extent is 0:0).  The device plan is built on the first
! call and cached in the ml_layout, as the reference calls macproject twice
! per time step (Source/advance.f90:466,937).
module mac_types_module

  use iso_c_binding
  implicit none

  integer, parameter :: dp_t = c_double

  ! physical boundary codes (FBoxLib bc_module values)
  integer, parameter :: PERIODIC = -1, INTERIOR = 0, INLET = 11, OUTLET = 12, &
                        SYMMETRY = 13, SLIP_WALL = 14, NO_SLIP_WALL = 15

  type fab
     real(dp_t), allocatable :: p(:,:,:)
  end type fab

  type multifab
     integer :: ng = 0
     integer :: nodal_dir = 0          ! 0: cell centred; d: edge array normal to d
     type(fab), allocatable :: fabs(:)
  end type multifab

  type ml_layout
     integer :: dim = 0
     integer :: nlevel = 1
     integer :: nboxes = 0
     integer, allocatable :: lo(:,:), hi(:,:)   ! (3, nboxes), inclusive, 0-based
     integer :: pd_lo(3) = 0, pd_hi(3) = 0     ! problem domain
     type(c_ptr) :: plan = c_null_ptr
     integer :: plan_ng(5) = -1
  end type ml_layout

  type bc_tower
     integer :: phys_bc(3,2) = PERIODIC       ! (direction, lo/hi)
  end type bc_tower

contains

  subroutine multifab_build(mf, mla, ng, nodal_dir)
    type(multifab), intent(inout) :: mf
    type(ml_layout), intent(in) :: mla
    integer, intent(in) :: ng, nodal_dir
    integer :: b, l(3), h(3), d
    mf%ng = ng
    mf%nodal_dir = nodal_dir
    if (allocated(mf%fabs)) deallocate(mf%fabs)
    allocate(mf%fabs(mla%nboxes))
    do b = 1, mla%nboxes
       l = 0
       h = 0
       do d = 1, mla%dim
          l(d) = mla%lo(d,b) - ng
          h(d) = mla%hi(d,b) + ng
          if (d == nodal_dir) h(d) = h(d) + 1
       end do
       allocate(mf%fabs(b)%p(l(1):h(1), l(2):h(2), l(3):h(3)))
       mf%fabs(b)%p = 0.0_dp_t
    end do
  end subroutine multifab_build

end module mac_types_module


module macproject_module

  use iso_c_binding
  use mac_types_module
  implicit none

  private
  public :: macproject, macproject_release
  public :: use_alt_energy_fix, mac_precision, mg_verbose, eps_mac, eps_mac_max, &
            mac_level_factor, last_stats, c_mac_stats

  ! runtime knobs (Source/_parameters:552-554, Source/mg_eps.f90:22-28)
  logical, save :: use_alt_energy_fix = .true.
  integer, save :: mac_precision = 0            ! 0 FP64 V-cycle, 1 FP32 V-cycle + FP64 refinement,
                                                ! 2 as 1 with bfloat16 finest-level weights and rhs
  integer, save :: mg_verbose = 0
  real(dp_t), save :: eps_mac = 1.d-10
  real(dp_t), save :: eps_mac_max = 1.d-8
  real(dp_t), save :: mac_level_factor = 10.d0

  type, bind(C) :: c_mac_layout
     integer(c_int) :: dim, nboxes
     type(c_ptr) :: box_lo, box_hi
     integer(c_int) :: domain_lo(3), domain_hi(3)
     real(c_double) :: dx(3)
     integer(c_int) :: phys_bc(2,3)
     integer(c_int) :: ng_umac, ng_phi, ng_rho, ng_macrhs, ng_beta0_cart
  end type c_mac_layout

  type, bind(C) :: c_mac_opts
     real(c_double) :: eps_rel, eps_abs
     integer(c_int) :: max_iter, nu1, nu2, nub
     real(c_double) :: bottom_eps
     integer(c_int) :: bottom_max_iter, max_levels, use_alt_energy_fix, precision, &
                       inner_cycles, device_ptrs, verbose
  end type c_mac_opts

  type, bind(C) :: c_mac_stats
     integer(c_int) :: iters, inner_cycles, converged, nlevels
     real(c_double) :: rhs_norm, res0, res, t_h2d_ms, t_setup_ms, t_solve_ms, &
                       t_correct_ms, t_d2h_ms, t_total_ms
  end type c_mac_stats

  type(c_mac_stats), save :: last_stats

  interface
     integer(c_int) function mac_default_opts(o) bind(C, name="mac_default_opts")
       import :: c_int, c_mac_opts
       type(c_mac_opts), intent(out) :: o
     end function mac_default_opts
     integer(c_int) function mac_plan_create(lay, o, plan) bind(C, name="mac_plan_create")
       import :: c_int, c_ptr, c_mac_layout, c_mac_opts
       type(c_mac_layout), intent(in) :: lay
       type(c_mac_opts), intent(in) :: o
       type(c_ptr), intent(out) :: plan
     end function mac_plan_create
     integer(c_int) function mac_plan_destroy(plan) bind(C, name="mac_plan_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: plan
     end function mac_plan_destroy
     integer(c_int) function mac_plan_set_opts(plan, o) bind(C, name="mac_plan_set_opts")
       import :: c_int, c_ptr, c_mac_opts
       type(c_ptr), value :: plan
       type(c_mac_opts), intent(in) :: o
     end function mac_plan_set_opts
     integer(c_int) function mac_project(plan, umac, phi, rho, macrhs, b0, b0e, cart, st) &
          bind(C, name="mac_project")
       import :: c_int, c_ptr, c_mac_stats
       type(c_ptr), value :: plan, umac, phi, rho, macrhs, b0, b0e, cart
       type(c_mac_stats), intent(out) :: st
     end function mac_project
  end interface

contains

  subroutine fail(msg, rc)
    character(len=*), intent(in) :: msg
    integer, intent(in) :: rc
    write(*,'(a,a,i0,a)') 'macproject: ', msg, rc, ''
    error stop 1
  end subroutine fail

  ! NOTE: as in the reference, phi is passed in (warm start) and out.
  subroutine macproject(mla, umac, phi, rho, dx, the_bc_tower, &
                        macrhs, beta0_1d, beta0_1d_edge, beta0_cart_edge)

    type(ml_layout), intent(inout) :: mla
    type(multifab), intent(inout), target :: umac(:,:)
    type(multifab), intent(inout), target :: phi(:)
    type(multifab), intent(in), target :: rho(:)
    real(dp_t), intent(in) :: dx(:,:)
    type(bc_tower), intent(in) :: the_bc_tower
    type(multifab), intent(in), optional, target :: macrhs(:)
    real(dp_t), intent(in), optional, target :: beta0_1d(:,0:)
    real(dp_t), intent(in), optional, target :: beta0_1d_edge(:,0:)
    type(multifab), intent(in), optional, target :: beta0_cart_edge(:,:)

    type(c_mac_layout) :: lay
    type(c_mac_opts) :: o
    type(c_mac_stats) :: st
    integer(c_int), allocatable, target :: blo(:), bhi(:)
    type(c_ptr), allocatable, target :: pu(:), pp(:), pr(:), pm(:), pc(:)
    real(dp_t), allocatable, target :: b0(:), b0e(:)
    type(c_ptr) :: cm, cc, cb0, cb0e
    integer :: nlevs, dm, nb, b, d, rc, ng(5), nr

    nlevs = size(phi)
    dm = mla%dim
    nb = mla%nboxes

    if (nlevs /= 1) call fail('only single-level projections are supported, nlevs = ', nlevs)
    ! option combinations, Source/macproject.f90:99-109
    if (present(beta0_1d) .and. present(beta0_cart_edge)) &
         call fail('CANT HAVE 1D and 3D DIV_COEFF IN MACPROJECT', 1)
    if (present(beta0_1d) .neqv. present(beta0_1d_edge)) &
         call fail('beta0_1d and beta0_1d_edge go together', 1)

    ng = [umac(1,1)%ng, phi(1)%ng, rho(1)%ng, 0, 0]
    if (present(macrhs)) ng(4) = macrhs(1)%ng
    if (present(beta0_cart_edge)) ng(5) = beta0_cart_edge(1,1)%ng

    ! the library's options: tolerances as the reference sets them
    rc = mac_default_opts(o)
    o%eps_rel = min(eps_mac * mac_level_factor**(nlevs-1), eps_mac_max)
    o%eps_abs = -1.0_dp_t
    o%use_alt_energy_fix = merge(1, 0, use_alt_energy_fix)
    o%precision = mac_precision
    o%verbose = mg_verbose
    o%device_ptrs = 0

    if (.not. c_associated(mla%plan) .or. any(mla%plan_ng /= ng)) then
       if (c_associated(mla%plan)) rc = mac_plan_destroy(mla%plan)
       allocate(blo(3*nb), bhi(3*nb))
       do b = 1, nb
          blo(3*(b-1)+1:3*b) = mla%lo(:,b)
          bhi(3*(b-1)+1:3*b) = mla%hi(:,b)
       end do
       lay%dim = dm
       lay%nboxes = nb
       lay%box_lo = c_loc(blo)
       lay%box_hi = c_loc(bhi)
       lay%domain_lo = mla%pd_lo
       lay%domain_hi = mla%pd_hi
       lay%dx = 1.0_dp_t
       lay%dx(1:dm) = dx(1,1:dm)
       do d = 1, 3
          lay%phys_bc(1,d) = the_bc_tower%phys_bc(d,1)
          lay%phys_bc(2,d) = the_bc_tower%phys_bc(d,2)
       end do
       lay%ng_umac = ng(1)
       lay%ng_phi = ng(2)
       lay%ng_rho = ng(3)
       lay%ng_macrhs = ng(4)
       lay%ng_beta0_cart = ng(5)
       rc = mac_plan_create(lay, o, mla%plan)
       if (rc /= 0) call fail('mac_plan_create failed, code ', rc)
       mla%plan_ng = ng
    else
       rc = mac_plan_set_opts(mla%plan, o)
       if (rc /= 0) call fail('mac_plan_set_opts failed, code ', rc)
    end if

    ! box-major pointer tables: umac(b*dm + d)
    allocate(pu(nb*dm), pp(nb), pr(nb))
    do b = 1, nb
       do d = 1, dm
          pu((b-1)*dm + d) = c_loc(umac(1,d)%fabs(b)%p)
       end do
       pp(b) = c_loc(phi(1)%fabs(b)%p)
       pr(b) = c_loc(rho(1)%fabs(b)%p)
    end do
    cm = c_null_ptr
    if (present(macrhs)) then
       allocate(pm(nb))
       do b = 1, nb
          pm(b) = c_loc(macrhs(1)%fabs(b)%p)
       end do
       cm = c_loc(pm)
    end if
    cc = c_null_ptr
    if (present(beta0_cart_edge)) then
       allocate(pc(nb*dm))
       do b = 1, nb
          do d = 1, dm
             pc((b-1)*dm + d) = c_loc(beta0_cart_edge(1,d)%fabs(b)%p)
          end do
       end do
       cc = c_loc(pc)
    end if
    cb0 = c_null_ptr
    cb0e = c_null_ptr
    if (present(beta0_1d)) then
       nr = mla%pd_hi(dm) - mla%pd_lo(dm) + 1
       allocate(b0(0:nr-1), b0e(0:nr))
       b0 = beta0_1d(1,0:nr-1)
       b0e = beta0_1d_edge(1,0:nr)
       cb0 = c_loc(b0)
       cb0e = c_loc(b0e)
    end if

    rc = mac_project(mla%plan, c_loc(pu), c_loc(pp), c_loc(pr), cm, cb0, cb0e, cc, st)
    last_stats = st
    if (mg_verbose > 0) then
       write(*,'(a,es15.8)') 'F90mg: Initial rhs                  = ', st%rhs_norm
       write(*,'(a,es15.8)') 'F90mg: Initial error (error0)       = ', st%res0
       write(*,'(a,i4,a,es15.8)') 'F90mg: Final Iter. ', st%iters, ' error/error0 = ', &
            st%res / max(st%res0, tiny(1.0_dp_t))
    end if
    if (rc /= 0) call fail('mac_project failed, code ', rc)

  end subroutine macproject

  subroutine macproject_release(mla)
    type(ml_layout), intent(inout) :: mla
    integer :: rc
    if (c_associated(mla%plan)) rc = mac_plan_destroy(mla%plan)
    mla%plan = c_null_ptr
    mla%plan_ng = -1
  end subroutine macproject_release

end module macproject_module
