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
     ! level 2 of a two-level hierarchy (nlevel = 2, refinement ratio 2): its
     ! boxes, in level-2 cell indices; its domain is the level-1 one refined
     integer :: nboxes2 = 0
     integer, allocatable :: lo2(:,:), hi2(:,:)
     type(c_ptr) :: plan = c_null_ptr
     integer :: plan_ng(5) = -1
     type(c_ptr) :: mlplan = c_null_ptr
     integer :: mlplan_ng(5) = -1
  end type ml_layout

  type bc_tower
     integer :: phys_bc(3,2) = PERIODIC       ! (direction, lo/hi)
  end type bc_tower

  ! FBoxLib's flux register argument of mac_multigrid; a single-level solve
  ! leaves it empty (the single-level callers discard it,
  ! Source/thermal_conduct.f90:366-368)
  type bndry_reg
     integer :: unused = 0
  end type bndry_reg

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

  ! multifab_build on level lev (1 or 2) of the layout
  subroutine multifab_build_level(mf, mla, lev, ng, nodal_dir)
    type(multifab), intent(inout) :: mf
    type(ml_layout), intent(in) :: mla
    integer, intent(in) :: lev, ng, nodal_dir
    integer :: b, l(3), h(3), d, nb
    if (lev == 1) then
       call multifab_build(mf, mla, ng, nodal_dir)
       return
    end if
    nb = mla%nboxes2
    mf%ng = ng
    mf%nodal_dir = nodal_dir
    if (allocated(mf%fabs)) deallocate(mf%fabs)
    allocate(mf%fabs(nb))
    do b = 1, nb
       l = 0
       h = 0
       do d = 1, mla%dim
          l(d) = mla%lo2(d,b) - ng
          h(d) = mla%hi2(d,b) + ng
          if (d == nodal_dir) h(d) = h(d) + 1
       end do
       allocate(mf%fabs(b)%p(l(1):h(1), l(2):h(2), l(3):h(3)))
       mf%fabs(b)%p = 0.0_dp_t
    end do
  end subroutine multifab_build_level

end module mac_types_module


module macproject_module

  use iso_c_binding
  use mac_types_module
  implicit none

  private
  public :: macproject, macproject_release, mac_plan_ensure, fail
  public :: use_alt_energy_fix, mac_precision, mg_verbose, eps_mac, eps_mac_max, &
            mac_level_factor, mg_cycle_type, mg_bottom_nu, mg_bottom_solver, mac_omega, last_stats, &
            c_mac_stats, c_mac_opts, mac_default_opts

  ! runtime knobs (Source/_parameters:370-392,552-554, Source/mg_eps.f90:22-28)
  logical, save :: use_alt_energy_fix = .true.
  integer, save :: mac_precision = 0            ! 0 FP64 V-cycle, 1 FP32 V-cycle + FP64 refinement,
                                                ! 2 as 1 with bfloat16 finest-level weights and rhs
  integer, save :: mg_verbose = 0
  integer, save :: mg_cycle_type = 3            ! 1 F-cycle, 2 W-cycle, 3 V-cycle
  integer, save :: mg_bottom_nu = 10            ! sweeps around the bottom solve
  integer, save :: mg_bottom_solver = -1        ! 0 smoothing only, 1 BiCGStab, 2 CG; -1 and 4
                                                ! (agglomerating, always on here): the library's CG
  real(dp_t), save :: mac_omega = 0.d0          ! over-relaxation of the red-black sweeps (no
                                                ! reference counterpart); 0: 1, Gauss-Seidel
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
                       inner_cycles, device_ptrs, verbose, cycle_type, bottom_solver
     real(c_double) :: omega
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
     integer(c_int) function mac_ml_plan_create(nlevs, lays, o, plan) bind(C, name="mac_ml_plan_create")
       import :: c_int, c_ptr, c_mac_layout, c_mac_opts
       integer(c_int), value :: nlevs
       type(c_mac_layout), intent(in) :: lays(*)
       type(c_mac_opts), intent(in) :: o
       type(c_ptr), intent(out) :: plan
     end function mac_ml_plan_create
     integer(c_int) function mac_ml_plan_destroy(plan) bind(C, name="mac_ml_plan_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: plan
     end function mac_ml_plan_destroy
     integer(c_int) function mac_ml_plan_set_opts(plan, o) bind(C, name="mac_ml_plan_set_opts")
       import :: c_int, c_ptr, c_mac_opts
       type(c_ptr), value :: plan
       type(c_mac_opts), intent(in) :: o
     end function mac_ml_plan_set_opts
     integer(c_int) function mac_ml_project(plan, umac, phi, rho, macrhs, b0, b0e, st) &
          bind(C, name="mac_ml_project")
       import :: c_int, c_ptr, c_mac_stats
       type(c_ptr), value :: plan, umac, phi, rho, macrhs, b0, b0e
       type(c_mac_stats), intent(out) :: st
     end function mac_ml_project
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

    type(c_mac_opts) :: o
    type(c_mac_stats) :: st
    type(c_ptr), allocatable, target :: pu(:), pp(:), pr(:), pm(:), pc(:)
    real(dp_t), allocatable, target :: b0(:), b0e(:)
    type(c_ptr) :: cm, cc, cb0, cb0e
    integer :: nlevs, dm, nb, b, d, rc, ng(5), nr

    nlevs = size(phi)
    dm = mla%dim
    nb = mla%nboxes

    if (nlevs == 2) then
       call macproject_2lev(mla, umac, phi, rho, dx, the_bc_tower, macrhs, beta0_1d, beta0_1d_edge, &
                            beta0_cart_edge)
       return
    end if
    if (nlevs /= 1) call fail('one or two levels are supported, nlevs = ', nlevs)
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
    o%cycle_type = mg_cycle_type
    o%nub = mg_bottom_nu
    if (mg_bottom_solver /= -1 .and. mg_bottom_solver /= 4) o%bottom_solver = mg_bottom_solver
    o%omega = mac_omega
    o%device_ptrs = 0

    call mac_plan_ensure(mla, dx, the_bc_tower, ng, o)

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

  ! The device plan of the layout, built on first use (with the ghost widths
  ! ng = [umac, phi, rho, macrhs, beta0_cart_edge] of the projection's
  ! arrays) and cached in the ml_layout; otherwise the options o are set on
  ! the cached plan.
  subroutine mac_plan_ensure(mla, dx, the_bc_tower, ng, o)
    type(ml_layout), intent(inout) :: mla
    real(dp_t), intent(in) :: dx(:,:)
    type(bc_tower), intent(in) :: the_bc_tower
    integer, intent(in) :: ng(5)
    type(c_mac_opts), intent(in) :: o
    type(c_mac_layout) :: lay
    integer(c_int), allocatable, target :: blo(:), bhi(:)
    integer :: dm, nb, b, d, rc
    dm = mla%dim
    nb = mla%nboxes
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
  end subroutine mac_plan_ensure

  subroutine macproject_release(mla)
    type(ml_layout), intent(inout) :: mla
    integer :: rc
    if (c_associated(mla%plan)) rc = mac_plan_destroy(mla%plan)
    mla%plan = c_null_ptr
    mla%plan_ng = -1
    if (c_associated(mla%mlplan)) rc = mac_ml_plan_destroy(mla%mlplan)
    mla%mlplan = c_null_ptr
    mla%mlplan_ng = -1
  end subroutine macproject_release

  ! macproject with nlevs = 2 (Source/macproject.f90:51-328, the nlevs > 1
  ! branches) through mac_ml_project: level 1 on mla's boxes, level 2 on its
  ! level-2 boxes; dx(n,:) and beta0_1d(n,0:nr_n-1) per level (nr_2 = 2 nr_1)
  subroutine macproject_2lev(mla, umac, phi, rho, dx, the_bc_tower, &
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

    type(c_mac_opts) :: o
    type(c_mac_stats) :: st
    type(c_mac_layout), target :: lays(2)
    integer(c_int), allocatable, target :: blo1(:), bhi1(:), blo2(:), bhi2(:)
    type(c_ptr), allocatable, target :: pu1(:), pu2(:), pp1(:), pp2(:), pr1(:), pr2(:), pm1(:), pm2(:)
    type(c_ptr), target :: tu(2), tp(2), tr(2), tm(2), tb(2), tbe(2)
    real(dp_t), allocatable, target :: b01(:), b0e1(:), b02(:), b0e2(:)
    type(c_ptr) :: cm, cb0, cb0e
    integer :: dm, nb1, nb2, b, d, rc, ng(5), nr, lev

    dm = mla%dim
    nb1 = mla%nboxes
    nb2 = mla%nboxes2
    if (mla%nlevel /= 2 .or. nb2 < 1) call fail('two-level projection without level-2 boxes, nlevel = ', &
                                                 mla%nlevel)
    if (present(beta0_cart_edge)) call fail('beta0_cart_edge with two levels is not supported', 2)
    if (present(beta0_1d) .neqv. present(beta0_1d_edge)) &
         call fail('beta0_1d and beta0_1d_edge go together', 1)

    ng = [umac(1,1)%ng, phi(1)%ng, rho(1)%ng, 0, 0]
    if (present(macrhs)) ng(4) = macrhs(1)%ng

    rc = mac_default_opts(o)
    o%eps_rel = min(eps_mac * mac_level_factor**(2-1), eps_mac_max)
    o%eps_abs = -1.0_dp_t
    o%use_alt_energy_fix = merge(1, 0, use_alt_energy_fix)
    o%precision = mac_precision
    o%verbose = mg_verbose
    o%cycle_type = mg_cycle_type
    o%nub = mg_bottom_nu
    if (mg_bottom_solver /= -1 .and. mg_bottom_solver /= 4) o%bottom_solver = mg_bottom_solver
    o%omega = mac_omega
    o%device_ptrs = 0

    if (.not. c_associated(mla%mlplan) .or. any(mla%mlplan_ng /= ng)) then
       if (c_associated(mla%mlplan)) rc = mac_ml_plan_destroy(mla%mlplan)
       allocate(blo1(3*nb1), bhi1(3*nb1), blo2(3*nb2), bhi2(3*nb2))
       do b = 1, nb1
          blo1(3*(b-1)+1:3*b) = mla%lo(:,b)
          bhi1(3*(b-1)+1:3*b) = mla%hi(:,b)
       end do
       do b = 1, nb2
          blo2(3*(b-1)+1:3*b) = mla%lo2(:,b)
          bhi2(3*(b-1)+1:3*b) = mla%hi2(:,b)
       end do
       do lev = 1, 2
          lays(lev)%dim = dm
          lays(lev)%dx = 1.0_dp_t
          lays(lev)%dx(1:dm) = dx(lev,1:dm)
          do d = 1, 3
             lays(lev)%phys_bc(1,d) = the_bc_tower%phys_bc(d,1)
             lays(lev)%phys_bc(2,d) = the_bc_tower%phys_bc(d,2)
          end do
          lays(lev)%domain_lo = 0
          lays(lev)%ng_umac = ng(1)
          lays(lev)%ng_phi = ng(2)
          lays(lev)%ng_rho = ng(3)
          lays(lev)%ng_macrhs = ng(4)
          lays(lev)%ng_beta0_cart = 0
       end do
       lays(1)%nboxes = nb1
       lays(1)%box_lo = c_loc(blo1)
       lays(1)%box_hi = c_loc(bhi1)
       lays(1)%domain_hi = mla%pd_hi
       lays(2)%nboxes = nb2
       lays(2)%box_lo = c_loc(blo2)
       lays(2)%box_hi = c_loc(bhi2)
       lays(2)%domain_hi = 0
       lays(2)%domain_hi(1:dm) = 2 * mla%pd_hi(1:dm) + 1
       rc = mac_ml_plan_create(2_c_int, lays, o, mla%mlplan)
       if (rc /= 0) call fail('mac_ml_plan_create failed, code ', rc)
       mla%mlplan_ng = ng
    else
       rc = mac_ml_plan_set_opts(mla%mlplan, o)
       if (rc /= 0) call fail('mac_ml_plan_set_opts failed, code ', rc)
    end if

    allocate(pu1(nb1*dm), pp1(nb1), pr1(nb1), pu2(nb2*dm), pp2(nb2), pr2(nb2))
    do b = 1, nb1
       do d = 1, dm
          pu1((b-1)*dm + d) = c_loc(umac(1,d)%fabs(b)%p)
       end do
       pp1(b) = c_loc(phi(1)%fabs(b)%p)
       pr1(b) = c_loc(rho(1)%fabs(b)%p)
    end do
    do b = 1, nb2
       do d = 1, dm
          pu2((b-1)*dm + d) = c_loc(umac(2,d)%fabs(b)%p)
       end do
       pp2(b) = c_loc(phi(2)%fabs(b)%p)
       pr2(b) = c_loc(rho(2)%fabs(b)%p)
    end do
    tu = [c_loc(pu1), c_loc(pu2)]
    tp = [c_loc(pp1), c_loc(pp2)]
    tr = [c_loc(pr1), c_loc(pr2)]
    cm = c_null_ptr
    if (present(macrhs)) then
       allocate(pm1(nb1), pm2(nb2))
       do b = 1, nb1
          pm1(b) = c_loc(macrhs(1)%fabs(b)%p)
       end do
       do b = 1, nb2
          pm2(b) = c_loc(macrhs(2)%fabs(b)%p)
       end do
       tm = [c_loc(pm1), c_loc(pm2)]
       cm = c_loc(tm)
    end if
    cb0 = c_null_ptr
    cb0e = c_null_ptr
    if (present(beta0_1d)) then
       nr = mla%pd_hi(dm) - mla%pd_lo(dm) + 1
       allocate(b01(0:nr-1), b0e1(0:nr), b02(0:2*nr-1), b0e2(0:2*nr))
       b01 = beta0_1d(1,0:nr-1)
       b0e1 = beta0_1d_edge(1,0:nr)
       b02 = beta0_1d(2,0:2*nr-1)
       b0e2 = beta0_1d_edge(2,0:2*nr)
       tb = [c_loc(b01), c_loc(b02)]
       tbe = [c_loc(b0e1), c_loc(b0e2)]
       cb0 = c_loc(tb)
       cb0e = c_loc(tbe)
    end if

    rc = mac_ml_project(mla%mlplan, c_loc(tu), c_loc(tp), c_loc(tr), cm, cb0, cb0e, st)
    last_stats = st
    if (mg_verbose > 0) then
       write(*,'(a,es15.8)') 'F90mg: Initial rhs                  = ', st%rhs_norm
       write(*,'(a,es15.8)') 'F90mg: Initial error (error0)       = ', st%res0
       write(*,'(a,i4,a,es15.8)') 'F90mg: Final Iter. ', st%iters, ' error/error0 = ', &
            st%res / max(st%res0, tiny(1.0_dp_t))
    end if
    if (rc /= 0) call fail('mac_ml_project failed, code ', rc)

  end subroutine macproject_2lev

end module macproject_module


! The solver-level drop-in: FBoxLib's cell-centred solve as MAESTRO calls it,
!   mac_multigrid(mla, rh, phi, fine_flx, solver_alpha, solver_beta, dx,
!                 the_bc_tower, bc_comp, stencil_order, rel_solver_eps,
!                 abs_solver_eps)
! (Source/mac_multigrid.f90:18-71): (alpha - D beta G) phi = rh on the
! library's mac_solve.  The elliptic BCs of component bc_comp follow
! ell_bc_level_build (Source/define_bc_tower.f90:296-402); the caller names
! its press_comp, foextrap_comp and hoextrap_comp indices (every other
! component above dm is a scalar: density, (rho h), species, temperature, pi,
! tracers).  Dirichlet values come from phi's ghost cells.
module mac_multigrid_module

  use iso_c_binding
  use mac_types_module
  use macproject_module, only: mac_plan_ensure, fail, c_mac_opts, c_mac_stats, mac_default_opts, &
                               mac_precision, mg_verbose, mg_cycle_type, mg_bottom_nu, &
                               mg_bottom_solver, mac_omega, last_stats
  implicit none

  private
  public :: mac_multigrid, press_comp, foextrap_comp, hoextrap_comp

  integer, save :: press_comp = 0, foextrap_comp = 0, hoextrap_comp = 0

  type, bind(C) :: c_mac_solve_in
     type(c_ptr) :: rh
     integer(c_int) :: ng_rh
     type(c_ptr) :: phi
     integer(c_int) :: ng_phi
     type(c_ptr) :: alpha
     integer(c_int) :: ng_alpha
     type(c_ptr) :: beta
     integer(c_int) :: ng_beta
     integer(c_int) :: ell_bc(2,3)
     real(c_double) :: eps_rel, eps_abs
  end type c_mac_solve_in

  interface
     integer(c_int) function mac_solve(plan, in, st) bind(C, name="mac_solve")
       import :: c_int, c_ptr, c_mac_solve_in, c_mac_stats
       type(c_ptr), value :: plan
       type(c_mac_solve_in), intent(in) :: in
       type(c_mac_stats), intent(out) :: st
     end function mac_solve
     integer(c_int) function mac_ell_bc(phys, comp) bind(C, name="mac_ell_bc")
       import :: c_int
       integer(c_int), value :: phys, comp
     end function mac_ell_bc
  end interface

contains

  subroutine mac_multigrid(mla, rh, phi, fine_flx, solver_alpha, solver_beta, dx, the_bc_tower, &
                           bc_comp, stencil_order, rel_solver_eps, abs_solver_eps)

    type(ml_layout), intent(inout) :: mla
    type(multifab), intent(inout), target :: rh(:), phi(:)
    type(bndry_reg), intent(inout) :: fine_flx(:)
    type(multifab), intent(in), target :: solver_alpha(:), solver_beta(:,:)
    real(dp_t), intent(in) :: dx(:,:)
    type(bc_tower), intent(in) :: the_bc_tower
    integer, intent(in) :: bc_comp, stencil_order
    real(dp_t), intent(in) :: rel_solver_eps, abs_solver_eps

    type(c_mac_opts) :: o
    type(c_mac_stats) :: st
    type(c_mac_solve_in) :: a
    type(c_ptr), allocatable, target :: pr(:), pp(:), pa(:), pb(:)
    integer :: dm, nb, b, d, s, rc, cls, ng(5)

    dm = mla%dim
    nb = mla%nboxes
    if (size(phi) /= 1) call fail('only single-level solves are supported, nlevs = ', size(phi))
    if (stencil_order /= 2) call fail('only stencil_order = 2 is supported, got ', stencil_order)
    if (bc_comp == press_comp) then
       cls = 0
    else if (bc_comp == foextrap_comp .or. bc_comp == hoextrap_comp) then
       cls = 2
    else if (bc_comp > dm) then
       cls = 1
    else
       call fail('velocity components have no solver BCs here, bc_comp = ', bc_comp)
    end if

    rc = mac_default_opts(o)
    o%precision = mac_precision
    o%verbose = mg_verbose
    o%cycle_type = mg_cycle_type
    o%nub = mg_bottom_nu
    if (mg_bottom_solver /= -1 .and. mg_bottom_solver /= 4) o%bottom_solver = mg_bottom_solver
    o%omega = mac_omega
    o%device_ptrs = 0
    ! a cached projection plan serves as it is; a new one gets the projection's usual widths
    if (c_associated(mla%plan)) then
       ng = mla%plan_ng
    else
       ng = [1, phi(1)%ng, 1, 0, 1]
    end if
    call mac_plan_ensure(mla, dx, the_bc_tower, ng, o)

    allocate(pr(nb), pp(nb), pa(nb), pb(nb*dm))
    do b = 1, nb
       pr(b) = c_loc(rh(1)%fabs(b)%p)
       pp(b) = c_loc(phi(1)%fabs(b)%p)
       pa(b) = c_loc(solver_alpha(1)%fabs(b)%p)
       do d = 1, dm
          pb((b-1)*dm + d) = c_loc(solver_beta(1,d)%fabs(b)%p)
       end do
    end do
    a%rh = c_loc(pr)
    a%ng_rh = rh(1)%ng
    a%phi = c_loc(pp)
    a%ng_phi = phi(1)%ng
    a%alpha = c_loc(pa)
    a%ng_alpha = solver_alpha(1)%ng
    a%beta = c_loc(pb)
    a%ng_beta = solver_beta(1,1)%ng
    a%ell_bc = -1
    do d = 1, dm
       do s = 1, 2
          a%ell_bc(s,d) = mac_ell_bc(the_bc_tower%phys_bc(d,s), cls)
       end do
    end do
    a%eps_rel = rel_solver_eps
    a%eps_abs = abs_solver_eps

    rc = mac_solve(mla%plan, a, st)
    last_stats = st
    if (mg_verbose > 0) then
       write(*,'(a,es15.8)') 'F90mg: Initial rhs                  = ', st%rhs_norm
       write(*,'(a,es15.8)') 'F90mg: Initial error (error0)       = ', st%res0
       write(*,'(a,i4,a,es15.8)') 'F90mg: Final Iter. ', st%iters, ' error/error0 = ', &
            st%res / max(st%res0, tiny(1.0_dp_t))
    end if
    if (rc /= 0) call fail('mac_solve failed, code ', rc)

  end subroutine mac_multigrid

end module mac_multigrid_module
