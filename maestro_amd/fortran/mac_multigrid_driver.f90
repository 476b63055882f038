! mac_multigrid_driver.f90 -- standalone Fortran caller of the drop-in
! mac_multigrid (macproject_hip.f90), as MAESTRO's thermal diffusion calls it
! (Source/thermal_conduct.f90:371-372).  It reads one problem from a binary
! stream file, solves (alpha - D beta G) phi = rh and writes phi.
!
! Input (stream, native endianness; arrays per box in Fortran order with
! their ghost cells):
!   int32  dim, nboxes, ng_rh, ng_phi, ng_alpha, ng_beta, precision, bc_comp
!   int32  domain_hi(3), phys_bc(3,2) (direction fastest)
!   int32  box lo(3), hi(3) for every box
!   real64 dx(3), rel_solver_eps, abs_solver_eps
!   per box: rh, phi, alpha, beta(d) for d = 1..dim
! Component indices of this driver: press_comp = 100, foextrap_comp = 101,
! hoextrap_comp = 102; any other bc_comp above dim is a scalar.
! Output: per box phi; then int32 iters, converged; real64 res0, res.
program mac_multigrid_driver

  use iso_c_binding
  use mac_types_module
  use macproject_module, only: mac_precision, last_stats, macproject_release
  use mac_multigrid_module
  implicit none

  type(ml_layout) :: mla
  type(bc_tower) :: bct
  type(multifab), allocatable :: rh(:), phi(:), alpha(:), beta(:,:)
  type(bndry_reg) :: flx(1)
  real(dp_t), allocatable :: dx(:,:)
  integer :: hdr(8), dhi(3), pbc(3,2), dim, nb, b, d
  real(dp_t) :: dx3(3), eps(2)
  character(len=512) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  dim = hdr(1)
  nb = hdr(2)
  read(u) dhi, pbc
  mla%dim = dim
  mla%nboxes = nb
  mla%pd_lo = 0
  mla%pd_hi = dhi
  allocate(mla%lo(3,nb), mla%hi(3,nb))
  do b = 1, nb
     read(u) mla%lo(:,b), mla%hi(:,b)
  end do
  read(u) dx3, eps
  allocate(dx(1,dim))
  dx(1,:) = dx3(1:dim)
  bct%phys_bc = pbc
  mac_precision = hdr(7)
  press_comp = 100
  foextrap_comp = 101
  hoextrap_comp = 102
  allocate(rh(1), phi(1), alpha(1), beta(1,dim))
  call multifab_build(rh(1), mla, hdr(3), 0)
  call multifab_build(phi(1), mla, hdr(4), 0)
  call multifab_build(alpha(1), mla, hdr(5), 0)
  do d = 1, dim
     call multifab_build(beta(1,d), mla, hdr(6), d)
  end do
  do b = 1, nb
     read(u) rh(1)%fabs(b)%p
     read(u) phi(1)%fabs(b)%p
     read(u) alpha(1)%fabs(b)%p
     do d = 1, dim
        read(u) beta(1,d)%fabs(b)%p
     end do
  end do
  close(u)

  call mac_multigrid(mla, rh, phi, flx, alpha, beta, dx, bct, hdr(8), 2, eps(1), eps(2))

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  do b = 1, nb
     write(u) phi(1)%fabs(b)%p
  end do
  write(u) int(last_stats%iters, 4), int(last_stats%converged, 4)
  write(u) last_stats%res0, last_stats%res
  close(u)
  call macproject_release(mla)

end program mac_multigrid_driver
