! macproject_driver.f90 -- standalone Fortran caller of the drop-in
! macproject (macproject_hip.f90), as MAESTRO's time step calls it
! (Source/advance.f90:466-467).  It reads one problem from a binary stream
! file, projects it (twice when asked, reusing the cached device plan like
! the predictor/corrector pair of a step) and writes the projected edge
! velocities and phi.
!
! Input (stream, native endianness; arrays per box in Fortran order with
! their ghost cells):
!   int32  dim, nboxes, ng_u, ng_phi, ng_rho, alt, has_beta0, precision, ncalls,
!          mg_cycle_type, mg_bottom_solver, mac_omega x 1000
!   int32  domain_hi(3), phys_bc(3,2) (direction fastest)
!   int32  box lo(3), hi(3) for every box
!   real64 dx(3)
!   per box: umac(d) for d = 1..dim, phi, rho
!   when has_beta0: beta0(0:nr-1), beta0_edge(0:nr)
! Output: per box umac(d) for d = 1..dim, phi; then int32 iters, converged;
! real64 res0, res.
program macproject_driver

  use iso_c_binding
  use mac_types_module
  use macproject_module
  implicit none

  type(ml_layout) :: mla
  type(bc_tower) :: bct
  type(multifab), allocatable :: umac(:,:), phi(:), rho(:)
  real(dp_t), allocatable :: beta0(:,:), beta0_edge(:,:), dx(:,:)
  integer :: hdr(12), dhi(3), pbc(3,2), dim, nb, b, d, nr, call_no
  real(dp_t) :: dx3(3)
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
  read(u) dx3
  allocate(dx(1,dim))
  dx(1,:) = dx3(1:dim)
  bct%phys_bc = pbc
  allocate(umac(1,dim), phi(1), rho(1))
  do d = 1, dim
     call multifab_build(umac(1,d), mla, hdr(3), d)
  end do
  call multifab_build(phi(1), mla, hdr(4), 0)
  call multifab_build(rho(1), mla, hdr(5), 0)
  do b = 1, nb
     do d = 1, dim
        read(u) umac(1,d)%fabs(b)%p
     end do
     read(u) phi(1)%fabs(b)%p
     read(u) rho(1)%fabs(b)%p
  end do
  use_alt_energy_fix = hdr(6) /= 0
  mac_precision = hdr(8)
  mg_cycle_type = hdr(10)
  mg_bottom_solver = hdr(11)
  mac_omega = real(hdr(12), dp_t) / 1000.0_dp_t
  if (hdr(7) /= 0) then
     nr = dhi(dim) + 1
     allocate(beta0(1,0:nr-1), beta0_edge(1,0:nr))
     read(u) beta0(1,:)
     read(u) beta0_edge(1,:)
  end if
  close(u)

  do call_no = 1, max(1, hdr(9))
     if (hdr(7) /= 0) then
        call macproject(mla, umac, phi, rho, dx, bct, &
                        beta0_1d=beta0, beta0_1d_edge=beta0_edge)
     else
        call macproject(mla, umac, phi, rho, dx, bct)
     end if
     write(*,'(a,i0,a,i0,a,es12.4)') 'call ', call_no, ': V-cycles ', last_stats%iters, &
          '  res/res0 ', last_stats%res / max(last_stats%res0, tiny(1.0_dp_t))
  end do

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  do b = 1, nb
     do d = 1, dim
        write(u) umac(1,d)%fabs(b)%p
     end do
     write(u) phi(1)%fabs(b)%p
  end do
  write(u) int(last_stats%iters), int(last_stats%converged)
  write(u) real(last_stats%res0, dp_t), real(last_stats%res, dp_t)
  close(u)
  call macproject_release(mla)

end program macproject_driver
