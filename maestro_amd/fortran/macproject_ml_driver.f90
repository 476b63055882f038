! macproject_ml_driver.f90 -- standalone Fortran caller of the drop-in
! macproject (macproject_hip.f90) on a two-level hierarchy, as MAESTRO's
! time step calls it with nlevs = 2 (Source/advance.f90:466-467).  It reads
! one problem from a binary stream file, projects it (twice when asked,
! reusing the cached device plan) and writes both levels' projected edge
! velocities and phi.
!
! Input (stream, native endianness; arrays per box in Fortran order with
! their ghost cells):
!   int32  dim, nboxes1, nboxes2, ng_u, ng_phi, ng_rho, alt, has_beta0,
!          precision, ncalls
!   int32  domain_hi(3) of level 1, phys_bc(3,2) (direction fastest)
!   int32  box lo(3), hi(3) for every box of level 1, then of level 2
!   real64 dx(3) of level 1
!   per level, per box: umac(d) for d = 1..dim, phi, rho
!   when has_beta0: level 1 beta0(0:nr-1), beta0_edge(0:nr), then level 2
!   beta0(0:2nr-1), beta0_edge(0:2nr)
! Output: per level, per box umac(d) for d = 1..dim, phi; then int32 iters,
! converged; real64 res0, res.
program macproject_ml_driver

  use iso_c_binding
  use mac_types_module
  use macproject_module
  implicit none

  type(ml_layout) :: mla
  type(bc_tower) :: bct
  type(multifab), allocatable :: umac(:,:), phi(:), rho(:)
  real(dp_t), allocatable :: beta0(:,:), beta0_edge(:,:), dx(:,:)
  integer :: hdr(10), dhi(3), pbc(3,2), dim, nb(2), b, d, nr, call_no, lev
  real(dp_t) :: dx3(3)
  character(len=512) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  dim = hdr(1)
  nb = hdr(2:3)
  read(u) dhi, pbc
  mla%dim = dim
  mla%nlevel = 2
  mla%nboxes = nb(1)
  mla%nboxes2 = nb(2)
  mla%pd_lo = 0
  mla%pd_hi = dhi
  allocate(mla%lo(3,nb(1)), mla%hi(3,nb(1)), mla%lo2(3,nb(2)), mla%hi2(3,nb(2)))
  do b = 1, nb(1)
     read(u) mla%lo(:,b), mla%hi(:,b)
  end do
  do b = 1, nb(2)
     read(u) mla%lo2(:,b), mla%hi2(:,b)
  end do
  read(u) dx3
  allocate(dx(2,dim))
  dx(1,:) = dx3(1:dim)
  dx(2,:) = dx3(1:dim) / 2
  bct%phys_bc = pbc
  allocate(umac(2,dim), phi(2), rho(2))
  do lev = 1, 2
     do d = 1, dim
        call multifab_build_level(umac(lev,d), mla, lev, hdr(4), d)
     end do
     call multifab_build_level(phi(lev), mla, lev, hdr(5), 0)
     call multifab_build_level(rho(lev), mla, lev, hdr(6), 0)
     do b = 1, nb(lev)
        do d = 1, dim
           read(u) umac(lev,d)%fabs(b)%p
        end do
        read(u) phi(lev)%fabs(b)%p
        read(u) rho(lev)%fabs(b)%p
     end do
  end do
  use_alt_energy_fix = hdr(7) /= 0
  mac_precision = hdr(9)
  if (hdr(8) /= 0) then
     nr = dhi(dim) + 1
     allocate(beta0(2,0:2*nr-1), beta0_edge(2,0:2*nr))
     beta0 = 0.0_dp_t
     beta0_edge = 0.0_dp_t
     read(u) beta0(1,0:nr-1)
     read(u) beta0_edge(1,0:nr)
     read(u) beta0(2,0:2*nr-1)
     read(u) beta0_edge(2,0:2*nr)
  end if
  close(u)

  do call_no = 1, max(1, hdr(10))
     if (hdr(8) /= 0) then
        call macproject(mla, umac, phi, rho, dx, bct, &
                        beta0_1d=beta0, beta0_1d_edge=beta0_edge)
     else
        call macproject(mla, umac, phi, rho, dx, bct)
     end if
     write(*,'(a,i0,a,i0,a,es12.4)') 'call ', call_no, ': composite steps ', last_stats%iters, &
          '  res/res0 ', last_stats%res / max(last_stats%res0, tiny(1.0_dp_t))
  end do

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  do lev = 1, 2
     do b = 1, nb(lev)
        do d = 1, dim
           write(u) umac(lev,d)%fabs(b)%p
        end do
        write(u) phi(lev)%fabs(b)%p
     end do
  end do
  write(u) int(last_stats%iters), int(last_stats%converged)
  write(u) real(last_stats%res0, dp_t), real(last_stats%res, dp_t)
  close(u)
  call macproject_release(mla)

end program macproject_ml_driver
