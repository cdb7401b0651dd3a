! swap_step_strict.f90 -- one time step of maind.f:385-479 written against
! the reference's routine names (module swap_ref), with host arrays that
! maind.f would own.  It shows the drop-in call sequence and is the
! executable the Fortran-boundary test runs:
!
!   swap_step_strict <parameter file> <output file>
!
! The parameter file holds, whitespace separated: ncx ncy ncz ppc dx dt
! ntsub alpha beta nu vsw vth seed nf_init beta_pu m.  With nf_init > 0 the
! step opens with the solar-wind pickup source as maind.f:379-383 does
! (Ionize_sw_mp at step m, ndot = 1e-6 nf_init omega_p, ring protons of
! weight beta_pu).  The output file gets, as stream records in this order:
! Ni_tot (INTEGER*4), then REAL*4 np(nx,ny,nz), up(nx,ny,nz,3),
! b1(nx,ny,nz,3), xp(Ni_tot,3), vp(Ni_tot,3), beta_p(Ni_tot), mrat(Ni_tot)
! after the step.
program swap_step_strict
  use swap_ref
  implicit none
  integer :: ncx, ncy, ncz, ppc, ntsub, n, m, u, mstep, nmax
  integer(8) :: seed
  real(8) :: dx, dt, alpha, beta, nu0, vsw, vth, nf_init, beta_pu
  real(8), parameter :: q = 1.6d-19, mion = 1.67d-27, b0_init = 0.2d-9
  real, allocatable :: beta_p(:), mrat(:)
  character(len=512) :: pfile, ofile
  real, allocatable :: b0(:,:,:,:), b1(:,:,:,:), b12(:,:,:,:), b1p2(:,:,:,:), bt(:,:,:,:)
  real, allocatable :: btc(:,:,:,:), up(:,:,:,:), aj(:,:,:,:), E(:,:,:,:)
  real, allocatable :: np(:,:,:), nu(:,:,:)
  real, allocatable :: xp(:,:), vp(:,:), vp1(:,:), vplus(:,:), vminus(:,:), Ep(:,:)
  real :: input_p(3)

  call get_command_argument(1, pfile)
  call get_command_argument(2, ofile)
  open(newunit=u, file=trim(pfile), status='old', action='read')
  read(u, *) ncx, ncy, ncz, ppc, dx, dt, ntsub, alpha, beta, nu0, vsw, vth, seed, nf_init, &
       beta_pu, mstep
  close(u)
  n = ppc * ncx * ncy * ncz
  nmax = n + n / 4 + 1000   ! Ni_max: room for the new ions

  call swap_ref_init(ncx, ncy, ncz, dx, dx, dx, dt, ntsub, alpha, beta, nu0, &
       int(nmax, 8), 0, seed, SWAP_SPECIES)
  ! inputs.f: omega_p = q b0_init / mion
  call swap_ref_set_sw(nf_init, q * b0_init / mion, vsw, beta_pu, mion)
  allocate(b0(nx,ny,nz,3), b1(nx,ny,nz,3), b12(nx,ny,nz,3), b1p2(nx,ny,nz,3), &
       bt(nx,ny,nz,3), btc(nx,ny,nz,3), up(nx,ny,nz,3), aj(nx,ny,nz,3), E(nx,ny,nz,3))
  allocate(np(nx,ny,nz), nu(nx,ny,nz))
  allocate(xp(nmax,3), vp(nmax,3), vp1(nmax,3), vplus(nmax,3), vminus(nmax,3), Ep(nmax,3))
  allocate(beta_p(nmax), mrat(nmax))
  ! uniform B0 along z (grd6_setup, initial.f:36-63); nu = nu_init (:79)
  b0 = 0.0
  b0(:,:,:,3) = 0.2e-9 * (1.6e-19 / 1.67e-27)
  b1 = b0; b12 = b0; b1p2 = b0; bt = b0
  nu = real(nu0)
  up = 0.0; aj = 0.0; E = 0.0; np = 0.0
  input_p = 0.0

  call swap_ref_load_maxwellian(n, vsw, vth)
  call swap_ref_get_particles(xp, vp, vp1)

  ! sources (maind.f:379-383)
  if (nf_init > 0d0 .and. swap_ref_ni_tot() < nmax) &
       call Ionize_sw_mp(np, vp, vp1, xp, mstep, input_p, up)

  ! maind.f:385-479
  call get_interp_weights(xp)
  call update_np(np)
  call update_up(vp, np, up)
  call curlB(bt, np, aj)
  call edge_to_center(bt, btc)
  call extrapol_up(up, vp, vp1, np)
  call get_Ep(Ep, aj, np, up, btc, nu)
  call get_vplus_vminus(Ep, btc, vp, vplus, vminus)
  call improve_up(vp1, vplus, vminus, up, np)
  call get_Ep(Ep, aj, np, up, btc, nu)
  call get_vplus_vminus(Ep, btc, vp, vplus, vminus)
  call get_vp_final(Ep, vp, vp1, vplus)
  call move_ion_half(xp, vp, vp1, input_p)
  call get_interp_weights(xp)
  call update_np(np)
  call update_up(vp, np, up)
  call check_time_step(bt, np)
  do m = 1, int(ntf)
     call curlB(bt, np, aj)
     call predict_B(b0, b1, b12, b1p2, bt, E, aj, up, np, nu)
     call correct_B(b0, b1, b1p2, E, aj, up, np, nu)
     call f_update_tlev(b1, b12, b1p2, bt, b0)
  end do
  call move_ion_half(xp, vp, vp1, input_p)

  n = swap_ref_ni_tot()
  call swap_ref_get_species(beta_p, mrat)
  open(newunit=u, file=trim(ofile), access='stream', form='unformatted', status='replace')
  write(u) int(n, 4)
  write(u) np, up, b1, xp(1:n,:), vp(1:n,:), beta_p(1:n), mrat(1:n)
  close(u)
  call swap_ref_finalize()
end program swap_step_strict
