! swap_step_strict.f90 -- one time step of maind.f:385-479 written against
! the reference's routine names (module swap_ref), with host arrays that
! maind.f would own.  It shows the drop-in call sequence and is the
! executable the Fortran-boundary test runs:
!
!   swap_step_strict <parameter file> <output file>
!
! The parameter file holds, whitespace separated: ncx ncy ncz ppc dx dt
! ntsub alpha beta nu vsw vth seed.  The output file gets, as REAL*4 stream
! records in this order: np(nx,ny,nz), up(nx,ny,nz,3), b1(nx,ny,nz,3) after
! the step, xp(n,3) and vp(n,3) after the step.
program swap_step_strict
  use swap_ref
  implicit none
  integer :: ncx, ncy, ncz, ppc, ntsub, n, m, u
  integer(8) :: seed
  real(8) :: dx, dt, alpha, beta, nu0, vsw, vth
  character(len=512) :: pfile, ofile
  real, allocatable :: b0(:,:,:,:), b1(:,:,:,:), b12(:,:,:,:), b1p2(:,:,:,:), bt(:,:,:,:)
  real, allocatable :: btc(:,:,:,:), up(:,:,:,:), aj(:,:,:,:), E(:,:,:,:)
  real, allocatable :: np(:,:,:), nu(:,:,:)
  real, allocatable :: xp(:,:), vp(:,:), vp1(:,:), vplus(:,:), vminus(:,:), Ep(:,:)
  real :: input_p(3)

  call get_command_argument(1, pfile)
  call get_command_argument(2, ofile)
  open(newunit=u, file=trim(pfile), status='old', action='read')
  read(u, *) ncx, ncy, ncz, ppc, dx, dt, ntsub, alpha, beta, nu0, vsw, vth, seed
  close(u)
  n = ppc * ncx * ncy * ncz

  call swap_ref_init(ncx, ncy, ncz, dx, dx, dx, dt, ntsub, alpha, beta, nu0, &
       int(n, 8), 0, seed)
  allocate(b0(nx,ny,nz,3), b1(nx,ny,nz,3), b12(nx,ny,nz,3), b1p2(nx,ny,nz,3), &
       bt(nx,ny,nz,3), btc(nx,ny,nz,3), up(nx,ny,nz,3), aj(nx,ny,nz,3), E(nx,ny,nz,3))
  allocate(np(nx,ny,nz), nu(nx,ny,nz))
  allocate(xp(n,3), vp(n,3), vp1(n,3), vplus(n,3), vminus(n,3), Ep(n,3))
  ! uniform B0 along z (grd6_setup, initial.f:36-63); nu = nu_init (:79)
  b0 = 0.0
  b0(:,:,:,3) = 0.2e-9 * (1.6e-19 / 1.67e-27)
  b1 = b0; b12 = b0; b1p2 = b0; bt = b0
  nu = real(nu0)
  up = 0.0; aj = 0.0; E = 0.0; np = 0.0
  input_p = 0.0

  call swap_ref_load_maxwellian(n, vsw, vth)
  call swap_ref_get_particles(xp, vp, vp1)

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

  open(newunit=u, file=trim(ofile), access='stream', form='unformatted', status='replace')
  write(u) np, up, b1, xp, vp
  close(u)
  call swap_ref_finalize()
end program swap_step_strict
