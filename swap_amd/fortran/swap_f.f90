! swap_f.f90 -- ISO_C_BINDING shim over the C-ABI of libswap_hip.so
! (include/swap_hip.h).
!
! Module swap_hip_c declares the C entry points.  Module swap_ref gives the
! hot-path routines of the reference under their own names and argument
! lists (gutsp.f, gutsf.f, grid_interp.f, boundary.f), so that maind.f can
! call them unchanged after `use swap_ref`:
!
!   get_interp_weights(xp)                     gutsp.f:1378
!   update_np(np)                              gutsp.f:1441
!   update_up(vp,np,up)                        gutsp.f:1649
!   extrapol_up(up,vp,vp1,np)                  gutsp.f:127
!   get_Ep(Ep,aj,np,up,btc,nu)                 gutsp.f:155
!   get_vplus_vminus(Ep,btc,vp,vplus,vminus)   gutsp.f:281
!   improve_up(vp1,vplus,vminus,up,np)         gutsp.f:407
!   get_vp_final(Ep,vp,vp1,vplus)              gutsp.f:433
!   move_ion_half(xp,vp,vp1,input_p)           gutsp.f:454
!   curlB(b1,np,aj)                            gutsf.f:178
!   curlE(E,curl_E)                            gutsf.f:249
!   predict_B(b0,b1,b12,b1p2,bt,E,aj,up,np,nu) gutsf.f:1115
!   correct_B(b0,b1,b1p2,E,aj,up,np,nu)        gutsf.f:1309
!   f_update_tlev(b1,b12,b1p2,bt,b0)           gutsf.f:10
!   check_time_step(bt,np)                     gutsf.f:1776
!   edge_to_center(bt,btc)                     grid_interp.f:10
!   edge_to_face(bt,btmf)                      grid_interp.f:67
!   face_to_center(v,vc)                       grid_interp.f:136
!   periodic(b), periodic_scalar(b)            boundary.f:9, 167
!   Ionize_sw_mp(np,vp,vp1,xp,m_tstep,input_p,up)     chem_rates.f:508
!   Ionize_pluto_mp(np,vp,vp1,xp,m_tstep,input_p,up)  chem_rates.f:290
!   res_chex(xp,vp)                                   chem_rates.f:805
!
! The sources read the reference's module inputs (nf_init, omega_p, vsw,
! beta_pu of inputs.f; the cloud's Qo, vrad, tau_photo, which the reference
! leaves undefined) from swap_ref_set_sw and swap_ref_set_cloud; they append
! to the host arrays, raise Ni_tot and keep the library's per-ion beta_p
! and mrat (swap_ref_get_species copies them to the host).
!
! These are the strict forms: every call moves its REAL*4 host arguments to
! the device, runs the routine there and moves the results (and the ghost
! layers the reference mutates in place) back, so the host arrays of
! maind.f stay authoritative.  swap_ref_step runs whole steps resident on
! the device instead (maind.f:385-479 fused).  Failures stop the program
! with the library's message, as the reference stops on input errors.
module swap_hip_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SWAP_FP32 = 32, SWAP_FP64 = 64
  integer(c_int), parameter :: SWAP_SPECIES = 1, SWAP_OUTFLOW_X = 2, SWAP_TRACK_IDS = 4
  integer(c_int), parameter :: SWAP_F_B1 = 0, SWAP_F_B12 = 1, SWAP_F_B1P2 = 2, &
       SWAP_F_BTC = 3, SWAP_F_NP = 4, SWAP_F_UP = 5, SWAP_F_AJ = 6, SWAP_F_E = 7, &
       SWAP_F_NU = 8, SWAP_F_AJC = 9, SWAP_F_UPC = 10, SWAP_F_TMP0 = 11, &
       SWAP_F_TMP1 = 12, SWAP_F_TMP2 = 13
  integer(c_int), parameter :: SWAP_P_XP = 0, SWAP_P_VP = 1, SWAP_P_VP1 = 2, &
       SWAP_P_EP = 3, SWAP_P_VPLUS = 4, SWAP_P_VMINUS = 5, SWAP_P_IJKP = 6, &
       SWAP_P_WGHT = 7, SWAP_P_CELL = 8, SWAP_P_FRAC = 9, SWAP_P_BETA = 10, &
       SWAP_P_MRAT = 11, SWAP_P_ID = 12
  integer(c_int), parameter :: SWAP_V_CUR = 0, SWAP_V_EXTRAP = 1, SWAP_V_VP1 = 2

  type, bind(C) :: swap_config
     integer(c_int) :: ncx, ncy, ncz
     real(c_double) :: dx, dy, dz
     real(c_double) :: dt
     integer(c_int) :: ntsub
     real(c_double) :: alpha, beta, nu
     integer(c_int64_t) :: capacity
     integer(c_int) :: precision, flags, sort_every
     integer(c_int64_t) :: seed
  end type swap_config

  ! the neutral cloud of swap_ionize_cloud
  type, bind(C) :: swap_cloud
     real(c_double) :: cx, cy, cz, N0, r0, H, tau_photo, sigma_chex, mrat_pu, beta_min
     integer(c_int) :: max_per_cell
     real(c_double) :: rmin
  end type swap_cloud

  ! Energy_diag's record (part_init.f:14-139), J
  type, bind(C) :: swap_energy
     real(c_double) :: input_E, Evp, Euf, EB1, EB1x, EB1y, EB1z, EE, total_E, norm_E
  end type swap_energy

  interface
     integer(c_int) function swap_create(cfg, out) bind(C, name="swap_create")
       import :: c_int, c_ptr, swap_config
       type(swap_config), intent(in) :: cfg
       type(c_ptr), intent(out) :: out
     end function swap_create
     integer(c_int) function swap_destroy(ctx) bind(C, name="swap_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_destroy
     type(c_ptr) function swap_last_error(ctx) bind(C, name="swap_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function swap_last_error
     integer(c_int) function swap_sync(ctx) bind(C, name="swap_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_sync
     integer(c_int) function swap_load_particles(ctx, n, xp, vp, vp1, beta_p, mrat) &
          bind(C, name="swap_load_particles")
       import :: c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       real(c_double), intent(in) :: xp(*), vp(*), vp1(*)
       type(c_ptr), value :: beta_p, mrat
     end function swap_load_particles
     integer(c_int) function swap_load_maxwellian(ctx, n, vsw, vth, beta_p, mrat) &
          bind(C, name="swap_load_maxwellian")
       import :: c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       real(c_double), value :: vsw, vth, beta_p, mrat
     end function swap_load_maxwellian
     integer(c_int) function swap_get_particle_array(ctx, which, host) &
          bind(C, name="swap_get_particle_array")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: which
       type(c_ptr), value :: host
     end function swap_get_particle_array
     integer(c_int) function swap_set_particle_array(ctx, which, host) &
          bind(C, name="swap_set_particle_array")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: which
       type(c_ptr), value :: host
     end function swap_set_particle_array
     integer(c_int) function swap_set_field_f32(ctx, field, host) &
          bind(C, name="swap_set_field_f32")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       real(c_float), intent(in) :: host(*)
     end function swap_set_field_f32
     integer(c_int) function swap_get_field_f32(ctx, field, host) &
          bind(C, name="swap_get_field_f32")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
       real(c_float), intent(out) :: host(*)
     end function swap_get_field_f32
     integer(c_int) function swap_periodic(ctx, field) bind(C, name="swap_periodic")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: field
     end function swap_periodic
     integer(c_int) function swap_get_interp_weights(ctx) bind(C, name="swap_get_interp_weights")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_get_interp_weights
     integer(c_int) function swap_update_np(ctx) bind(C, name="swap_update_np")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_update_np
     integer(c_int) function swap_update_up(ctx, vsrc) bind(C, name="swap_update_up")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: vsrc
     end function swap_update_up
     integer(c_int) function swap_get_Ep(ctx) bind(C, name="swap_get_Ep")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_get_Ep
     integer(c_int) function swap_get_vplus_vminus(ctx) bind(C, name="swap_get_vplus_vminus")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_get_vplus_vminus
     integer(c_int) function swap_improve_up(ctx) bind(C, name="swap_improve_up")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_improve_up
     integer(c_int) function swap_get_vp_final(ctx) bind(C, name="swap_get_vp_final")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_get_vp_final
     integer(c_int) function swap_move_ion_half(ctx) bind(C, name="swap_move_ion_half")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_move_ion_half
     integer(c_int) function swap_curlB(ctx, b_in, aj_out) bind(C, name="swap_curlB")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: b_in, aj_out
     end function swap_curlB
     integer(c_int) function swap_curlE(ctx, e_in, out) bind(C, name="swap_curlE")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: e_in, out
     end function swap_curlE
     integer(c_int) function swap_predict_B(ctx) bind(C, name="swap_predict_B")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_predict_B
     integer(c_int) function swap_correct_B(ctx) bind(C, name="swap_correct_B")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_correct_B
     integer(c_int) function swap_check_time_step(ctx, ntf, dtsub) &
          bind(C, name="swap_check_time_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: ntf
       real(c_double), intent(out) :: dtsub
     end function swap_check_time_step
     integer(c_int) function swap_edge_to_center(ctx, fin, fout) &
          bind(C, name="swap_edge_to_center")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: fin, fout
     end function swap_edge_to_center
     integer(c_int) function swap_edge_to_face(ctx, fin, fout) bind(C, name="swap_edge_to_face")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: fin, fout
     end function swap_edge_to_face
     integer(c_int) function swap_face_to_center(ctx, fin, fout) &
          bind(C, name="swap_face_to_center")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: fin, fout
     end function swap_face_to_center
     integer(c_int) function swap_prime(ctx) bind(C, name="swap_prime")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function swap_prime
     integer(c_int) function swap_step(ctx, nsteps) bind(C, name="swap_step")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function swap_step
     integer(c_int) function swap_energy_diag(ctx, mion, b0, out) &
          bind(C, name="swap_energy_diag")
       import :: c_int, c_ptr, c_double, swap_energy
       type(c_ptr), value :: ctx
       real(c_double), value :: mion
       real(c_double), intent(in) :: b0(3)
       type(swap_energy), intent(out) :: out
     end function swap_energy_diag
     integer(c_int) function swap_set_input_energy(ctx, mion, input_E) &
          bind(C, name="swap_set_input_energy")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: mion, input_E
     end function swap_set_input_energy
     integer(c_int) function swap_ionize_sw(ctx, step, ndot, vsw, beta_pu, injected) &
          bind(C, name="swap_ionize_sw")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: step
       real(c_double), value :: ndot, vsw, beta_pu
       integer(c_int64_t), intent(out) :: injected
     end function swap_ionize_sw
     integer(c_int) function swap_ionize_cloud(ctx, step, cloud, injected, chex) &
          bind(C, name="swap_ionize_cloud")
       import :: c_int, c_ptr, c_int64_t, swap_cloud
       type(c_ptr), value :: ctx
       integer(c_int), value :: step
       type(swap_cloud), intent(in) :: cloud
       integer(c_int64_t), intent(out) :: injected, chex
     end function swap_ionize_cloud
     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface
end module swap_hip_c


module swap_ref
  use iso_c_binding
  use swap_hip_c
  implicit none
  private :: chk, fset, fget, pset, pget, xp_set, xp_get, pull_new, redeposit

  type(c_ptr), save :: swap_ctx_h = c_null_ptr
  integer, save :: nx = 0, ny = 0, nz = 0   ! grid incl. ghosts (dimensions.f:5)
  integer, save :: Ni_tot = 0               ! particles on the device
  real, save :: dtsub = 0.0, ntf = 0.0      ! subcycle step and count (gutsf.f:1776)
  integer, save :: Ni_max = 0               ! particle capacity (dimensions.f)
  real(8), save :: beta_s = 0d0, mion_s = 1.67d-27   ! beta (misc.f:462) and mion (inputs.f)
  ! Ionize_sw_mp's inputs: nf_init, omega_p, vsw, beta_pu (inputs.f)
  real(8), save :: sw_nf_init = 0d0, sw_omega_p = 0d0, sw_vsw = 0d0, sw_beta_pu = 1d0
  type(swap_cloud), save :: cloud_s
  integer, save :: src_step = 0             ! m_tstep of the last source call

contains

  subroutine chk(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = swap_last_error(swap_ctx_h)
    n = 0
    if (c_associated(p)) n = int(c_strlen(p))
    if (n > 0) then
       call c_f_pointer(p, msg, [n])
       write(*, '(a, a, i0, a, 999a)') what, ' failed (', rc, '): ', msg
    else
       write(*, '(a, a, i0, a)') what, ' failed (', rc, ')'
    end if
    stop 1
  end subroutine chk

  ! -------- set-up --------
  ! flags: SWAP_SPECIES keeps per-ion beta_p and mrat (needed by the
  ! sources), SWAP_TRACK_IDS per-ion ids; capacity is the reference's Ni_max
  subroutine swap_ref_init(ncx, ncy, ncz, dx, dy, delz, dt, ntsub, alpha, beta, nu, &
       capacity, sort_every, seed, flags)
    integer, intent(in) :: ncx, ncy, ncz, ntsub, sort_every
    real(8), intent(in) :: dx, dy, delz, dt, alpha, beta, nu
    integer(8), intent(in) :: capacity, seed
    integer, intent(in), optional :: flags
    type(swap_config) :: cfg
    cfg%ncx = ncx; cfg%ncy = ncy; cfg%ncz = ncz
    cfg%dx = dx; cfg%dy = dy; cfg%dz = delz; cfg%dt = dt
    cfg%ntsub = ntsub; cfg%alpha = alpha; cfg%beta = beta; cfg%nu = nu
    cfg%capacity = capacity; cfg%precision = SWAP_FP32; cfg%flags = 0
    if (present(flags)) cfg%flags = flags
    cfg%sort_every = sort_every; cfg%seed = seed
    call chk(swap_create(cfg, swap_ctx_h), 'swap_create')
    nx = ncx + 2; ny = ncy + 2; nz = ncz + 2
    dtsub = real(dt / ntsub); ntf = real(ntsub)
    Ni_max = int(capacity)
    beta_s = beta
  end subroutine swap_ref_init

  ! the inputs of Ionize_sw_mp (inputs.f: nf_init, omega_p, vsw, beta_pu; the
  ! ion mass mion of the input_p account)
  subroutine swap_ref_set_sw(nf_init, omega_p, vsw, beta_pu, mion)
    real(8), intent(in) :: nf_init, omega_p, vsw, beta_pu, mion
    sw_nf_init = nf_init; sw_omega_p = omega_p; sw_vsw = vsw; sw_beta_pu = beta_pu
    mion_s = mion
  end subroutine swap_ref_set_sw

  ! the neutral cloud of Ionize_pluto_mp and res_chex: n_n(r) = N0 (r0/r)^2
  ! exp(-r/H) about (cx, cy, cz), r >= rmin (the reference's Qo, vrad,
  ! tau_photo and S_radius are undefined, chem_rates.f:350-352)
  subroutine swap_ref_set_cloud(cx, cy, cz, N0, r0, H, tau_photo, sigma_chex, mrat_pu, &
       beta_min, max_per_cell, rmin)
    real(8), intent(in) :: cx, cy, cz, N0, r0, H, tau_photo, sigma_chex, mrat_pu, beta_min, rmin
    integer, intent(in) :: max_per_cell
    cloud_s%cx = cx; cloud_s%cy = cy; cloud_s%cz = cz; cloud_s%N0 = N0; cloud_s%r0 = r0
    cloud_s%H = H; cloud_s%tau_photo = tau_photo; cloud_s%sigma_chex = sigma_chex
    cloud_s%mrat_pu = mrat_pu; cloud_s%beta_min = beta_min
    cloud_s%max_per_cell = max_per_cell; cloud_s%rmin = rmin
  end subroutine swap_ref_set_cloud

  ! the library's per-ion beta_p and mrat (global.f) into host arrays
  subroutine swap_ref_get_species(beta_p, mrat)
    real, intent(inout) :: beta_p(:), mrat(:)
    real(c_float), allocatable, target :: t(:)
    allocate(t(Ni_tot))
    call chk(swap_get_particle_array(swap_ctx_h, SWAP_P_BETA, c_loc(t)), 'get beta_p')
    beta_p(1:Ni_tot) = t
    call chk(swap_get_particle_array(swap_ctx_h, SWAP_P_MRAT, c_loc(t)), 'get mrat')
    mrat(1:Ni_tot) = t
  end subroutine swap_ref_get_species

  ! ions on the device (the reference's Ni_tot)
  integer function swap_ref_ni_tot()
    swap_ref_ni_tot = Ni_tot
  end function swap_ref_ni_tot

  subroutine swap_ref_finalize()
    if (c_associated(swap_ctx_h)) call chk(swap_destroy(swap_ctx_h), 'swap_destroy')
    swap_ctx_h = c_null_ptr
  end subroutine swap_ref_finalize

  ! seeded drifting-Maxwellian load (load_Maxwellian, part_init.f:146-232)
  subroutine swap_ref_load_maxwellian(n, vsw, vth)
    integer, intent(in) :: n
    real(8), intent(in) :: vsw, vth
    call chk(swap_load_maxwellian(swap_ctx_h, int(n, c_int64_t), vsw, vth, 1d0, 1d0), &
         'load_maxwellian')
    Ni_tot = n
  end subroutine swap_ref_load_maxwellian

  ! host particle arrays (Ni_max,3) -> device; Ni_tot = n
  subroutine swap_ref_load(n, xp, vp, vp1)
    integer, intent(in) :: n
    real, intent(in) :: xp(:,:), vp(:,:), vp1(:,:)
    real(c_double), allocatable :: x(:,:), v(:,:), v1(:,:)
    allocate(x(n,3), v(n,3), v1(n,3))
    x = real(xp(1:n,:), c_double); v = real(vp(1:n,:), c_double)
    v1 = real(vp1(1:n,:), c_double)
    call chk(swap_load_particles(swap_ctx_h, int(n, c_int64_t), x, v, v1, c_null_ptr, &
         c_null_ptr), 'load_particles')
    Ni_tot = n
  end subroutine swap_ref_load

  ! device particles -> host arrays (Ni_max,3), rows 1..Ni_tot
  subroutine swap_ref_get_particles(xp, vp, vp1)
    real, intent(inout) :: xp(:,:), vp(:,:), vp1(:,:)
    call xp_get(xp)
    call pget(SWAP_P_VP, vp)
    call pget(SWAP_P_VP1, vp1)
  end subroutine swap_ref_get_particles

  ! resident time stepping: maind.f:385-479 for nsteps steps
  subroutine swap_ref_step(nsteps)
    integer, intent(in) :: nsteps
    call chk(swap_prime(swap_ctx_h), 'prime')
    call chk(swap_step(swap_ctx_h, int(nsteps, c_int)), 'step')
  end subroutine swap_ref_step

  ! -------- host <-> device helpers --------
  subroutine fset(id, a)
    integer(c_int), intent(in) :: id
    real(c_float), intent(in) :: a(*)
    call chk(swap_set_field_f32(swap_ctx_h, id, a), 'set_field')
  end subroutine fset
  subroutine fget(id, a)
    integer(c_int), intent(in) :: id
    real(c_float), intent(out) :: a(*)
    call chk(swap_get_field_f32(swap_ctx_h, id, a), 'get_field')
  end subroutine fget
  subroutine pset(id, a)   ! (Ni_max,3) REAL*4 -> device (Ni_tot,3)
    integer(c_int), intent(in) :: id
    real, intent(in) :: a(:,:)
    real(c_float), allocatable, target :: t(:,:)
    allocate(t(Ni_tot, size(a, 2)))
    t = a(1:Ni_tot, :)
    call chk(swap_set_particle_array(swap_ctx_h, id, c_loc(t)), 'set_particle_array')
  end subroutine pset
  subroutine pget(id, a)
    integer(c_int), intent(in) :: id
    real, intent(inout) :: a(:,:)
    real(c_float), allocatable, target :: t(:,:)
    allocate(t(Ni_tot, size(a, 2)))
    call chk(swap_get_particle_array(swap_ctx_h, id, c_loc(t)), 'get_particle_array')
    a(1:Ni_tot, :) = t
  end subroutine pget
  subroutine xp_set(xp)
    real, intent(in) :: xp(:,:)
    real(c_double), allocatable, target :: t(:,:)
    allocate(t(Ni_tot, 3))
    t = real(xp(1:Ni_tot, :), c_double)
    call chk(swap_set_particle_array(swap_ctx_h, SWAP_P_XP, c_loc(t)), 'set xp')
  end subroutine xp_set
  subroutine xp_get(xp)
    real, intent(inout) :: xp(:,:)
    real(c_double), allocatable, target :: t(:,:)
    allocate(t(Ni_tot, 3))
    call chk(swap_get_particle_array(swap_ctx_h, SWAP_P_XP, c_loc(t)), 'get xp')
    xp(1:Ni_tot, :) = real(t)
  end subroutine xp_get

  ! -------- boundary.f --------
  subroutine periodic(b)
    real, intent(inout) :: b(nx,ny,nz,3)
    call fset(SWAP_F_TMP0, b)
    call chk(swap_periodic(swap_ctx_h, SWAP_F_TMP0), 'periodic')
    call fget(SWAP_F_TMP0, b)
  end subroutine periodic

  subroutine periodic_scalar(b)
    real, intent(inout) :: b(nx,ny,nz)
    call fset(SWAP_F_NP, b)
    call chk(swap_periodic(swap_ctx_h, SWAP_F_NP), 'periodic_scalar')
    call fget(SWAP_F_NP, b)
  end subroutine periodic_scalar

  ! -------- grid_interp.f (each makes its input periodic first) --------
  subroutine edge_to_center(bt, btc)
    real, intent(inout) :: bt(nx,ny,nz,3)
    real, intent(out) :: btc(nx,ny,nz,3)
    call fset(SWAP_F_B1, bt)
    call chk(swap_periodic(swap_ctx_h, SWAP_F_B1), 'periodic')
    call chk(swap_edge_to_center(swap_ctx_h, SWAP_F_B1, SWAP_F_BTC), 'edge_to_center')
    call fget(SWAP_F_BTC, btc)
    call fget(SWAP_F_B1, bt)
  end subroutine edge_to_center

  subroutine edge_to_face(bt, btmf)
    real, intent(inout) :: bt(nx,ny,nz,3)
    real, intent(out) :: btmf(nx,ny,nz,3)
    call fset(SWAP_F_B1, bt)
    call chk(swap_periodic(swap_ctx_h, SWAP_F_B1), 'periodic')
    call chk(swap_edge_to_face(swap_ctx_h, SWAP_F_B1, SWAP_F_TMP0), 'edge_to_face')
    call fget(SWAP_F_TMP0, btmf)
    call fget(SWAP_F_B1, bt)
  end subroutine edge_to_face

  subroutine face_to_center(v, vc)
    real, intent(inout) :: v(nx,ny,nz,3)
    real, intent(out) :: vc(nx,ny,nz,3)
    call fset(SWAP_F_TMP0, v)
    call chk(swap_periodic(swap_ctx_h, SWAP_F_TMP0), 'periodic')
    call chk(swap_face_to_center(swap_ctx_h, SWAP_F_TMP0, SWAP_F_TMP1), 'face_to_center')
    call fget(SWAP_F_TMP1, vc)
    call fget(SWAP_F_TMP0, v)
  end subroutine face_to_center

  ! -------- gutsf.f --------
  subroutine curlB(b1, np, aj)
    real, intent(inout) :: b1(nx,ny,nz,3)
    real, intent(in) :: np(nx,ny,nz)
    real, intent(out) :: aj(nx,ny,nz,3)
    call fset(SWAP_F_B1, b1)
    call fset(SWAP_F_NP, np)
    call chk(swap_periodic(swap_ctx_h, SWAP_F_B1), 'periodic')
    call chk(swap_curlB(swap_ctx_h, SWAP_F_B1, SWAP_F_AJ), 'curlB')
    call fget(SWAP_F_AJ, aj)
    call fget(SWAP_F_B1, b1)
  end subroutine curlB

  subroutine curlE(E, curl_E)
    real, intent(in) :: E(nx,ny,nz,3)
    real, intent(out) :: curl_E(nx,ny,nz,3)
    call fset(SWAP_F_E, E)
    call chk(swap_curlE(swap_ctx_h, SWAP_F_E, SWAP_F_TMP0), 'curlE')
    call fget(SWAP_F_TMP0, curl_E)
  end subroutine curlE

  subroutine predict_B(b0, b1, b12, b1p2, bt, E, aj, up, np, nu)
    real, intent(in) :: b0(nx,ny,nz,3), b1(nx,ny,nz,3), b12(nx,ny,nz,3), bt(nx,ny,nz,3)
    real, intent(out) :: b1p2(nx,ny,nz,3), E(nx,ny,nz,3)
    real, intent(inout) :: aj(nx,ny,nz,3), up(nx,ny,nz,3)
    real, intent(in) :: np(nx,ny,nz), nu(nx,ny,nz)
    ! bt is the total field and equals b1 (f_update_tlev, gutsf.f:22-32)
    call fset(SWAP_F_B1, bt)
    call fset(SWAP_F_B12, b12)
    call fset(SWAP_F_AJ, aj)
    call fset(SWAP_F_UP, up)
    call fset(SWAP_F_NP, np)
    call fset(SWAP_F_NU, nu)
    call chk(swap_predict_B(swap_ctx_h), 'predict_B')
    call fget(SWAP_F_B1P2, b1p2)
    call fget(SWAP_F_E, E)
  end subroutine predict_B

  subroutine correct_B(b0, b1, b1p2, E, aj, up, np, nu)
    real, intent(in) :: b0(nx,ny,nz,3), b1(nx,ny,nz,3)
    real, intent(inout) :: b1p2(nx,ny,nz,3)
    real, intent(out) :: E(nx,ny,nz,3), aj(nx,ny,nz,3)
    real, intent(inout) :: up(nx,ny,nz,3)
    real, intent(in) :: np(nx,ny,nz), nu(nx,ny,nz)
    call fset(SWAP_F_B1, b1)
    call fset(SWAP_F_B1P2, b1p2)
    call fset(SWAP_F_UP, up)
    call fset(SWAP_F_NP, np)
    call fset(SWAP_F_NU, nu)
    call chk(swap_correct_B(swap_ctx_h), 'correct_B')
    call fget(SWAP_F_B1P2, b1p2)
    call fget(SWAP_F_E, E)
    call fget(SWAP_F_AJ, aj)
  end subroutine correct_B

  subroutine f_update_tlev(b1, b12, b1p2, bt, b0)
    real, intent(inout) :: b1(nx,ny,nz,3), b12(nx,ny,nz,3), bt(nx,ny,nz,3)
    real, intent(in) :: b1p2(nx,ny,nz,3), b0(nx,ny,nz,3)
    bt = b1p2
    b12 = b1
    b1 = b1p2
  end subroutine f_update_tlev

  ! maind.f:420-421 resets dtsub = dt/ntsub and ntf = ntsub before the call;
  ! the library does the same and returns the halved values
  subroutine check_time_step(bt, np)
    real, intent(in) :: bt(nx,ny,nz,3), np(nx,ny,nz)
    integer(c_int) :: n
    real(c_double) :: d
    call fset(SWAP_F_B1, bt)
    call fset(SWAP_F_NP, np)
    call chk(swap_check_time_step(swap_ctx_h, n, d), 'check_time_step')
    ntf = real(n)
    dtsub = real(d)
  end subroutine check_time_step

  ! -------- gutsp.f --------
  subroutine get_interp_weights(xp)
    real, intent(in) :: xp(:,:)
    call xp_set(xp)
    call chk(swap_get_interp_weights(swap_ctx_h), 'get_interp_weights')
  end subroutine get_interp_weights

  subroutine update_np(np)
    real, intent(out) :: np(nx,ny,nz)
    call chk(swap_update_np(swap_ctx_h), 'update_np')
    call fget(SWAP_F_NP, np)
  end subroutine update_np

  subroutine update_up(vp, np, up)
    real, intent(in) :: vp(:,:), np(nx,ny,nz)
    real, intent(out) :: up(nx,ny,nz,3)
    call pset(SWAP_P_VP, vp)
    call fset(SWAP_F_NP, np)
    call chk(swap_update_up(swap_ctx_h, SWAP_V_CUR), 'update_up')
    call fget(SWAP_F_UP, up)
  end subroutine update_up

  subroutine extrapol_up(up, vp, vp1, np)
    real, intent(out) :: up(nx,ny,nz,3)
    real, intent(in) :: vp(:,:), vp1(:,:), np(nx,ny,nz)
    call pset(SWAP_P_VP, vp)
    call pset(SWAP_P_VP1, vp1)
    call fset(SWAP_F_NP, np)
    call chk(swap_update_up(swap_ctx_h, SWAP_V_EXTRAP), 'extrapol_up')
    call fget(SWAP_F_UP, up)
  end subroutine extrapol_up

  subroutine get_Ep(Ep, aj, np, up, btc, nu)
    real, intent(inout) :: Ep(:,:)
    real, intent(in) :: aj(nx,ny,nz,3), up(nx,ny,nz,3), btc(nx,ny,nz,3)
    real, intent(in) :: np(nx,ny,nz), nu(nx,ny,nz)
    call fset(SWAP_F_AJ, aj)
    call fset(SWAP_F_UP, up)
    call fset(SWAP_F_BTC, btc)
    call chk(swap_get_Ep(swap_ctx_h), 'get_Ep')
    call pget(SWAP_P_EP, Ep)
  end subroutine get_Ep

  subroutine get_vplus_vminus(Ep, btc, vp, vplus, vminus)
    real, intent(in) :: Ep(:,:), btc(nx,ny,nz,3), vp(:,:)
    real, intent(inout) :: vplus(:,:), vminus(:,:)
    call pset(SWAP_P_EP, Ep)
    call fset(SWAP_F_BTC, btc)
    call pset(SWAP_P_VP, vp)
    call chk(swap_get_vplus_vminus(swap_ctx_h), 'get_vplus_vminus')
    call pget(SWAP_P_VPLUS, vplus)
    call pget(SWAP_P_VMINUS, vminus)
  end subroutine get_vplus_vminus

  subroutine improve_up(vp1, vplus, vminus, up, np)
    real, intent(inout) :: vp1(:,:)
    real, intent(in) :: vplus(:,:), vminus(:,:), np(nx,ny,nz)
    real, intent(out) :: up(nx,ny,nz,3)
    call pset(SWAP_P_VPLUS, vplus)
    call pset(SWAP_P_VMINUS, vminus)
    call fset(SWAP_F_NP, np)
    call chk(swap_improve_up(swap_ctx_h), 'improve_up')
    call pget(SWAP_P_VP1, vp1)
    call fget(SWAP_F_UP, up)
  end subroutine improve_up

  subroutine get_vp_final(Ep, vp, vp1, vplus)
    real, intent(in) :: Ep(:,:), vplus(:,:)
    real, intent(inout) :: vp(:,:), vp1(:,:)
    call pset(SWAP_P_EP, Ep)
    call pset(SWAP_P_VP, vp)
    call pset(SWAP_P_VPLUS, vplus)
    call chk(swap_get_vp_final(swap_ctx_h), 'get_vp_final')
    call pget(SWAP_P_VP, vp)
    call pget(SWAP_P_VP1, vp1)
  end subroutine get_vp_final

  ! -------- chem_rates.f --------
  ! the host rows l0..Ni_tot of xp, vp, vp1 from the device (ions just added)
  subroutine pull_new(l0, xp, vp, vp1)
    integer, intent(in) :: l0
    real, intent(inout) :: xp(:,:), vp(:,:), vp1(:,:)
    real(c_double), allocatable, target :: x(:,:)
    real(c_float), allocatable, target :: v(:,:)
    allocate(x(Ni_tot, 3), v(Ni_tot, 3))
    call chk(swap_get_particle_array(swap_ctx_h, SWAP_P_XP, c_loc(x)), 'get xp')
    xp(l0:Ni_tot, :) = real(x(l0:Ni_tot, :))
    call chk(swap_get_particle_array(swap_ctx_h, SWAP_P_VP, c_loc(v)), 'get vp')
    vp(l0:Ni_tot, :) = v(l0:Ni_tot, :)
    call chk(swap_get_particle_array(swap_ctx_h, SWAP_P_VP1, c_loc(v)), 'get vp1')
    vp1(l0:Ni_tot, :) = v(l0:Ni_tot, :)
  end subroutine pull_new

  ! the re-deposit that ends the sources (chem_rates.f:641-643)
  subroutine redeposit(np, up)
    real, intent(out) :: np(nx,ny,nz), up(nx,ny,nz,3)
    call chk(swap_get_interp_weights(swap_ctx_h), 'get_interp_weights')
    call chk(swap_update_np(swap_ctx_h), 'update_np')
    call fget(SWAP_F_NP, np)
    call chk(swap_update_up(swap_ctx_h, SWAP_V_CUR), 'update_up')
    call fget(SWAP_F_UP, up)
  end subroutine redeposit

  ! Ionize_sw_mp: dNi = ndot dt beta beta_pu dx dy delz nz with
  ! ndot = 1e-6 nf_init omega_p (chem_rates.f:551-562), ring-distribution
  ! protons appended (:567-628), input_p += m v / beta * beta_p as the
  ! reference accumulates it (:614-619), then the re-deposit (:641-643)
  subroutine Ionize_sw_mp(np, vp, vp1, xp, m_tstep, input_p, up)
    real, intent(inout) :: np(nx,ny,nz), vp(:,:), vp1(:,:), xp(:,:), input_p(3)
    real, intent(inout) :: up(nx,ny,nz,3)
    integer, intent(in) :: m_tstep
    integer(c_int64_t) :: inj
    integer :: l0, m
    real(8) :: ndot
    call xp_set(xp)
    call pset(SWAP_P_VP, vp)
    call pset(SWAP_P_VP1, vp1)
    ndot = 1d-6 * sw_nf_init * sw_omega_p
    src_step = m_tstep
    call chk(swap_ionize_sw(swap_ctx_h, int(m_tstep, c_int), ndot, sw_vsw, sw_beta_pu, inj), &
         'Ionize_sw_mp')
    l0 = Ni_tot + 1
    Ni_tot = Ni_tot + int(inj)
    if (inj > 0) then
       call pull_new(l0, xp, vp, vp1)
       do m = 1, 3
          input_p(m) = input_p(m) + real(mion_s * sum(real(vp(l0:Ni_tot, m), 8)) / beta_s * &
               sw_beta_pu)
       end do
    end if
    call redeposit(np, up)
  end subroutine Ionize_sw_mp

  ! Ionize_pluto_mp, the cloud form (chem_rates.f:290-503): photoionization
  ! of the neutral cloud, ions at rest in their node's cell with variable
  ! weights, then the re-deposit (:496-498); no charge exchange here
  subroutine Ionize_pluto_mp(np, vp, vp1, xp, m_tstep, input_p, up)
    real, intent(inout) :: np(nx,ny,nz), vp(:,:), vp1(:,:), xp(:,:), input_p(3)
    real, intent(inout) :: up(nx,ny,nz,3)
    integer, intent(in) :: m_tstep
    integer(c_int64_t) :: inj, nchex
    type(swap_cloud) :: cl
    integer :: l0
    call xp_set(xp)
    call pset(SWAP_P_VP, vp)
    call pset(SWAP_P_VP1, vp1)
    src_step = m_tstep
    cl = cloud_s
    cl%sigma_chex = 0d0
    call chk(swap_ionize_cloud(swap_ctx_h, int(m_tstep, c_int), cl, inj, nchex), &
         'Ionize_pluto_mp')
    l0 = Ni_tot + 1
    Ni_tot = Ni_tot + int(inj)
    if (inj > 0) call pull_new(l0, xp, vp, vp1)
    call redeposit(np, up)
  end subroutine Ionize_pluto_mp

  ! res_chex (chem_rates.f:805-843): an ion hit by charge exchange, with
  ! probability dt n_n sigma |v|, stops and becomes a pickup ion (mrat of the
  ! cloud's ions).  The neutral density is the cloud's profile (the
  ! reference's constant 10000e15 stands where its profile is commented out).
  ! The draws are keyed by the step: m_tstep, by default the step of the last
  ! source call (the reference draws from its running pad_ranf stream).
  subroutine res_chex(xp, vp, m_tstep)
    real, intent(in) :: xp(:,:)
    real, intent(inout) :: vp(:,:)
    integer, intent(in), optional :: m_tstep
    integer(c_int64_t) :: inj, nchex
    type(swap_cloud) :: cl
    integer :: st
    st = src_step
    if (present(m_tstep)) st = m_tstep
    call xp_set(xp)
    call pset(SWAP_P_VP, vp)
    cl = cloud_s
    cl%tau_photo = huge(1d0)   ! no photoionization
    call chk(swap_ionize_cloud(swap_ctx_h, int(st, c_int), cl, inj, nchex), 'res_chex')
    if (inj /= 0) call chk(-1_c_int, 'res_chex: unexpected new ions')
    call pget(SWAP_P_VP, vp)
  end subroutine res_chex

  subroutine move_ion_half(xp, vp, vp1, input_p)
    real, intent(inout) :: xp(:,:)
    real, intent(in) :: vp(:,:), vp1(:,:)
    real, intent(inout) :: input_p(3)
    call xp_set(xp)
    call pset(SWAP_P_VP, vp)
    call chk(swap_move_ion_half(swap_ctx_h), 'move_ion_half')
    call xp_get(xp)
  end subroutine move_ion_half

end module swap_ref
