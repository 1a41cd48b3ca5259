!+
! Module bmx_bmad_hooks
!
! The Bmad side of the drop-in: the two plugin hooks through which Bmad's
! own programs hand tracking to libbmx, and the flattener of a branch.  This
! file uses Bmad's modules (bmad/modules/bmad.f90), so it compiles inside a
! Bmad build only, after bmx_c_binding.f90 and bmx_shim_core.f90 (which hold
! everything that does not need Bmad and are compiled and tested on their
! own).  A program installs the hooks where it installs its own:
!
!   call bmx_hooks_init(lat, ix_branch = 0)      ! flatten + upload once
!   track1_bunch_hook_ptr => gpu_track1_bunch_hook  ! long_term_tracking.f90:18
!   track_many_hook_ptr   => gpu_track_many_hook    ! dynamic_aperture.f90:49
!
! gpu_track1_bunch_hook (track1_bunch_hook_def,
! bmad/modules/bmad_routine_interface.f90:3927-3936): called by track1_bunch
! for every element (bmad/multiparticle/beam_mod.f90:263-266).  At the turn's
! first element (the one after the bunch's start, bunch%ix_ele) it packs
! bunch%particle(:)%vec/state/beta/t, tracks the whole turn on the GPU,
! unpacks, and returns finished = .true.; for the turn's other elements it
! returns finished = .true. with no work.  With bmad_com's radiation
! switches on, bmx_hooks_init hands libbmx the maps radiation_map_setup made
! for every element (ele%rad_map%rm0 / %rm1), which libbmx applies where
! track1 does (track1.f90:198-201, :316-319); its fluctuations come from its
! own counter-based stream (seed: bmx_hooks_init's rad_seed).  Refuses
! (finished = .false., the CPU path runs) what the device does not model: a
! lattice libbmx refused at bmx_hooks_init, centroid or bunch_track output,
! and ltt's ramping and split-bend radiation points
! (lt_tracking_mod.f90:3199-3214, :3383-3446), which drive track1_bunch
! element by element.
!
! gpu_track_many_hook (track_many_hook_def, :3949-3959): the contract of the
! DA program's own hook (bsim/dynamic_aperture/track_many_hook.f90:27-106):
! orbit(ix_start) through one turn back to ix_end, track_state =
! moving_forward$ or the element where the particle was lost.  One particle
! per call is how dynamic_aperture_point drives it; a whole scan in one
! launch is bmx_track_grid / bmx_track_ics (bmx/da.py).
!-

module bmx_bmad_hooks

use bmad
use bmx_c_binding
use bmx_shim_core

implicit none

type(c_ptr), save :: bmx_hook_ctx = c_null_ptr, bmx_hook_lat = c_null_ptr
integer, save :: bmx_hook_branch = 0, bmx_hook_n_ele = 0
logical, save :: bmx_hook_ok = .false.

contains

!------------------------------------------------------------------------
! Flatten lat%branch(ix_branch)%ele(1:n_ele_track) (after bmad_parser and
! lattice_bookkeeper) and upload it.  Sets bmx_hook_ok; err for an element
! or species the device path refuses (the message is printed).
subroutine bmx_hooks_init(lat, ix_branch, err, rad_seed)
  type(lat_struct), target :: lat
  integer, optional :: ix_branch
  integer(c_int64_t), optional :: rad_seed
  logical err
  type(branch_struct), pointer :: branch
  type(ele_struct), pointer :: ele
  type(bmx_ele_rec), allocatable :: recs(:)
  type(bmx_rad_map), allocatable :: maps(:)
  type(bmx_ref) :: ref
  type(bmx_config) :: cfg
  type(bmx_radiation) :: rad
  real(rp) :: a(0:n_pole_maxx), b(0:n_pole_maxx), ae(0:n_pole_maxx), be(0:n_pole_maxx)
  integer i, rc
  logical ok, rad_on, rerr

  err = .true.
  bmx_hook_ok = .false.
  bmx_hook_branch = integer_option(0, ix_branch)
  branch => lat%branch(bmx_hook_branch)
  bmx_hook_n_ele = branch%n_ele_track
  allocate(recs(bmx_hook_n_ele))
  do i = 1, bmx_hook_n_ele
    ele => branch%ele(i)
    a = 0; b = 0; ae = 0; be = 0
    if (associated(ele%a_pole)) then
      a = ele%a_pole; b = ele%b_pole
    endif
    if (associated(ele%a_pole_elec)) then
      ae = ele%a_pole_elec; be = ele%b_pole_elec
    endif
    call bmx_flatten_element(ele%key, ele%value(1:bmad_num_ele_attrib), real(ele%ref_time, rp), ele%is_on, &
                             ele%orientation, ele%field_master, ele%scale_multipoles, ele%aperture_type, &
                             ele%aperture_at, associated(ele%a_pole), a, b, associated(ele%a_pole_elec), ae, be, &
                             recs(i), ok)
    if (.not. ok) then
      print '(a, i0, 2a)', 'bmx_hooks_init: element ', i, ' has a key libbmx does not track: ', trim(ele%name)
      return
    endif
  enddo
  ref%species = bmx_species_from_bmad(branch%param%particle)
  if (ref%species == 0) return
  cfg%aperture_limit_on = merge(1, 0, bmad_com%aperture_limit_on)
  cfg%max_aperture_limit = bmad_com%max_aperture_limit
  cfg%absolute_time_tracking = merge(1, 0, bmad_com%absolute_time_tracking)
  if (.not. c_associated(bmx_hook_ctx)) then
    rc = bmx_ctx_create(bmx_hook_ctx, 0, 0)
    if (rc /= 0) return
  endif
  if (c_associated(bmx_hook_lat)) call bmx_lattice_destroy(bmx_hook_lat)
  rad_on = bmad_com%radiation_damping_on .or. bmad_com%radiation_fluctuations_on
  if (rad_on) then
    ! the maps track1_radiation would use (radiation_mod.f90:95-109), set up
    ! where they are stale; C reads the matrices row-major, hence transpose
    allocate(maps(2 * bmx_hook_n_ele))
    do i = 1, bmx_hook_n_ele
      ele => branch%ele(i)
      if (ele%value(l$) == 0) cycle
      rerr = .false.
      if (.not. associated(ele%rad_map)) then
        call radiation_map_setup(ele, rerr)
      elseif (ele%rad_map%stale) then
        call radiation_map_setup(ele, rerr)
      endif
      if (rerr .or. .not. associated(ele%rad_map)) cycle
      maps(2*i-1)%ref_orb = ele%rad_map%rm0%ref_orb
      maps(2*i-1)%damp_dmat = transpose(ele%rad_map%rm0%damp_dmat)
      maps(2*i-1)%xfer_damp_vec = ele%rad_map%rm0%xfer_damp_vec
      maps(2*i-1)%stoc_mat = transpose(ele%rad_map%rm0%stoc_mat)
      maps(2*i)%ref_orb = ele%rad_map%rm1%ref_orb
      maps(2*i)%damp_dmat = transpose(ele%rad_map%rm1%damp_dmat)
      maps(2*i)%xfer_damp_vec = ele%rad_map%rm1%xfer_damp_vec
      maps(2*i)%stoc_mat = transpose(ele%rad_map%rm1%stoc_mat)
    enddo
    rad%damping_on = merge(1, 0, bmad_com%radiation_damping_on)
    rad%fluctuations_on = merge(1, 0, bmad_com%radiation_fluctuations_on)
    rad%zero_average = merge(1, 0, bmad_com%radiation_zero_average)
    rad%synch_rad_scale = bmad_com%synch_rad_scale
    if (present(rad_seed)) rad%seed = rad_seed
    rc = bmx_lattice_upload_rad(bmx_hook_ctx, recs, int(bmx_hook_n_ele, c_int32_t), c_null_ptr, 1, maps, rad, &
                                ref, cfg, bmx_hook_lat)
  else
    rc = bmx_lattice_upload(bmx_hook_ctx, recs, int(bmx_hook_n_ele, c_int32_t), c_null_ptr, 1, ref, cfg, bmx_hook_lat)
  endif
  if (rc /= 0) then
    print '(2a)', 'bmx_hooks_init: ', bmx_error_string(bmx_hook_ctx)
    return
  endif
  bmx_hook_ok = .true.
  err = .false.
end subroutine

!------------------------------------------------------------------------
subroutine gpu_track1_bunch_hook(bunch, ele, err, centroid, direction, finished, bunch_track)
  type(bunch_struct), target :: bunch
  type(ele_struct), target :: ele
  type(coord_struct), optional :: centroid(0:)
  type(bunch_track_struct), optional :: bunch_track
  integer, optional :: direction
  logical err, finished
  type(bmx_shim_bunch), target :: b
  type(bmx_stats) :: stats
  real(rp), allocatable :: vec(:,:), beta(:), t(:)
  integer, allocatable :: state(:)
  integer n, i, dir, ix0, first, rc

  err = .false.
  finished = .false.
  if (.not. bmx_hook_ok .or. present(centroid) .or. present(bunch_track)) return
  if (ele%ix_branch /= bmx_hook_branch) return
  dir = integer_option(1, direction)
  ! the turn's first element (forward: the one after the bunch's start;
  ! backward: the start itself, beam_mod.f90:131-147) runs the whole turn;
  ! the others have nothing left to do (a finished hook leaves bunch%ix_ele
  ! at the start, beam_mod.f90:263-266)
  ix0 = bunch%ix_ele
  if (dir == 1) then
    first = modulo(ix0, bmx_hook_n_ele) + 1
  else
    first = merge(ix0, bmx_hook_n_ele, ix0 > 0)
  endif
  if (ele%ix_ele /= first) then
    finished = .true.
    return
  endif
  n = size(bunch%particle)
  allocate(vec(6, n), beta(n), t(n), state(n))
  do i = 1, n
    vec(:, i) = bunch%particle(i)%vec
    beta(i) = bunch%particle(i)%beta
    t(i) = real(bunch%particle(i)%t, rp)
    state(i) = bunch%particle(i)%state
  enddo
  call bmx_shim_pack(vec, state, beta, t, b)
  rc = bmx_shim_track(bmx_hook_ctx, bmx_hook_lat, ix0, ix0, 1, dir, b, stats, .true.)
  if (rc /= 0) then
    err = .true.
    return
  endif
  call bmx_shim_unpack(b, vec, state, beta, t)
  do i = 1, n
    bunch%particle(i)%vec = vec(:, i)
    bunch%particle(i)%beta = beta(i)
    bunch%particle(i)%t = t(i)
    bunch%particle(i)%state = state(i)
  enddo
  finished = .true.
end subroutine

!------------------------------------------------------------------------
subroutine gpu_track_many_hook(finished, lat, orbit, ix_start, ix_end, direction, ix_branch, track_state)
  type(lat_struct), target :: lat
  type(coord_struct) orbit(0:)
  integer ix_start, ix_end, direction
  integer, optional :: ix_branch, track_state
  logical finished
  real(c_double) :: v(6), beta
  integer state, ts, rc

  finished = .false.
  if (.not. bmx_hook_ok .or. direction /= 1 .or. ix_start /= ix_end) return
  if (integer_option(0, ix_branch) /= bmx_hook_branch) return
  v = orbit(ix_start)%vec
  beta = orbit(ix_start)%beta
  state = orbit(ix_start)%state
  rc = bmx_shim_track_one(bmx_hook_ctx, bmx_hook_lat, ix_start, 1, v, beta, state, ts)
  if (rc /= 0) return
  orbit(ix_end) = orbit(ix_start)
  orbit(ix_end)%vec = v
  orbit(ix_end)%beta = beta
  orbit(ix_end)%state = state
  if (present(track_state)) track_state = ts
  finished = .true.
end subroutine

end module
