!+
! Module bmx_shim_core
!
! The Bmad-independent half of the drop-in shim: what the Bmad hooks
! (bmx_bmad_hooks.f90) need that does not need Bmad's own modules.
!
!   bmx_flatten_element  one element, given as Bmad keeps it (ele%key,
!                        ele%value(1:num_ele_attrib$), the ele_struct flags,
!                        ele%a_pole/b_pole and their electric twins), into a
!                        bmx_ele_rec -- the flattener of the shim
!                        (bmad/code/attribute_bookkeeper.f90:296-404 leaves
!                        every value the record needs in ele%value)
!   bmx_shim_pack /      bunch%particle(:)%vec/state/beta/t (array of
!   bmx_shim_unpack      structures) to the structure of arrays libbmx tracks
!                        and back (bmad/modules/bmad_struct.f90:546-575)
!   bmx_shim_track       bmx_track_bunch on a packed bunch
!   bmx_shim_track_one   one particle through whole turns: the
!                        track_many_hook_def contract (track_state =
!                        moving_forward$ or the element where it was lost)
!
! The Bmad enumerations below restate the values of bmad_struct.f90 and
! particle_species_mod.f90 (cited at each block); they are what a Bmad
! program's ele%key, ele%value(:) and species integers mean.
!-

module bmx_shim_core

use, intrinsic :: iso_c_binding
use bmx_c_binding

implicit none

! element keys (bmad/modules/bmad_struct.f90:1681-1693)
integer, parameter :: bmad_drift = 1, bmad_sbend = 2, bmad_quadrupole = 3, bmad_sextupole = 5, &
                      bmad_rfcavity = 9, bmad_marker = 14, bmad_kicker = 15, bmad_octupole = 17, &
                      bmad_multipole = 19, bmad_ab_multipole = 22, bmad_monitor = 31, bmad_instrument = 32, &
                      bmad_hkicker = 33, bmad_vkicker = 34, bmad_rcollimator = 35, bmad_ecollimator = 36, &
                      bmad_pipe = 44, bmad_thick_multipole = 67

! ele%value(:) indices (bmad/modules/bmad_struct.f90:1775-1881); several
! share an index and mean different things for different keys
integer, parameter :: bmad_num_ele_attrib = 75
integer, parameter :: iv_l = 1, iv_ref_tilt = 3, iv_kick = 3, iv_k1 = 4, iv_k2 = 5, iv_k3 = 6, iv_g = 6, &
                      iv_dg = 7, iv_voltage = 8, iv_fringe_type = 10, iv_fringe_at = 11, &
                      iv_exact_multipoles = 14, iv_r0_elec = 15, iv_rf_frequency = 15, iv_r0_mag = 16, &
                      iv_e1 = 19, iv_e2 = 20, iv_fint = 21, iv_fintx = 22, iv_hgap = 23, iv_hgapx = 24, &
                      iv_phi0 = 24, iv_px_aperture_width2 = 24, iv_px_aperture_center = 25, &
                      iv_py_aperture_width2 = 26, iv_phi0_multipass = 26, iv_py_aperture_center = 27, &
                      iv_z_aperture_width2 = 28, iv_z_aperture_center = 29, iv_pz_aperture_width2 = 30, &
                      iv_pz_aperture_center = 31, iv_hkick = 39, iv_vkick = 40, iv_coupler_phase = 44, &
                      iv_coupler_angle = 45, iv_coupler_strength = 46, iv_coupler_at = 47, &
                      iv_delta_ref_time = 50, iv_p0c_start = 51, iv_e_tot_start = 52, iv_p0c = 53, &
                      iv_e_tot = 54, iv_x_pitch_tot = 55, iv_y_pitch_tot = 56, iv_x_offset_tot = 57, &
                      iv_y_offset_tot = 58, iv_z_offset_tot = 59, iv_tilt_tot = 60, iv_ref_tilt_tot = 61, &
                      iv_num_steps = 66, iv_x1_limit = 71, iv_x2_limit = 72, iv_y1_limit = 73, iv_y2_limit = 74

! species (sim_utils/interfaces/particle_species_mod.f90:60-72)
integer, parameter :: bmad_positron = 1, bmad_electron = -1, bmad_proton = 2, bmad_antiproton = -2, &
                      bmad_antimuon = 3, bmad_muon = -3

! track_many's track_state for a particle that survived (bmad/modules/bmad_struct.f90)
integer, parameter :: bmad_moving_forward = -9

! A bunch in libbmx's structure of arrays (the arrays the bmx_particles of
! bmx_shim_track points to)
type bmx_shim_bunch
  real(c_double), allocatable :: x(:), px(:), y(:), py(:), z(:), pz(:), beta(:), t(:)
  integer(c_int32_t), allocatable :: state(:), ix_lost(:), turn_lost(:)
end type

contains

!------------------------------------------------------------------------
! libbmx key of a Bmad key, -1 for keys the device path does not track
function bmx_key_from_bmad(key) result(k)
  integer, intent(in) :: key
  integer(c_int32_t) :: k
  select case (key)
  case (bmad_marker);          k = bmx_key_marker
  case (bmad_drift);           k = bmx_key_drift
  case (bmad_quadrupole);      k = bmx_key_quadrupole
  case (bmad_sbend);           k = bmx_key_sbend
  case (bmad_sextupole);       k = bmx_key_sextupole
  case (bmad_octupole);        k = bmx_key_octupole
  case (bmad_thick_multipole); k = bmx_key_thick_multipole
  case (bmad_multipole);       k = bmx_key_multipole
  case (bmad_ab_multipole);    k = bmx_key_ab_multipole
  case (bmad_rfcavity);        k = bmx_key_rfcavity
  case (bmad_hkicker);         k = bmx_key_hkicker
  case (bmad_vkicker);         k = bmx_key_vkicker
  case (bmad_kicker);          k = bmx_key_kicker
  case (bmad_instrument);      k = bmx_key_instrument
  case (bmad_monitor);         k = bmx_key_monitor
  case (bmad_ecollimator);     k = bmx_key_ecollimator
  case (bmad_rcollimator);     k = bmx_key_rcollimator
  case (bmad_pipe);            k = bmx_key_pipe
  case default;                k = -1
  end select
end function

! libbmx species of a Bmad species, 0 for others
function bmx_species_from_bmad(species) result(s)
  integer, intent(in) :: species
  integer(c_int32_t) :: s
  select case (species)
  case (bmad_electron);   s = bmx_electron
  case (bmad_positron);   s = bmx_positron
  case (bmad_proton);     s = bmx_proton
  case (bmad_antiproton); s = bmx_antiproton
  case (bmad_muon);       s = bmx_muon
  case (bmad_antimuon);   s = bmx_antimuon
  case default;           s = 0
  end select
end function

!------------------------------------------------------------------------
! One element as Bmad keeps it -> bmx_ele_rec.  value: ele%value(1:75);
! ref_time: ele%ref_time; the poles: ele%a_pole(0:21) etc., has_mag /
! has_elec: associated(ele%a_pole) / associated(ele%a_pole_elec) (for a
! multipole element a_pole holds KnL and b_pole Tn, as in Bmad).  ok is
! .false. for a key the device path does not track.
subroutine bmx_flatten_element(key, value, ref_time, is_on, orientation, field_master, scale_multipoles, &
                               aperture_type, aperture_at, has_mag, a_pole, b_pole, has_elec, a_pole_elec, &
                               b_pole_elec, rec, ok)
  integer, intent(in) :: key, orientation, aperture_type, aperture_at
  real(c_double), intent(in) :: value(bmad_num_ele_attrib), ref_time
  logical, intent(in) :: is_on, field_master, scale_multipoles, has_mag, has_elec
  real(c_double), intent(in) :: a_pole(0:bmx_n_pole-1), b_pole(0:bmx_n_pole-1)
  real(c_double), intent(in) :: a_pole_elec(0:bmx_n_pole-1), b_pole_elec(0:bmx_n_pole-1)
  type(bmx_ele_rec), intent(out) :: rec
  logical, intent(out) :: ok
  integer(c_int32_t) :: k

  k = bmx_key_from_bmad(key)
  ok = (k >= 0)
  if (.not. ok) return
  rec%key = k
  rec%is_on = merge(1, 0, is_on)
  rec%orientation = orientation
  rec%field_master = merge(1, 0, field_master)
  rec%scale_multipoles = merge(1, 0, scale_multipoles)
  rec%aperture_type = aperture_type
  rec%aperture_at = aperture_at
  rec%num_steps = max(1, nint(value(iv_num_steps)))
  rec%exact_multipoles = nint(value(iv_exact_multipoles))
  if (rec%exact_multipoles == 0) rec%exact_multipoles = 1  ! off$
  rec%fringe_type = nint(value(iv_fringe_type))
  rec%fringe_at = nint(value(iv_fringe_at))
  rec%l = value(iv_l)
  rec%p0c = value(iv_p0c);  rec%e_tot = value(iv_e_tot)
  rec%p0c_start = value(iv_p0c_start);  rec%e_tot_start = value(iv_e_tot_start)
  rec%delta_ref_time = value(iv_delta_ref_time);  rec%ref_time = ref_time
  rec%x_offset = value(iv_x_offset_tot);  rec%y_offset = value(iv_y_offset_tot);  rec%z_offset = value(iv_z_offset_tot)
  rec%x_pitch = value(iv_x_pitch_tot);  rec%y_pitch = value(iv_y_pitch_tot)
  rec%x1_limit = value(iv_x1_limit);  rec%x2_limit = value(iv_x2_limit)
  rec%y1_limit = value(iv_y1_limit);  rec%y2_limit = value(iv_y2_limit)
  rec%hkick = value(iv_hkick);  rec%vkick = value(iv_vkick)
  rec%has_mag_poles = merge(1, 0, has_mag)
  rec%has_elec_poles = merge(1, 0, has_elec)
  if (has_mag) then
    rec%a = a_pole;  rec%b = b_pole
  endif
  if (has_elec) then
    rec%a_elec = a_pole_elec;  rec%b_elec = b_pole_elec
  endif
  select case (key)
  case (bmad_sbend)
    rec%g = value(iv_g);  rec%dg = value(iv_dg);  rec%k1 = value(iv_k1);  rec%k2 = value(iv_k2)
    rec%e1 = value(iv_e1);  rec%e2 = value(iv_e2)
    rec%fint = value(iv_fint);  rec%fintx = value(iv_fintx);  rec%hgap = value(iv_hgap);  rec%hgapx = value(iv_hgapx)
    rec%ref_tilt = value(iv_ref_tilt_tot)
    rec%r0_mag = value(iv_r0_mag);  rec%r0_elec = value(iv_r0_elec)
  case (bmad_rfcavity)
    rec%voltage = value(iv_voltage);  rec%rf_frequency = value(iv_rf_frequency)
    rec%phi0 = value(iv_phi0);  rec%phi0_multipass = value(iv_phi0_multipass)
    rec%coupler_strength = value(iv_coupler_strength);  rec%coupler_angle = value(iv_coupler_angle)
    rec%coupler_phase = value(iv_coupler_phase);  rec%coupler_at = nint(value(iv_coupler_at))
    rec%tilt = value(iv_tilt_tot)
  case (bmad_hkicker, bmad_vkicker)
    rec%kick = value(iv_kick);  rec%tilt = value(iv_tilt_tot)
    rec%r0_mag = value(iv_r0_mag);  rec%r0_elec = value(iv_r0_elec)
  case (bmad_ecollimator, bmad_rcollimator)
    rec%tilt = value(iv_tilt_tot)
    rec%px_aperture_width2 = value(iv_px_aperture_width2);  rec%px_aperture_center = value(iv_px_aperture_center)
    rec%py_aperture_width2 = value(iv_py_aperture_width2);  rec%py_aperture_center = value(iv_py_aperture_center)
    rec%z_aperture_width2 = value(iv_z_aperture_width2);  rec%z_aperture_center = value(iv_z_aperture_center)
    rec%pz_aperture_width2 = value(iv_pz_aperture_width2);  rec%pz_aperture_center = value(iv_pz_aperture_center)
    rec%r0_mag = value(iv_r0_mag);  rec%r0_elec = value(iv_r0_elec)
  case default
    rec%k1 = value(iv_k1);  rec%k2 = value(iv_k2);  rec%k3 = value(iv_k3)
    rec%tilt = value(iv_tilt_tot)
    rec%r0_mag = value(iv_r0_mag);  rec%r0_elec = value(iv_r0_elec)
  end select
end subroutine

!------------------------------------------------------------------------
! bunch%particle(1:n) -> structure of arrays.  vec(6, n), state(n), beta(n),
! t(n) are the particle records' vec, state, beta and t.
subroutine bmx_shim_pack(vec, state, beta, t, b)
  real(c_double), intent(in) :: vec(:,:), beta(:), t(:)
  integer, intent(in) :: state(:)
  type(bmx_shim_bunch), intent(inout) :: b
  integer n
  n = size(vec, 2)
  if (allocated(b%x)) deallocate(b%x, b%px, b%y, b%py, b%z, b%pz, b%beta, b%t, b%state, b%ix_lost, b%turn_lost)
  allocate(b%x(n), b%px(n), b%y(n), b%py(n), b%z(n), b%pz(n), b%beta(n), b%t(n), b%state(n), b%ix_lost(n), &
           b%turn_lost(n))
  b%x = vec(1,:);  b%px = vec(2,:);  b%y = vec(3,:);  b%py = vec(4,:);  b%z = vec(5,:);  b%pz = vec(6,:)
  b%beta = beta;  b%t = t;  b%state = state
  b%ix_lost = -1;  b%turn_lost = -1
end subroutine

subroutine bmx_shim_unpack(b, vec, state, beta, t)
  type(bmx_shim_bunch), intent(in) :: b
  real(c_double), intent(inout) :: vec(:,:), beta(:), t(:)
  integer, intent(inout) :: state(:)
  vec(1,:) = b%x;  vec(2,:) = b%px;  vec(3,:) = b%y;  vec(4,:) = b%py;  vec(5,:) = b%z;  vec(6,:) = b%pz
  beta = b%beta;  t = b%t;  state = b%state
end subroutine

!------------------------------------------------------------------------
! bmx_track_bunch on a packed bunch (host arrays; with_time: the particle
! times go in and come back).
function bmx_shim_track(ctx, lat, ix_start, ix_end, n_turns, direction, b, stats, with_time) result(rc)
  type(c_ptr), intent(in) :: ctx, lat
  integer, intent(in) :: ix_start, ix_end, n_turns, direction
  type(bmx_shim_bunch), target, intent(inout) :: b
  type(bmx_stats), intent(out) :: stats
  logical, intent(in) :: with_time
  integer(c_int32_t) :: rc
  type(bmx_particles) :: p
  p%x = c_loc(b%x);  p%px = c_loc(b%px);  p%y = c_loc(b%y);  p%py = c_loc(b%py);  p%z = c_loc(b%z)
  p%pz = c_loc(b%pz);  p%beta = c_loc(b%beta);  p%state = c_loc(b%state)
  p%ix_lost = c_loc(b%ix_lost);  p%turn_lost = c_loc(b%turn_lost)
  if (with_time) p%t = c_loc(b%t)
  p%n = size(b%x, kind = c_int64_t)
  p%on_device = 0
  rc = bmx_track_bunch(ctx, lat, int(ix_start, c_int32_t), int(ix_end, c_int32_t), int(n_turns, c_int32_t), &
                       int(direction, c_int32_t), p, stats)
end function

!------------------------------------------------------------------------
! One particle (vec, beta, alive on entry) through n_turns whole turns from
! element ix_start: track_state = bmad_moving_forward if it survives, else
! the Bmad index (1-based) of the element where it was lost; the
! track_many_hook_def contract (bmad/modules/bmad_routine_interface.f90:
! 3949-3959, bsim/dynamic_aperture/track_many_hook.f90:94-102).
function bmx_shim_track_one(ctx, lat, ix_start, n_turns, vec, beta, state, track_state) result(rc)
  type(c_ptr), intent(in) :: ctx, lat
  integer, intent(in) :: ix_start, n_turns
  real(c_double), intent(inout) :: vec(6), beta
  integer, intent(inout) :: state
  integer, intent(out) :: track_state
  integer(c_int32_t) :: rc
  type(bmx_shim_bunch), target :: b
  type(bmx_stats) :: stats
  real(c_double) :: v2(6, 1), be(1), tt(1)
  integer :: st(1)
  v2(:, 1) = vec;  be(1) = beta;  tt(1) = 0;  st(1) = state
  call bmx_shim_pack(v2, st, be, tt, b)
  rc = bmx_shim_track(ctx, lat, ix_start, ix_start, n_turns, 1, b, stats, .false.)
  call bmx_shim_unpack(b, v2, st, be, tt)
  vec = v2(:, 1);  beta = be(1);  state = st(1)
  if (state == bmx_alive) then
    track_state = bmad_moving_forward
  else
    track_state = b%ix_lost(1) + 1
  endif
end function

end module
