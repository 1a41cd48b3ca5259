!+
! Module bmx_c_binding
!
! ISO_C_BINDING interface to libbmx (include/bmx.h): interoperable mirrors of
! the ABI structs and bind(C) interfaces of every entry point.  A Bmad program
! uses it to hand bunches to the MI355X tracker from the plugin hooks
! track1_bunch_hook_ptr (bmad/modules/bmad_routine_interface.f90:3927-3936)
! and track_many_hook_ptr (:3949-3959).  This module depends on nothing from
! Bmad, so it compiles on its own; bmx_bmad_hooks.f90 holds the Bmad-side
! adapters.
!-

module bmx_c_binding

use, intrinsic :: iso_c_binding

implicit none

integer(c_int32_t), parameter :: bmx_n_pole = 22
integer(c_int32_t), parameter :: bmx_noalloc_max_particles = 65536
integer(c_int32_t), parameter :: bmx_attr_k1 = 1, bmx_attr_k2 = 2, bmx_attr_k3 = 3

! particle states (bmad/modules/bmad_struct.f90:520-527)
integer(c_int32_t), parameter :: bmx_alive = 1, bmx_lost = 2, bmx_lost_neg_x = 3, bmx_lost_pos_x = 4, &
                                 bmx_lost_neg_y = 5, bmx_lost_pos_y = 6, bmx_lost_z = 7, bmx_lost_pz = 8

! element keys
integer(c_int32_t), parameter :: bmx_key_marker = 0, bmx_key_drift = 1, bmx_key_quadrupole = 2, &
     bmx_key_sbend = 3, bmx_key_sextupole = 4, bmx_key_octupole = 5, bmx_key_thick_multipole = 6, &
     bmx_key_multipole = 7, bmx_key_ab_multipole = 8, bmx_key_rfcavity = 9, bmx_key_hkicker = 10, &
     bmx_key_vkicker = 11, bmx_key_kicker = 12, bmx_key_instrument = 13, bmx_key_monitor = 14, &
     bmx_key_ecollimator = 15, bmx_key_rcollimator = 16, bmx_key_pipe = 17

integer(c_int32_t), parameter :: bmx_electron = 1, bmx_positron = 2, bmx_proton = 3, bmx_antiproton = 4, &
                                 bmx_muon = 5, bmx_antimuon = 6

type, bind(C) :: bmx_ele_rec
  integer(c_int32_t) :: key = 0, is_on = 1, orientation = 1, num_steps = 1
  integer(c_int32_t) :: fringe_type = 1, fringe_at = 3, aperture_type = 2, aperture_at = 2
  integer(c_int32_t) :: field_master = 0, scale_multipoles = 1, has_mag_poles = 0, has_elec_poles = 0
  integer(c_int32_t) :: coupler_at = 2, exact_multipoles = 1, reserved(2) = 0
  real(c_double) :: l = 0, k1 = 0, k2 = 0, k3 = 0, g = 0, dg = 0, e1 = 0, e2 = 0, fint = 0, fintx = 0, hgap = 0, hgapx = 0
  real(c_double) :: tilt = 0, ref_tilt = 0
  real(c_double) :: x_offset = 0, y_offset = 0, z_offset = 0, x_pitch = 0, y_pitch = 0
  real(c_double) :: hkick = 0, vkick = 0, kick = 0
  real(c_double) :: voltage = 0, rf_frequency = 0, phi0 = 0, phi0_multipass = 0
  real(c_double) :: coupler_strength = 0, coupler_angle = 0, coupler_phase = 0
  real(c_double) :: x1_limit = 0, x2_limit = 0, y1_limit = 0, y2_limit = 0
  real(c_double) :: p0c = 0, e_tot = 0, p0c_start = 0, e_tot_start = 0, delta_ref_time = 0, ref_time = 0
  real(c_double) :: r0_mag = 0, r0_elec = 0
  real(c_double) :: px_aperture_width2 = 0, px_aperture_center = 0, py_aperture_width2 = 0, py_aperture_center = 0
  real(c_double) :: z_aperture_width2 = 0, z_aperture_center = 0, pz_aperture_width2 = 0, pz_aperture_center = 0
  real(c_double) :: a(0:bmx_n_pole-1) = 0, b(0:bmx_n_pole-1) = 0
  real(c_double) :: a_elec(0:bmx_n_pole-1) = 0, b_elec(0:bmx_n_pole-1) = 0
end type

type, bind(C) :: bmx_ref
  integer(c_int32_t) :: species = bmx_electron, tracked_species = 0
end type

type, bind(C) :: bmx_config
  integer(c_int32_t) :: aperture_limit_on = 1, direction = 0
  real(c_double) :: max_aperture_limit = 1d3
  integer(c_int32_t) :: absolute_time_tracking = 0, reserved = 0
end type

type, bind(C) :: bmx_particles
  type(c_ptr) :: x = c_null_ptr, px = c_null_ptr, y = c_null_ptr, py = c_null_ptr, z = c_null_ptr, pz = c_null_ptr
  type(c_ptr) :: beta = c_null_ptr, state = c_null_ptr, ix_lost = c_null_ptr, turn_lost = c_null_ptr
  type(c_ptr) :: t = c_null_ptr
  integer(c_int64_t) :: n = 0
  integer(c_int32_t) :: on_device = 0, reserved = 0
end type

type, bind(C) :: bmx_stats
  integer(c_int64_t) :: n_live = 0, n_lost = 0
  real(c_double) :: kernel_ms = 0, pet = 0
end type

type, bind(C) :: bmx_grid
  integer(c_int32_t) :: n_angle = 0, n_amp = 0, n_pz = 0, turns_per_block = 0
  real(c_double) :: amp_max_x = 0, amp_ratio_y = 1, pz_min = 0, pz_max = 0, angle_min = 0, angle_max = 0
end type

type, bind(C) :: bmx_schedule
  integer(c_int32_t) :: turns_per_block = 32, graph_depth = 1
  real(c_double) :: probe_pet_per_s = 0
end type

! rad_map_struct (bmad/modules/bmad_struct.f90:974-980) without xfer_damp_mat;
! C row-major matrices: damp_dmat(c, r) here holds the reference's damp_dmat(r, c),
! so pass transpose(ele%rad_map%rm0%damp_dmat) (see bmx_bmad_hooks.f90)
type, bind(C) :: bmx_rad_map
  real(c_double) :: ref_orb(6) = 0
  real(c_double) :: damp_dmat(6, 6) = 0
  real(c_double) :: xfer_damp_vec(6) = 0
  real(c_double) :: stoc_mat(6, 6) = 0
end type

type, bind(C) :: bmx_radiation
  integer(c_int32_t) :: damping_on = 0, fluctuations_on = 0, zero_average = 0, reserved = 0
  real(c_double) :: synch_rad_scale = 1
  integer(c_int64_t) :: seed = 0
end type

interface
  function bmx_abi_version() bind(C, name = 'bmx_abi_version') result(v)
    import
    integer(c_int32_t) :: v
  end function

  function bmx_sizeof(which) bind(C, name = 'bmx_sizeof') result(s)
    import
    integer(c_int32_t), value :: which
    integer(c_int32_t) :: s
  end function

  function bmx_ctx_create(ctx, device, flags) bind(C, name = 'bmx_ctx_create') result(rc)
    import
    type(c_ptr) :: ctx
    integer(c_int32_t), value :: device
    integer(c_int32_t), value :: flags
    integer(c_int32_t) :: rc
  end function

  subroutine bmx_ctx_destroy(ctx) bind(C, name = 'bmx_ctx_destroy')
    import
    type(c_ptr), value :: ctx
  end subroutine

  function bmx_last_error(ctx) bind(C, name = 'bmx_last_error') result(msg)
    import
    type(c_ptr), value :: ctx
    type(c_ptr) :: msg
  end function

  function bmx_lattice_upload_rad(ctx, eles, n_ele_total, variant_offsets, n_variants, maps, rad, ref, cfg, lat) &
                                  bind(C, name = 'bmx_lattice_upload_rad') result(rc)
    import
    type(c_ptr), value :: ctx
    type(bmx_ele_rec) :: eles(*)
    integer(c_int32_t), value :: n_ele_total
    type(c_ptr), value :: variant_offsets
    integer(c_int32_t), value :: n_variants
    type(bmx_rad_map) :: maps(*)
    type(bmx_radiation) :: rad
    type(bmx_ref) :: ref
    type(bmx_config) :: cfg
    type(c_ptr) :: lat
    integer(c_int32_t) :: rc
  end function

  function bmx_lattice_set_rad_stream(lat, seed, next_call) bind(C, name = 'bmx_lattice_set_rad_stream') result(rc)
    import
    type(c_ptr), value :: lat
    integer(c_int64_t), value :: seed, next_call
    integer(c_int32_t) :: rc
  end function

  function bmx_lattice_upload(ctx, eles, n_ele_total, variant_offsets, n_variants, ref, cfg, lat) &
                              bind(C, name = 'bmx_lattice_upload') result(rc)
    import
    type(c_ptr), value :: ctx
    type(bmx_ele_rec) :: eles(*)
    integer(c_int32_t), value :: n_ele_total
    type(c_ptr), value :: variant_offsets
    integer(c_int32_t), value :: n_variants
    type(bmx_ref) :: ref
    type(bmx_config) :: cfg
    type(c_ptr) :: lat
    integer(c_int32_t) :: rc
  end function

  function bmx_lattice_upload_scaled(ctx, eles, n_ele, n_variants, attr, sel, n_sel, factors, ref, cfg, lat) &
                                     bind(C, name = 'bmx_lattice_upload_scaled') result(rc)
    import
    type(c_ptr), value :: ctx
    type(bmx_ele_rec) :: eles(*)
    integer(c_int32_t), value :: n_ele, n_variants, attr
    integer(c_int32_t) :: sel(*)
    integer(c_int32_t), value :: n_sel
    real(c_double) :: factors(*)
    type(bmx_ref) :: ref
    type(bmx_config) :: cfg
    type(c_ptr) :: lat
    integer(c_int32_t) :: rc
  end function

  subroutine bmx_lattice_destroy(lat) bind(C, name = 'bmx_lattice_destroy')
    import
    type(c_ptr), value :: lat
  end subroutine

  function bmx_lattice_check(eles, n_ele, ref, cfg, n_ops, msg, msg_len) &
                             bind(C, name = 'bmx_lattice_check') result(rc)
    import
    type(bmx_ele_rec) :: eles(*)
    integer(c_int32_t), value :: n_ele
    type(bmx_ref) :: ref
    type(bmx_config) :: cfg
    integer(c_int64_t), intent(out) :: n_ops
    character(kind=c_char) :: msg(*)
    integer(c_int32_t), value :: msg_len
    integer(c_int32_t) :: rc
  end function

  function bmx_lattice_cells(lat, cell_len, n_cells) bind(C, name = 'bmx_lattice_cells') result(rc)
    import
    type(c_ptr), value :: lat
    integer(c_int32_t), intent(out) :: cell_len
    integer(c_int64_t), intent(out) :: n_cells
    integer(c_int32_t) :: rc
  end function

  function bmx_track_bunch(ctx, lat, ix_start, ix_end, n_turns, direction, p, stats) &
                           bind(C, name = 'bmx_track_bunch') result(rc)
    import
    type(c_ptr), value :: ctx, lat
    integer(c_int32_t), value :: ix_start, ix_end, n_turns, direction
    type(bmx_particles) :: p
    type(bmx_stats) :: stats
    integer(c_int32_t) :: rc
  end function

  function bmx_track_bunch_noalloc(ctx, lat, ix_start, ix_end, n_turns, direction, p, stats) &
                                   bind(C, name = 'bmx_track_bunch_noalloc') result(rc)
    import
    type(c_ptr), value :: ctx, lat
    integer(c_int32_t), value :: ix_start, ix_end, n_turns, direction
    type(bmx_particles) :: p
    type(bmx_stats) :: stats
    integer(c_int32_t) :: rc
  end function

  function bmx_track_multi(ctx, lat, part_offsets, n_turns, p, stats) bind(C, name = 'bmx_track_multi') result(rc)
    import
    type(c_ptr), value :: ctx, lat
    integer(c_int64_t) :: part_offsets(*)
    integer(c_int32_t), value :: n_turns
    type(bmx_particles) :: p
    type(bmx_stats) :: stats
    integer(c_int32_t) :: rc
  end function

  function bmx_track_grid(ctx, lat, grid, ref_orbit, n_turns, survival_turn, da_area, stats) &
                          bind(C, name = 'bmx_track_grid') result(rc)
    import
    type(c_ptr), value :: ctx, lat
    type(bmx_grid) :: grid
    type(c_ptr), value :: ref_orbit      ! c_null_ptr or c_loc of real(c_double) :: orb(6, n_pz, n_variants)
    integer(c_int32_t), value :: n_turns
    integer(c_int32_t) :: survival_turn(*)
    real(c_double) :: da_area(*)
    type(bmx_stats) :: stats
    integer(c_int32_t) :: rc
  end function

  function bmx_track_ics(ctx, lat, ics, ic_offsets, n_turns, turns_per_block, survival_turn, stats) &
                         bind(C, name = 'bmx_track_ics') result(rc)
    import
    type(c_ptr), value :: ctx, lat
    real(c_double) :: ics(*)              ! (n, 6): x(1:n), px(1:n), ... pz(1:n)
    integer(c_int64_t) :: ic_offsets(*)
    integer(c_int32_t), value :: n_turns, turns_per_block
    integer(c_int32_t) :: survival_turn(*)
    type(bmx_stats) :: stats
    integer(c_int32_t) :: rc
  end function

  function bmx_bunch_moments(ctx, p, moments) bind(C, name = 'bmx_bunch_moments') result(rc)
    import
    type(c_ptr), value :: ctx
    type(bmx_particles) :: p
    real(c_double) :: moments(28)
    integer(c_int32_t) :: rc
  end function

  function bmx_ctx_set_schedule(ctx, s) bind(C, name = 'bmx_ctx_set_schedule') result(rc)
    import
    type(c_ptr), value :: ctx
    type(bmx_schedule) :: s
    integer(c_int32_t) :: rc
  end function

  function bmx_ctx_get_schedule(ctx, s) bind(C, name = 'bmx_ctx_get_schedule') result(rc)
    import
    type(c_ptr), value :: ctx
    type(bmx_schedule) :: s
    integer(c_int32_t) :: rc
  end function

  function bmx_tune_grid(ctx, lat, grid, ref_orbit, n_turns, best) bind(C, name = 'bmx_tune_grid') result(rc)
    import
    type(c_ptr), value :: ctx, lat
    type(bmx_grid) :: grid
    type(c_ptr), value :: ref_orbit
    integer(c_int32_t), value :: n_turns
    type(bmx_schedule) :: best
    integer(c_int32_t) :: rc
  end function

  function bmx_closed_orbit(ctx, lat, n_dim, orbit, status, stats) bind(C, name = 'bmx_closed_orbit') result(rc)
    import
    type(c_ptr), value :: ctx, lat
    integer(c_int32_t), value :: n_dim
    real(c_double) :: orbit(*)            ! (6, n_variants)
    integer(c_int32_t) :: status(*)
    type(bmx_stats) :: stats
    integer(c_int32_t) :: rc
  end function

  function bmx_one_turn_matrix(ctx, lat, orbit, d_orb, mat6, status, stats) &
                               bind(C, name = 'bmx_one_turn_matrix') result(rc)
    import
    type(c_ptr), value :: ctx, lat
    real(c_double), intent(in) :: orbit(*)
    type(c_ptr), value :: d_orb
    real(c_double), intent(out) :: mat6(*)   ! (6, 6, n_variants), each matrix the transpose of mat6(i, j)
    integer(c_int32_t), intent(out) :: status(*)
    type(bmx_stats) :: stats
    integer(c_int32_t) :: rc
  end function

  ! DA surrogate held by the library: parameters per layer W(out, in) stored
  ! row-major (in Fortran: W(in, out) column-major) then b(out), FP32; X, Y,
  ! knobs, geo and kb are device pointers
  function bmx_surrogate_create(ctx, n_dims, dims, params, n_params, net) &
                                bind(C, name = 'bmx_surrogate_create') result(rc)
    import
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: n_dims
    integer(c_int32_t), intent(in) :: dims(*)
    real(c_float), intent(in) :: params(*)
    integer(c_int64_t), intent(out) :: n_params
    type(c_ptr) :: net
    integer(c_int32_t) :: rc
  end function

  subroutine bmx_surrogate_destroy(net) bind(C, name = 'bmx_surrogate_destroy')
    import
    type(c_ptr), value :: net
  end subroutine

  function bmx_surrogate_set_params(net, params) bind(C, name = 'bmx_surrogate_set_params') result(rc)
    import
    type(c_ptr), value :: net
    real(c_float), intent(in) :: params(*)
    integer(c_int32_t) :: rc
  end function

  function bmx_surrogate_get_params(net, params) bind(C, name = 'bmx_surrogate_get_params') result(rc)
    import
    type(c_ptr), value :: net
    real(c_float), intent(out) :: params(*)
    integer(c_int32_t) :: rc
  end function

  function bmx_surrogate_train(net, x, y, n, epochs, batch, lr, lr_final, seed, flags, epoch_loss) &
                               bind(C, name = 'bmx_surrogate_train') result(rc)
    import
    type(c_ptr), value :: net, x, y
    integer(c_int64_t), value :: n
    integer(c_int32_t), value :: epochs, batch
    real(c_float), value :: lr, lr_final
    integer(c_int64_t), value :: seed
    integer(c_int32_t), value :: flags
    real(c_float), intent(out) :: epoch_loss(*)
    integer(c_int32_t) :: rc
  end function

  function bmx_surrogate_infer(net, x_bf16, n, y) bind(C, name = 'bmx_surrogate_infer') result(rc)
    import
    type(c_ptr), value :: net, x_bf16, y
    integer(c_int64_t), value :: n
    integer(c_int32_t) :: rc
  end function

  function bmx_surrogate_screen(net, n_lat, n_knob, knobs, knob_range, n_ic, n_geo, geo, n_amp, alive_level, kb) &
                                bind(C, name = 'bmx_surrogate_screen') result(rc)
    import
    type(c_ptr), value :: net
    integer(c_int32_t), value :: n_lat, n_knob
    type(c_ptr), value :: knobs
    real(c_float), value :: knob_range
    integer(c_int32_t), value :: n_ic, n_geo
    type(c_ptr), value :: geo
    integer(c_int32_t), value :: n_amp
    real(c_float), value :: alive_level
    type(c_ptr), value :: kb
    integer(c_int32_t) :: rc
  end function

  ! exact transfer matrices (make_mat6 / transfer_matrix_calc over a range):
  ! vec_in, vec_out (6, n) as coord_struct%vec; mat6 (6, 6, n) holds in
  ! Fortran order the transpose of each C row-major matrix, so ele%mat6 =
  ! transpose(mat6(:, :, i))
  function bmx_transfer_matrix(ctx, lat, ix_start, ix_end, n_turns, n, vec_in, vec_out, mat6, state) &
                               bind(C, name = 'bmx_transfer_matrix') result(rc)
    import
    type(c_ptr), value :: ctx, lat
    integer(c_int32_t), value :: ix_start, ix_end, n_turns
    integer(c_int64_t), value :: n
    real(c_double), intent(in) :: vec_in(*)
    real(c_double), intent(out) :: vec_out(*), mat6(*)
    integer(c_int32_t), intent(out) :: state(*)
    integer(c_int32_t) :: rc
  end function
end interface

contains

!+
! Function bmx_error_string(ctx) result(str)
! Copy of the context's last error message as a Fortran string.
!-
function bmx_error_string(ctx) result(str)
  type(c_ptr), value :: ctx
  character(:), allocatable :: str
  character(kind=c_char), pointer :: chars(:)
  type(c_ptr) :: p
  integer n
  interface
    function c_strlen(s) bind(C, name = 'strlen') result(n)
      import
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface
  p = bmx_last_error(ctx)
  n = int(c_strlen(p))
  call c_f_pointer(p, chars, [n])
  allocate(character(n) :: str)
  str = transfer(chars(1:n), str)
end function

end module bmx_c_binding
