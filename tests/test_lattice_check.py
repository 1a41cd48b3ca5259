"""bmx_lattice_check on the CPU: the host-side compile of an element list
for the device path accepts what the device tracks and names what it
rejects, with no device present.  The accepted set of the reference's
tracking_method_test elements is pinned, so a change in device coverage shows
up here."""
import json
import os

import pytest

from bmx import abi, lattice, tracker
from helpers import SPECIES, ele_from_spec

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _check(eles, species=abi.ELECTRON, tracked_species=None, direction=1):
    eles = lattice.set_reference([lattice.EleRec_copy(e) for e in eles], species, e_tot=1e9)
    return tracker.check_lattice(eles, species, tracked_species=tracked_species, direction=direction)


def _n_ops(e):
    ok, n, msg = _check([e])
    assert ok, msg
    return n


def test_plain_elements_compile():
    ok, n, msg = _check([lattice.ele("drift", l=1.0), lattice.ele("quadrupole", l=0.5, k1=0.8, num_steps=1)])
    assert ok and msg == "" and n >= 2


@pytest.mark.parametrize("fringe_at,extra", [(abi.BOTH_ENDS, 2), (abi.ENTRANCE_END, 1), (abi.EXIT_END, 1),
                                             (abi.NO_END, 0)])
@pytest.mark.parametrize("fringe_type", [abi.FRINGE_HARD_EDGE_ONLY, abi.FRINGE_FULL])
def test_quad_and_sextupole_hard_edge_ops(fringe_type, fringe_at, extra):
    """hard_multipole_edge_kick (fringe_mod.f90:714-721) adds one op per edge
    that fringe_at selects, for fringe_type hard_edge_only and full."""
    for key, kw in (("quadrupole", dict(k1=0.8)), ("sextupole", dict(k2=5.0))):
        base = _n_ops(lattice.ele(key, l=0.4, num_steps=2, **kw))
        n = _n_ops(lattice.ele(key, l=0.4, num_steps=2, fringe_type=fringe_type, fringe_at=fringe_at, **kw))
        assert n == base + extra, (key, n, base)


def test_quad_soft_edge_only_is_a_no_op():
    """soft_quadrupole_edge_kick has no strength without fq1/fq2
    (fringe_mod.f90:605-607): no ops."""
    base = _n_ops(lattice.ele("quadrupole", l=0.5, k1=0.8, num_steps=1))
    assert _n_ops(lattice.ele("quadrupole", l=0.5, k1=0.8, num_steps=1,
                              fringe_type=abi.FRINGE_SOFT_EDGE_ONLY)) == base


def test_octupole_fringe_is_ignored():
    """apply_element_edge_kick has no magnetic edge for octupoles."""
    base = _n_ops(lattice.ele("octupole", l=0.3, k3=50.0, num_steps=2))
    assert _n_ops(lattice.ele("octupole", l=0.3, k3=50.0, num_steps=2, fringe_type=abi.FRINGE_FULL)) == base


@pytest.mark.parametrize("kw,extra", [(dict(ref_tilt=0.3), 1), (dict(x_offset=1e-3), 2),
                                      (dict(x_offset=1e-3, ref_tilt=0.3, y_pitch=1e-4), 2)])
def test_misaligned_bend_adds_frame_changes(kw, extra):
    """offset_particle set and unset around the bend: two tilt_coords for
    ref_tilt alone (the element end folds into the second), two frame changes
    with offsets or pitches (the element end stays an op of its own)."""
    base = _n_ops(lattice.ele("sbend", l=1.0, angle=0.02, num_steps=1))
    assert _n_ops(lattice.ele("sbend", l=1.0, angle=0.02, num_steps=1, **kw)) == base + extra


@pytest.mark.parametrize("e,needle", [
    (lattice.ele("sbend", l=1.0, angle=0.02, num_steps=1, b_elec={0: 1.0, 6: 1.0}), "6 in a bend"),
    (lattice.ele("quadrupole", l=0.5, k1=0.8, num_steps=1, b_elec={0: 1.0, 7: 1.0}), "more than 7 orders"),
    (lattice.ele("drift", l=1.0, orientation=2), "orientation must be +1 or -1"),
    (lattice.ele("sbend", l=1.0, angle=0.02, num_steps=1, fringe_type=abi.FRINGE_FULL), "fringe_type full"),
    (lattice.ele("sbend", l=1.0, angle=0.02, num_steps=1, fringe_type=abi.FRINGE_SAD_FULL), "fringe_type full"),
])
def test_rejections_name_the_feature(e, needle):
    ok, n, msg = _check([lattice.ele("drift", l=0.5), e])
    assert not ok and n == 0
    assert msg.startswith("element 1 ") and needle in msg, msg


def test_invalid_arguments():
    lib = abi.load_library()
    import ctypes as C
    msg = C.create_string_buffer(64)
    ref = abi.Ref(abi.ELECTRON, 0)
    assert lib.bmx_lattice_check(None, 0, C.byref(ref), None, None, msg, len(msg)) == -1
    assert msg.value == b"invalid arguments"


@pytest.mark.parametrize("fringe_at,n_fringe", [(abi.BOTH_ENDS, 2), (abi.ENTRANCE_END, 1), (abi.NO_END, 0)])
def test_electric_multipole_ops(fringe_at, n_fringe):
    """A quadrupole with electric multipoles: an electric kick after every
    magnetic kick (n_step + 1 of them, ab_multipole_kicks) and an electric
    longitudinal fringe at each edge fringe_at selects, whatever fringe_type
    (has_elec_poles makes the edge kicks run)."""
    kw = dict(l=0.6, k1=0.5, num_steps=4, fringe_at=fringe_at)
    base = _n_ops(lattice.ele("quadrupole", b={2: 1.0}, **kw))
    n = _n_ops(lattice.ele("quadrupole", b={2: 1.0}, a_elec={1: 2e4}, b_elec={0: -2e3}, **kw))
    assert n == base + 5 + n_fringe


def test_electric_poles_of_thin_multipoles_are_ignored():
    """track1_bmad tracks multipole / ab_multipole elements with the kt kicks
    of their magnetic poles alone."""
    for key in ("multipole", "ab_multipole"):
        kw = dict(knl={2: 1.0}) if key == "multipole" else dict(b={2: 1.0})
        assert _n_ops(lattice.ele(key, a_elec={1: 1e4}, **kw)) == _n_ops(lattice.ele(key, **kw))


def test_golden_elements_the_device_accepts():
    """The device path takes every one of the reference's tracking_method_test
    elements."""
    tm = json.load(open(os.path.join(GOLD, "tracking_method_bmad_standard.json")))
    sp = SPECIES[tm["particle"]]
    accepted, rejected = set(), {}
    for name, spec in tm["elements"].items():
        ok, _, msg = tracker.check_lattice([ele_from_spec(spec, sp, e_tot=tm["e_tot"])], sp)
        if ok:
            accepted.add(name)
        else:
            rejected[name] = msg
    assert not rejected, rejected
    assert len(accepted) == len(tm["elements"]) == 24


def test_phase_space_limits_only_on_collimators():
    """check_aperture_limit reads the phase-space limits of ecollimator and
    rcollimator alone: on a drift they change nothing; a collimator with
    them gets its own element-end op (and an entrance check where
    aperture_at asks for one), even with no x, y limits."""
    ps = dict(px_aperture_width2=1e-3, pz_aperture_width2=2e-3)
    assert _n_ops(lattice.ele("drift", l=1.0, **ps)) == _n_ops(lattice.ele("drift", l=1.0))
    for key in ("rcollimator", "ecollimator"):
        assert _n_ops(lattice.ele(key, l=0.5, **ps)) == _n_ops(lattice.ele(key, l=0.5)) + 1
        both = _n_ops(lattice.ele(key, l=0.5, aperture_at=abi.BOTH_ENDS, **ps))
        assert both == _n_ops(lattice.ele(key, l=0.5)) + 2


def test_tracked_species_must_be_the_antiparticle():
    """A lattice compiles for the reference species or its antiparticle; a
    species of another mass is refused before any element."""
    eles = [lattice.ele("quadrupole", l=0.5, k1=0.8, num_steps=1)]
    for sp, tr in ((abi.ELECTRON, abi.POSITRON), (abi.POSITRON, abi.ELECTRON), (abi.PROTON, abi.ANTIPROTON),
                   (abi.ELECTRON, abi.ELECTRON)):
        ok, n, msg = _check(eles, sp, tr)
        assert ok and n == _check(eles, sp)[1], msg
    ok, n, msg = _check(eles, abi.ELECTRON, abi.PROTON)
    assert not ok and n == 0 and "antiparticle" in msg, msg


@pytest.mark.parametrize("key,kw", [("quadrupole", dict(k1=0.8)), ("sextupole", dict(k2=5.0)),
                                    ("rfcavity", dict(voltage=1e6, rf_frequency=5e8)),
                                    ("multipole", dict(knl={2: 1.0}, l=0.0))])
def test_reversed_elements_compile_to_as_many_ops(key, kw):
    """Orientation -1 changes constants and which physical end an edge is,
    not the op sequence: a misaligned reversed element's frame changes are
    one affine op each, like the forward element's offsets."""
    kw = dict(dict(l=0.4, num_steps=2), **kw)
    mis = dict(x_offset=1e-3, y_pitch=2e-4, z_offset=1e-3, tilt=0.1)
    for extra in ({}, mis, dict(mis, fringe_type=abi.FRINGE_FULL, fringe_at=abi.BOTH_ENDS)):
        if key in ("rfcavity", "multipole") and "fringe_type" in extra:
            continue
        assert _n_ops(lattice.ele(key, orientation=-1, **kw, **extra)) == _n_ops(lattice.ele(key, **kw, **extra))


def test_reversed_element_edges_follow_the_physical_end():
    """fringe_at = entrance_end on a reversed quadrupole: the hard edge is
    still one op (met at the second track edge); an aperture at the entrance
    end of a reversed drift is checked at its second track edge, where the
    element end op carries it."""
    q = dict(l=0.4, k1=0.8, num_steps=2, fringe_type=abi.FRINGE_FULL, fringe_at=abi.ENTRANCE_END)
    assert _n_ops(lattice.ele("quadrupole", orientation=-1, **q)) == _n_ops(lattice.ele("quadrupole", **q))
    ap = dict(l=1.0, x1_limit=0.01, x2_limit=0.01, y1_limit=0.01, y2_limit=0.01, aperture_at=abi.ENTRANCE_END)
    fwd, rev = _n_ops(lattice.ele("drift", **ap)), _n_ops(lattice.ele("drift", orientation=-1, **ap))
    assert fwd == rev + 1  # forward: an entrance aperture op; reversed: folded into the drift's end


def test_golden_elements_accepted_reversed_for_antiparticles():
    """The -Anti_O variant of tracking_method_test: every element except sbend
    and rfcavity reversed and tracked as the antiparticle."""
    tm = json.load(open(os.path.join(GOLD, "tracking_method_bmad_standard.json")))
    sp = SPECIES[tm["particle"]]
    anti = abi.POSITRON + abi.ELECTRON - sp
    for name, spec in tm["elements"].items():
        e = ele_from_spec(spec, sp, e_tot=tm["e_tot"])
        flip = spec["key"] not in ("sbend", "rfcavity")
        if flip:
            e.orientation = -1
        ok, _, msg = tracker.check_lattice([e], sp, tracked_species=anti if flip else None)
        assert ok, (name, msg)


def test_golden_elements_accepted_backward():
    """The -Anti_D and -Anti_OD variants (direction = -1): every element,
    rf cavities included (their phase on relative or absolute time), compiles
    for backward tracking."""
    tm = json.load(open(os.path.join(GOLD, "tracking_method_bmad_standard.json")))
    sp = SPECIES[tm["particle"]]
    for orient in (1, -1):
        for name, spec in tm["elements"].items():
            e = ele_from_spec(spec, sp, e_tot=tm["e_tot"])
            if spec["key"] not in ("sbend", "rfcavity"):
                e.orientation = orient
            anti = abi.POSITRON + abi.ELECTRON - sp if -e.orientation == -1 else None
            for abs_time in (False, True):
                ok, _, msg = tracker.check_lattice([e], sp, tracked_species=anti, direction=-1,
                                                   absolute_time_tracking=abs_time)
                assert ok, (name, msg)


def test_backward_programs():
    """Backward quadrupoles carry track_a_quadrupole's -2 c beta delta_ref_time
    after every body step; thick multipoles run their drifts and kicks as
    separate ops; a direction other than +1 or -1 is refused."""
    q = lattice.ele("quadrupole", l=0.6, k1=0.5, num_steps=3, b={2: 1.0})
    assert _n_ops_dir(q, -1) == _n_ops_dir(q, 1) + 3
    sx = lattice.ele("sextupole", l=0.4, k2=5.0, num_steps=4)
    assert _n_ops_dir(sx, -1) == 4 + 5  # 4 drifts and 5 kicks (the element end folds into the last kick)
    assert _n_ops_dir(sx, 1) == 1
    ok, _, msg = _check([lattice.ele("drift", l=1.0)], direction=2)
    assert not ok and "direction must be" in msg


def _n_ops_dir(e, direction):
    ok, n, msg = _check([e], direction=direction)
    assert ok, msg
    return n
