"""Extract golden vectors for the bmad_standard path from the reference's
regression tests into small JSON fixtures under tests/golden/.

Run in a container that has the reference tree (default /root/reference);
the JSON files it writes are committed and are all the tests read.

Sources (relative to the bmad-ecosystem root):
  regression_tests/tracking_method_test/output.correct  (Bmad_Standard rows)
  regression_tests/tracking_method_test/tracking_method_test.bmad (element attributes, start orbit)
  regression_tests/fringe_test/output.correct, fringe_test.bmad
  regression_tests/aperture_test/output.correct, aperture_test.bmad

The element attributes below are transcribed from tracking_method_test.bmad
(lines cited) into this build's element-record vocabulary; defaults are the
ones bmad/parsing/set_ele_defaults.f90 applies.
"""
import json
import math
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

PI = math.pi

# tracking_method_test.bmad: element definitions used by the Bmad_Standard rows
# this build covers.  q0 / sb0 base attributes (lines 62, 79) are merged in.
Q0 = dict(a={0: 0.01, 1: 0.1}, vkick=-0.0001, b_elec={0: -2e3}, a_elec={1: 2e4})
SB0 = dict(a={0: 0.01, 1: 0.1}, vkick=-0.0001, b_elec={0: -2e3}, a_elec={1: 2e4}, r0_mag=2, r0_elec=3)
ELEMENTS = {
    # line 18
    "AB_MULTIPOLE1": dict(key="ab_multipole", a={2: 1e2}, b={2: 1e2}, x_offset=0.02, tilt=0.1),
    # line 25
    "DRIFT1": dict(key="drift", l=4.5, num_steps=10, x_offset=0.05, y_pitch=0.001, tilt=0.1),
    # line 27
    "ECOLLIMATOR1": dict(key="ecollimator", l=4.5, x_limit=0.2, y_limit=0.1, num_steps=10, x_offset=0.02,
                         y_pitch=0.001, tilt=0.1, hkick=0.01, vkick=0.002),
    # line 40
    "HKICKER1": dict(key="hkicker", l=4.5, kick=0.005, num_steps=10, x_offset=0.05, y_pitch=0.001, tilt=0.1),
    # line 41
    "INSTRUMENT1": dict(key="instrument", l=4.5, num_steps=10),
    # line 42
    "KICKER1": dict(key="kicker", l=4.5, b={0: 0.002}, vkick=0.003, num_steps=10, x_offset=0.05, y_pitch=0.001,
                    tilt=0.1),
    # line 56
    "MONITOR1": dict(key="monitor", l=4.5, num_steps=10),
    # line 57: t1 with no value = pi/4
    "MULTIPOLE1": dict(key="multipole", knl={1: 1.0, 3: 4.5}, tn={1: PI / 4, 3: 0.31 * PI}, x_offset=0.05),
    # line 59: bare tilt = pi/8
    "OCTUPOLE1": dict(key="octupole", l=4.5, k3=1e3, tilt=PI / 8, num_steps=10, x_offset=0.01, y_pitch=0.001),
    # lines 62-67: bare tilt = pi/4
    "QUADRUPOLE1": dict(key="quadrupole", l=0.6, k1=1, tilt=PI / 4, num_steps=10, x_offset=0.05, y_pitch=0.001, **Q0),
    "QUADRUPOLE2": dict(key="quadrupole", l=0.6, k1=-1, tilt=PI / 4, num_steps=10, x_offset=0.05, y_pitch=0.001, **Q0),
    "QUADRUPOLE3": dict(key="quadrupole", l=0.6, k1=1e-20, tilt=PI / 4, num_steps=10, x_offset=0.05, y_pitch=0.001,
                        **Q0),
    "QUADRUPOLE4": dict(key="quadrupole", l=0.6, k1=1, tilt=PI / 4, num_steps=10, x_offset=0.05, y_pitch=0.001,
                        fringe_type="full", **Q0),
    "QUADRUPOLE5": dict(key="quadrupole", l=0.6, k1=-1, tilt=PI / 4, num_steps=10, x_offset=0.05, y_pitch=0.001,
                        fringe_type="full", **Q0),
    # line 68
    "RCOLLIMATOR1": dict(key="rcollimator", l=4.5, x_limit=0.2, y_limit=0.1, num_steps=10, x_offset=0.02,
                         y_pitch=0.001, tilt=0.1),
    # lines 69-70
    "RFCAVITY1": dict(key="rfcavity", l=4.5, rf_frequency=500e6, voltage=1e5, phi0=0.1, num_steps=10, x_offset=0.05,
                      y_pitch=0.001, tilt=0.1),
    # lines 71-72: cavity_type = traveling_wave changes nothing in bmad_standard rfcavity tracking
    # (track_a_rfcavity.f90 and e_accel_field.f90 never read it)
    "RFCAVITY2": dict(key="rfcavity", l=4.5, rf_frequency=500e6, voltage=1e5, phi0=0.1, num_steps=10, x_offset=0.05,
                      y_pitch=0.001, tilt=0.1, coupler_at="both_ends", coupler_strength=0.1),
    # lines 73-74 (zero length)
    "RFCAVITY3": dict(key="rfcavity", rf_frequency=500e6, voltage=1e5, phi0=0.1, num_steps=10, x_offset=0.05,
                      y_pitch=0.001, tilt=0.1, a={2: 0.02}),
    # line 84
    "SBEND3": dict(key="sbend", l=0.6, k1=0.3, k2=3, num_steps=10, x_offset=0.05, y_pitch=0.001, ref_tilt=0.1, **SB0),
    # lines 88-89: an rbend, settled into an sbend by the parser rule below (rbend_as_sbend)
    "RBEND6": None,
    # line 91
    "SBEND7": dict(key="sbend", l=0.6, g=0.01, e1=0.1, e2=0.2, num_steps=10, x_offset=0.05, y_pitch=0.001,
                   ref_tilt=0.1, **SB0),
    # line 93: bare tilt = pi/6
    "SEXTUPOLE1": dict(key="sextupole", l=0.6, k2=1, tilt=PI / 6, num_steps=10, x_offset=0.05, y_pitch=0.001),
    # line 129: no num_steps -> ds_step = default_ds_step 0.2 -> nint(22.5) = 23 steps
    "THICK_MULTIPOLE1": dict(key="thick_multipole", l=4.5, a={2: 1}, b={3: 20}, x_offset=0.02, tilt=0.1,
                             num_steps=23),
    # line 130
    "VKICKER1": dict(key="vkicker", l=4.5, kick=0.005, num_steps=10, x_offset=0.05, y_pitch=0.001, tilt=0.1),
}

def rbend_as_sbend(l_chord, g, e1, e2, **kw):
    """An rbend given by chord length and g, as bmad/parsing/bmad_parser_mod.f90:5155-5191
    settles it (fiducial_pt none): arc angle a = 2 asin(g l_chord / 2), arc length
    l = l_chord a / (2 sin(a / 2)), angle = g l, e1 and e2 each + angle / 2, key sbend."""
    a = 2.0 * math.asin(0.5 * l_chord * g)
    l = l_chord * a / (2.0 * math.sin(0.5 * a))
    angle = g * l
    return dict(key="sbend", l=l, g=angle / l, e1=e1 + 0.5 * angle, e2=e2 + 0.5 * angle, **kw)


# rb0 (line 80) has sb0's attributes; line 88-89: l = 0.6 is the chord length of an rbend
ELEMENTS["RBEND6"] = rbend_as_sbend(0.6, 0.01, 0.1, 0.2, dg=-0.0099999, num_steps=10, x_offset=0.05,
                                    y_pitch=0.001, ref_tilt=0.1, **SB0)

VARIANTS = {"": (1, 1, False), "-Anti_O": (1, -1, False), "-Anti_D": (-1, 1, False), "-Anti_OD": (-1, -1, False),
            "-Abs_Time": (1, 1, True)}


def tracking_method():
    path = os.path.join(REF, "regression_tests/tracking_method_test/output.correct")
    rows = []
    pat = re.compile(r'^"([A-Z0-9_]+)((?:-Anti_OD|-Anti_O|-Anti_D|-Abs_Time)?): Bmad_Standard"\s+ABS\s+(\S+)\s+(.*)$')
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            m = pat.match(line.strip())
            if not m or m.group(1) not in ELEMENTS:
                continue
            vals = [float(v) for v in m.group(4).split()]
            rows.append(dict(name=m.group(1), variant=m.group(2), line=lineno, tol=float(m.group(3)), vec=vals[:6],
                             col7=vals[6]))
    return dict(
        source="regression_tests/tracking_method_test/output.correct",
        lattice="regression_tests/tracking_method_test/tracking_method_test.bmad",
        # tracking_method_test.bmad:9-14 (start orbit), :141 (l1[e_tot] = 1e6); default species positron
        particle="positron", e_tot=1e6, start=[0.001, 0.002, 0.003, 0.004, 0.005, 0.006],
        printed_digits=11, elements=ELEMENTS, variants={k: list(v) for k, v in VARIANTS.items()}, rows=rows)


def fringe():
    path = os.path.join(REF, "regression_tests/fringe_test/output.correct")
    out = {}
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            m = re.match(r'^"([A-Za-z0-9-]+)"\s+ABS\s+(\S+)\s+(.*)$', line.strip())
            if m:
                out[m.group(1)] = dict(line=lineno, tol=float(m.group(2)), vals=[float(v) for v in m.group(3).split()])
    # fringe_test.bmad: electron, e_tot 1e6, start (0.01, 0.02, 0.03, 0.04, 0, 0.5)
    # b: sbend, e1 = 0.1, e2 = -0.2, k1 = 0.1, g = 0.1, fint = 0.5, hgap = 0.4 (fintx, hgapx default to fint, hgap)
    return dict(source="regression_tests/fringe_test/output.correct", particle="electron", e_tot=1e6,
                start=[0.01, 0.02, 0.03, 0.04, 0.0, 0.5],
                element=dict(key="sbend", e1=0.1, e2=-0.2, k1=0.1, g=0.1, fint=0.5, hgap=0.4, fintx=0.5, hgapx=0.4),
                rows=out)


def aperture():
    path = os.path.join(REF, "regression_tests/aperture_test/output.correct")
    out = {}
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            m = re.match(r'^"((?:EC|RC)_[SU]_[A-Za-z]+[-+])"\s+ABS\s+(\S+)\s+(.*)$', line.strip())
            if m:
                out[m.group(1)] = dict(line=lineno, tol=float(m.group(2)), vals=[float(v) for v in m.group(3).split()])
    # aperture_test.bmad: antimuon at p0c = 3.094353005e9; ec / rc attributes (lines 22-35)
    coll = dict(px_aperture_width2=0.01, px_aperture_center=0.01, py_aperture_width2=0.015, z_aperture_width2=0.2,
                z_aperture_center=-0.6, pz_aperture_width2=0.03, pz_aperture_center=0.02, x1_limit=1e-10,
                x2_limit=0.02, y_limit=0.015)
    return dict(source="regression_tests/aperture_test/output.correct", particle="antimuon", p0c=3.094353005e9,
                elements={"EC": dict(key="ecollimator", **coll), "RC": dict(key="rcollimator", **coll)}, rows=out)


def multipole():
    """MULT rows of regression_tests/multipole_test for the lord elements
    QQ, BB, MM: a1, a3, a5, b2, b4, b6 of multipole_ele_to_ab (magnetic), then
    the electric call's ix_pole_max, for use_ele_tilt (first digit 2) and
    include_kicks (second digit 2) (multipole_test.f90:34-47)."""
    path = os.path.join(REF, "regression_tests/multipole_test/output.correct")
    out = {}
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            m = re.match(r'^"MULT-([12])([12])-(QQ|BB|MM)"\s+ABS\s+(\S+)\s+(.*)$', line.strip())
            if m:
                vals = m.group(5).split()
                out[f"{m.group(3)}-{m.group(1)}{m.group(2)}"] = dict(
                    line=lineno, use_ele_tilt=m.group(1) == "2", include_kicks=m.group(2) == "2",
                    vals=[float(v) for v in vals[:6]], ix_elec=int(vals[6]))
    # multipole.bmad:9-26 (e_tot 1e9, electron); bb's "a5 = 5, a5 = 6" leaves a4 = 0 and a5 = 6
    elements = {
        "QQ": dict(key="quadrupole", l=1e-3, k1=0.1, tilt=PI / 4, scale_multipoles=0,
                   a={0: 1, 1: 2, 2: 3, 3: 4, 4: 5, 5: 6}, b={0: 0.1, 1: 0.2, 2: 0.3, 3: 0.4, 4: 0.5, 5: 0.6},
                   a_elec={0: 1e-2, 1: 2e-2, 2: 3e-2, 3: 4e-2, 4: 5e-2, 5: 6e-2},
                   b_elec={0: 1e-3, 1: 2e-3, 2: 3e-3, 3: 4e-3, 4: 5e-3, 5: 6e-3}, num_steps=1),
        "BB": dict(key="sbend", l=1e-3, g=1e-2, k1=0.1, k2=0.2, ref_tilt=PI / 2,
                   a={0: 1, 1: 2, 2: 3, 3: 4, 5: 6}, b={0: 0.1, 1: 0.2, 2: 0.3, 3: 0.4, 4: 0.5, 5: 0.6},
                   a_elec={0: 1e-2, 1: 2e-2, 2: 3e-2, 3: 4e-2, 4: 5e-2, 5: 6e-2},
                   b_elec={0: 1e-3, 1: 2e-3, 2: 3e-3, 3: 4e-3, 4: 5e-3, 5: 6e-3}, num_steps=1),
        "MM": dict(key="multipole", knl={6: 1, 1: 2, 2: 3, 3: 4, 4: 5, 5: 6},
                   tn={0: 0.1, 1: 0.2, 2: 0.3, 3: 0.4, 4: 0.5, 5: 0.6}),
    }
    return dict(source="regression_tests/multipole_test/output.correct", particle="electron", e_tot=1e9,
                printed_digits=5, elements=elements, rows=out)


# mat6_calc_method_test.bmad: the elements of this path, as that file defines
# them (they differ from tracking_method_test.bmad's in several values).
# q0 (line 61) and sb0 / rb0 (lines 79-80) are merged in; sb0 sets no r0.
MQ0 = dict(a={0: 0.01, 1: 0.1}, vkick=-0.0001, b_elec={0: -2e3}, a_elec={1: 2e5})
MSB0 = dict(a={0: 0.001, 1: 0.1}, vkick=-0.0001, b_elec={0: -2e3}, a_elec={1: 2e3})
MAT6_ELEMENTS = {
    "AB_MULTIPOLE1": dict(key="ab_multipole", a={2: 1e3}, b={2: 1e3}, x_offset=0.02, tilt=0.1),
    "DRIFT1": dict(key="drift", l=4.5, num_steps=10, x_offset=0.05, y_pitch=0.001, tilt=0.1),
    "ECOLLIMATOR1": dict(key="ecollimator", l=4.5, x_limit=0.2, y_limit=0.1, num_steps=10, x_offset=0.02,
                         y_pitch=0.001, tilt=0.1, hkick=0.01, vkick=0.002),
    "HKICKER1": dict(key="hkicker", l=4.5, kick=0.005, num_steps=10, x_offset=0.05, y_pitch=0.001, tilt=0.1),
    "INSTRUMENT1": dict(key="instrument", l=4.5, num_steps=10),
    "KICKER1": dict(key="kicker", l=4.5, hkick=0.003, vkick=0.003, num_steps=10, x_offset=0.05, y_pitch=0.001,
                    tilt=0.1),
    "MONITOR1": dict(key="monitor", l=4.5, num_steps=10),
    # t1 with no value = pi/4
    "MULTIPOLE1": dict(key="multipole", knl={1: 1.0, 3: 4.5}, tn={1: PI / 4, 3: 0.31 * PI}, x_offset=0.05),
    # bare tilt = pi/8
    "OCTUPOLE1": dict(key="octupole", l=4.5, k3=1e3, tilt=PI / 8, num_steps=10, x_offset=0.01, y_pitch=0.001),
    # quadrupole1 has tilt = 0.1, the others a bare tilt = pi/4
    "QUADRUPOLE1": dict(key="quadrupole", l=0.6, k1=1, tilt=0.1, num_steps=10, x_offset=0.05, y_pitch=0.001, **MQ0),
    "QUADRUPOLE2": dict(key="quadrupole", l=0.6, k1=-1, tilt=PI / 4, num_steps=10, x_offset=0.05, y_pitch=0.001,
                        **MQ0),
    "QUADRUPOLE3": dict(key="quadrupole", l=0.6, k1=1e-20, tilt=PI / 4, num_steps=10, x_offset=0.05, y_pitch=0.001,
                        **MQ0),
    "QUADRUPOLE4": dict(key="quadrupole", l=0.6, k1=1, tilt=PI / 4, num_steps=10, x_offset=0.05, y_pitch=0.001,
                        fringe_type="full", **MQ0),
    "QUADRUPOLE5": dict(key="quadrupole", l=0.6, k1=-1, tilt=PI / 4, num_steps=10, x_offset=0.05, y_pitch=0.001,
                        fringe_type="full", **MQ0),
    "RCOLLIMATOR1": dict(key="rcollimator", l=4.5, x_limit=0.2, y_limit=0.1, num_steps=10, x_offset=0.02,
                         y_pitch=0.001, tilt=0.1),
    "RFCAVITY1": dict(key="rfcavity", l=4.5, rf_frequency=500e6, voltage=5e5, phi0=0.1, num_steps=10, x_offset=0.05,
                      y_pitch=0.001, tilt=0.1, b={2: 0.002}),
    "RFCAVITY2": dict(key="rfcavity", l=4.5, rf_frequency=500e6, voltage=5e5, phi0=0.1, num_steps=10, x_offset=0.05,
                      y_pitch=0.001, tilt=0.1, coupler_at="both_ends", coupler_strength=0.1),
    "SBEND3": dict(key="sbend", l=0.6, k1=0.3, k2=3, num_steps=10, x_offset=0.05, y_pitch=0.001, ref_tilt=0.1,
                   **MSB0),
    "RBEND6": None,
    "SBEND7": dict(key="sbend", l=0.6, g=0.01, e1=0.1, e2=0.2, num_steps=10, x_offset=0.05, y_pitch=0.001,
                   ref_tilt=0.1, **dict(MSB0, a={0: 0.001, 1: 10, 3: 100})),
    # bare tilt = pi/6
    "SEXTUPOLE1": dict(key="sextupole", l=0.6, k2=1, tilt=PI / 6, num_steps=10, x_offset=0.05, y_pitch=0.001),
    # no num_steps -> ds_step = default_ds_step 0.2 -> nint(22.5) = 23 steps
    "THICK_MULTIPOLE1": dict(key="thick_multipole", l=4.5, a={2: 1}, b={2: 20}, x_offset=0.02, tilt=0.1,
                             num_steps=23),
    "VKICKER1": dict(key="vkicker", l=4.5, kick=0.005, num_steps=10, x_offset=0.05, y_pitch=0.001, tilt=0.1),
}
MAT6_ELEMENTS["RBEND6"] = rbend_as_sbend(0.6, 0.01, 0.1, 0.2, dg=-0.009999999, num_steps=10, x_offset=0.05,
                                         y_pitch=0.001, ref_tilt=0.1, **MSB0)


def mat6():
    """regression_tests/mat6_calc_method_test/output.correct: the
    Bmad_Standard transfer matrix rows (make_mat6 with the make_matrix
    branches of the tracking routines), vec0 and symplectic error of each
    element above, at particle_start (mat6_calc_method_test.bmad:9-14), e_tot
    1 MeV (:2), default species positron."""
    path = os.path.join(REF, "regression_tests/mat6_calc_method_test/output.correct")
    pat = re.compile(r'^"([A-Z0-9_]+):Bmad_Standard:(MatrixRow[1-6]|Vector|Symp_Err)"\s+ABS\s+(\S+)\s+(.*)$')
    rows = {}
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            m = pat.match(line.strip())
            if not m or m.group(1) not in MAT6_ELEMENTS:
                continue
            rows.setdefault(m.group(1), {})[m.group(2)] = dict(line=lineno, tol=float(m.group(3)),
                                                              vals=[float(v) for v in m.group(4).split()])
    return dict(source="regression_tests/mat6_calc_method_test/output.correct",
                lattice="regression_tests/mat6_calc_method_test/mat6_calc_method_test.bmad",
                particle="positron", e_tot=1e6, start=[0.001, 0.002, 0.003, 0.004, 0.005, 0.006],
                printed_digits=14, elements=MAT6_ELEMENTS, rows=rows)


# nonlin_test/small_ring.bmad (lines 28-40): the ring whose file records the
# coupled periodic Twiss Bmad computed for it (beginning[...], lines 7-22).
# rf is off for the Twiss (voltage 0 here), as the recorded values need.
SMALL_RING = [
    dict(key="sextupole", l=0.01, num_steps=1),
    dict(key="drift", l=0.3),
    dict(key="quadrupole", l=0.5, tilt=0.05, k1=0.6, num_steps=1),
    dict(key="drift", l=0.3),
    dict(key="sbend", l=3.0, g=1.30899693899574704e-1, e1=0.07, e2=0.07, num_steps=1),
    dict(key="drift", l=0.3),
    dict(key="quadrupole", l=0.5, k1=-0.5, num_steps=1),
    dict(key="drift", l=0.3),
    dict(key="sextupole", l=0.01, num_steps=1),
    dict(key="rfcavity", l=1.2, voltage=0.0, rf_frequency=2.5e8, num_steps=10),
    dict(key="sbend", l=3.0, g=1.30899693899574704e-1, e1=0.07, e2=0.07, num_steps=1),
]


def nonlin():
    """The beginning[...] Twiss of nonlin_test/small_ring.bmad (positron, p0c 1 GeV)."""
    path = os.path.join(REF, "regression_tests/nonlin_test/small_ring.bmad")
    twiss = {}
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            m = re.match(r"^beginning\[(\w+)\]\s*=\s*(\S+)", line.strip())
            if m:
                twiss[m.group(1)] = dict(value=float(m.group(2)), line=lineno)
    return dict(source="regression_tests/nonlin_test/small_ring.bmad", particle="positron", p0c=1e9,
                elements=SMALL_RING, twiss=twiss, rows=twiss)


# closed_orbit_test/ad05f_pos_RF.lat and ad05f_neg_RF.lat: an SLS-2 ring
# (2.4 GeV e-, 12 cavities at harmonic 96) whose 6-d closed orbit with
# radiation damping on (the test leaves bmad_com%radiation_damping_on set,
# closed_orbit_test.f90:28-31) and synchrotron tune the test records.
# Element attributes (lines 6-151) and lines (lines 155-213) as data; the
# "_inv" bends swap e1 and e2; sbend[ds_step] = 0.005, sextupole and
# quad::* num_steps = 16 (lines 210-212); a cavity's ds_step is
# bmad_com%default_ds_step = 0.2 m, so 0.88 m is tracked in nint(4.4) = 4
# slices.  The pos_RF file halves every sextupole's k2 (line 222) and the
# neg_RF file negates the voltage (line 6).
AD05F_DRIFTS = dict(d0=0.1, d1=0.2, d2=0.4, dq=0.3, dc=0.38, dc1=0.21, dc2=0.21, ds1=0.15, ds2=0.15, dm1=0.15,
                    dm2=0.1, dl1=0.35, dl2=0.35, dsmgp=1.0, dm=3.06, dl=5.34, dqs=0.05, de=0.1, doc=0.05,
                    no_sas=0.1, no_sam=0.1, no_sal=0.1, gj=0.0, gb=0.0, gm=0.0)
AD05F_QUADS = dict(qf=(0.2, 7.420488), qf1=(0.15, 0), qcs=(0.15, 7.600167), qds=(0.1, 1.093151),
                   qcs1=(0.15, 6.549262), qs1=(0.15, 0), qs2=(0.15, -7.385729), qs3=(0.2, 7.467984),
                   qm1=(0.1, 1.537587), qm2=(0.15, -8.479469), qm3=(0.2, 7.204039), ql1=(0.1, -1.283269),
                   ql2=(0.1, -5.033152), ql3=(0.2, 6.820985), ql4=(0.1, -5.664977), qd=(0.15, -7.129892),
                   qes=(0.1, -2.892556))
AD05F_SEXTUPOLES = dict(sf=650.94069, sfc=638.68885, sfc1=482.44199, sd1=-328.93798, sd2=-420.96168,
                        sd3=-726.04434, sd4=-603.7521, sde=-305.55462, sd=-848.21693)
# bs00..bs14: (l, e1, e2), angle 0.004420919 each
AD05F_BS = [(0.0062, 0, 0.004420919), (0.00748, -0.004420919, 0.0088435833), (0.00897, -0.0088435833, 0.013264502),
            (0.01072, -0.013264502, 0.017685421), (0.01275, -0.017685421, 0.022108086),
            (0.0151, -0.022108086, 0.026529005), (0.01782, -0.026529005, 0.030949924),
            (0.02094, -0.030949924, 0.035372588), (0.02452, -0.035372588, 0.039793507),
            (0.02861, -0.039793507, 0.044214426), (0.03328, -0.044214426, 0.04863709),
            (0.03859, -0.04863709, 0.053058009), (0.04461, -0.053058009, 0.057478928),
            (0.05142, -0.057478928, 0.061901593), (0.0591, -0.061901593, 0.066322512)]
AD05F_MULTIPOLES = dict(oyy0=3364.6247, oxy0=-2626.9988, ocyy0=1.0, ocxy0=-2002.6493, ocxx0=1082.4127)
AD05F_MARKERS = ["no_oxxm0", "no_oxxs0", "no_oxyl0", "no_ocyy0", "osbend", "onbend", "obm", "obm2", "ol3"]
AD05F_LINES = {
    "mhv": "ch mon cv", "mhv_inv": "cv mon ch", "bn": "bn00 bn01", "bn_inv": "bn01_inv bn00_inv",
    "ds": "dsend dsmgp", "oxy": "doc oxy0 doc", "oyy": "doc oyy0 doc", "ocxy": "doc ocxy0 doc",
    "ocyy": "doc no_ocyy0 doc", "ocxx": "doc ocxx0 doc",
    "bs": " ".join(f"bs{i:02d}" for i in range(15)),
    "bs_inv": " ".join(f"bs{i:02d}_inv" for i in range(14, -1, -1)),
    "hc": "bs gb d0 gj qd dqs mhv sd dqs qf dqs sf ocxx d1 bahq1 d2 ocxy dq sd1 mhv dqs ocyy gj gb bn_inv obm bn "
          "gb gj dqs mhv sd2 dc",
    "hc_inv": "dc sd2 mhv_inv dqs gj gb bn_inv obm bn gb gj ocyy dqs mhv_inv sd1 dq ocxy d2 bahq1_inv d1 ocxx sf "
              "dqs qf dqs sd mhv_inv dqs qd gj d0 gb bs_inv",
    "ts": "qcs dqs sfc dqs bahq2 dc sd3 mhv dqs gj gb bn_inv obm2 bn gb gj dqs mhv sd4 dqs qds dc1 qcs1 dqs sfc1 "
          "dqs bahq3 dc2 qes dqs sde mhv dqs gj gb bn_inv gb gj oyy mhv",
    "ts_inv": "mhv_inv oyy gj gb bn gb gj dqs mhv_inv sde dqs qes dc2 bahq3_inv dqs sfc1 dqs qcs1 dc1 qds dqs sd4 "
              "mhv_inv dqs gj gb bn_inv obm2 bn gb gj dqs mhv_inv sd3 dc bahq2_inv dqs sfc dqs qcs",
    "ms": "no_sas dqs qs2 ds2 oxy qs3 mhv gm ds", "ms_inv": "ds gm mhv_inv qs3 oxy ds2 qs2 dqs no_sas",
    "mm": "qm1 dm1 no_sam dqs qm2 dm2 oxy qm3 mhv gm dm", "mm_inv": "dm gm mhv_inv qm3 oxy dm2 qm2 dqs no_sam dm1 qm1",
    "ml": "ql1 dl1 ql2 dl2 no_sal oxy ol3 ql3 dl1 ql4 mhv gm dl",
    "ml_inv": "dl gm mhv_inv ql4 dl1 ql3 ol3 oxy no_sal dl2 ql2 dl1 ql1",
    "hs": "osbend hc ts ms", "hs_inv": "ms_inv ts_inv hc_inv osbend",
    "hm": "osbend hc ts mm", "hm_inv": "mm_inv ts_inv hc_inv osbend",
    "hl": "osbend hc ts ml", "hl_inv": "ml_inv ts_inv hc_inv osbend",
    "six": "hl_inv hs hs_inv hm", "six_inv": "hm_inv hs hs_inv hl", "per": "six six_inv", "ring": "per per per",
}


def ad05f_elements():
    """Element definitions of the ad05f files in the element-record vocabulary."""
    el = {}
    for n, l in AD05F_DRIFTS.items():
        el[n] = dict(key="drift", l=l)
    for n, (l, k1) in AD05F_QUADS.items():
        el[n] = dict(key="quadrupole", l=l, k1=k1, num_steps=16)
    for n, k2 in AD05F_SEXTUPOLES.items():
        el[n] = dict(key="sextupole", l=0.1, k2=k2, num_steps=16)
    bends = {f"bs{i:02d}": dict(l=l, angle=0.004420919, e1=e1, e2=e2, k1=0.0) for i, (l, e1, e2) in enumerate(AD05F_BS)}
    bends["bn00"] = dict(l=0.1026, angle=0.021629865, e1=0.0, e2=0.021629865, k1=0.0)
    bends["bn01"] = dict(l=0.2774, angle=0.02472608, e1=-0.021629865, e2=0.046355945, k1=-1.98)
    bends["bahq1"] = dict(l=0.2, angle=-0.0113446401, k1=3.240843, e1=-0.00567232007, e2=-0.00567232007)
    bends["bahq2"] = dict(l=0.2, angle=-0.0136135682, k1=4.757043, e1=-0.00680678408, e2=-0.00680678408)
    bends["bahq3"] = dict(l=0.2, angle=-0.0113446401, k1=6.480930, e1=-0.00567232007, e2=-0.00567232007)
    for n, b in list(bends.items()):
        bends[n + "_inv"] = dict(b, e1=b["e2"], e2=b["e1"])
    for n, b in bends.items():
        # fringe_type = basic_bend, hgap = fint = 0; sbend[ds_step] = 0.005
        el[n] = dict(key="sbend", num_steps=max(1, int(math.floor(b["l"] / 0.005 + 0.5))), **b)
    for n, k3l in AD05F_MULTIPOLES.items():
        el[n] = dict(key="multipole", knl={3: k3l})
    for n in AD05F_MARKERS:
        el[n] = dict(key="marker")
    el["mon"] = dict(key="monitor")
    el["ch"] = dict(key="hkicker")
    el["cv"] = dict(key="vkicker")
    # dsend: rfcav (voltage 0.7e7 / 12 V, harmon 96, l 0.88); its sign per file below
    el["dsend"] = dict(key="rfcavity", l=0.88, voltage=0.7e7 / 12, num_steps=4)
    return el


def closed_orbit():
    """Closed orbit z and synchrotron tune of the two ad05f rings (closed_orbit_test)."""
    path = os.path.join(REF, "regression_tests/closed_orbit_test/output.correct")
    rows = {}
    with open(path) as f:
        for lineno, line in enumerate(f, 1):
            m = re.match(r'^"(ad05f_(pos|neg)_RF\.lat-\d-[AB] Z(_tune)?)"\s+ABS\s+(\S+)\s+(\S+)', line.strip())
            if m:
                rows[m.group(1)] = dict(value=float(m.group(5)), tol=float(m.group(4)), line=lineno)
    lines = {k: v.split() for k, v in AD05F_LINES.items()}
    return dict(source="regression_tests/closed_orbit_test (ad05f_pos_RF.lat, ad05f_neg_RF.lat, output.correct)",
                particle="electron", e_tot=2.4e9, harmon=96, elements=ad05f_elements(), lines=lines, use="ring",
                lattices={"ad05f_pos_RF.lat": dict(voltage_sign=1, sextupole_scale=0.5),
                          "ad05f_neg_RF.lat": dict(voltage_sign=-1, sextupole_scale=1.0)},
                rows=rows)


if __name__ == "__main__":
    for name, fn in (("tracking_method_bmad_standard.json", tracking_method), ("fringe_test.json", fringe),
                     ("aperture_test.json", aperture), ("multipole_test.json", multipole),
                     ("mat6_calc_method_test.json", mat6), ("nonlin_small_ring.json", nonlin),
                     ("closed_orbit_ad05f.json", closed_orbit)):
        d = fn()
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        print(name, len(d["rows"]))
