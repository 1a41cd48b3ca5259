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


if __name__ == "__main__":
    for name, fn in (("tracking_method_bmad_standard.json", tracking_method), ("fringe_test.json", fringe),
                     ("aperture_test.json", aperture), ("multipole_test.json", multipole),
                     ("mat6_calc_method_test.json", mat6), ("nonlin_small_ring.json", nonlin)):
        d = fn()
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        print(name, len(d["rows"]))
