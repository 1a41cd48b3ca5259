"""Elements written the way Bmad keeps them, for the tests of the Fortran
shim's flattener (bmad-ecosystem_amd/fortran/bmx_shim_core.f90): the Bmad
element key, the ele_struct flags and ele%value(1:75) with Bmad's attribute
indices (bmad/modules/bmad_struct.f90:1681-1693 keys, :1775-1881 indices),
plus ele%ref_time and the pole arrays.  This is the inverse of the
flattener, written from the same table, so that a round trip through the
Fortran code must give back the records byte for byte."""
import struct

import numpy as np

from bmx import abi

BMAD_KEY = {abi.KEY_MARKER: 14, abi.KEY_DRIFT: 1, abi.KEY_QUADRUPOLE: 3, abi.KEY_SBEND: 2, abi.KEY_SEXTUPOLE: 5,
            abi.KEY_OCTUPOLE: 17, abi.KEY_THICK_MULTIPOLE: 67, abi.KEY_MULTIPOLE: 19, abi.KEY_AB_MULTIPOLE: 22,
            abi.KEY_RFCAVITY: 9, abi.KEY_HKICKER: 33, abi.KEY_VKICKER: 34, abi.KEY_KICKER: 15,
            abi.KEY_INSTRUMENT: 32, abi.KEY_MONITOR: 31, abi.KEY_ECOLLIMATOR: 36, abi.KEY_RCOLLIMATOR: 35,
            abi.KEY_PIPE: 44}

# ele%value indices shared by every key (*_tot values: the offsets the flattener reads)
COMMON = dict(l=1, num_steps=66, exact_multipoles=14, fringe_type=10, fringe_at=11, p0c=53, e_tot=54,
              p0c_start=51, e_tot_start=52, delta_ref_time=50, x_offset=57, y_offset=58, z_offset=59,
              x_pitch=55, y_pitch=56, x1_limit=71, x2_limit=72, y1_limit=73, y2_limit=74, hkick=39, vkick=40)
BY_KEY = {
    abi.KEY_SBEND: dict(g=6, dg=7, k1=4, k2=5, e1=19, e2=20, fint=21, fintx=22, hgap=23, hgapx=24, ref_tilt=61,
                        r0_mag=16, r0_elec=15),
    abi.KEY_RFCAVITY: dict(voltage=8, rf_frequency=15, phi0=24, phi0_multipass=26, coupler_strength=46,
                           coupler_angle=45, coupler_phase=44, coupler_at=47, tilt=60),
    abi.KEY_HKICKER: dict(kick=3, tilt=60, r0_mag=16, r0_elec=15),
    abi.KEY_VKICKER: dict(kick=3, tilt=60, r0_mag=16, r0_elec=15),
    abi.KEY_ECOLLIMATOR: dict(tilt=60, px_aperture_width2=24, px_aperture_center=25, py_aperture_width2=26,
                              py_aperture_center=27, z_aperture_width2=28, z_aperture_center=29,
                              pz_aperture_width2=30, pz_aperture_center=31, r0_mag=16, r0_elec=15),
}
BY_KEY[abi.KEY_RCOLLIMATOR] = BY_KEY[abi.KEY_ECOLLIMATOR]
DEFAULT = dict(k1=4, k2=5, k3=6, tilt=60, r0_mag=16, r0_elec=15)
INT_ATTRS = {"num_steps", "exact_multipoles", "fringe_type", "fringe_at", "coupler_at"}


def bmad_record(e):
    """bytes of one element in the layout the Fortran driver reads: 9 int32
    (Bmad key, is_on, orientation, field_master, scale_multipoles,
    aperture_type, aperture_at, has_mag, has_elec), value(1:75), ref_time,
    a_pole, b_pole, a_pole_elec, b_pole_elec (22 each)."""
    value = np.zeros(75)
    table = dict(COMMON)
    table.update(BY_KEY.get(e.key, DEFAULT))
    for name, ix in table.items():
        v = getattr(e, name)
        value[ix - 1] = float(v)
    ints = [BMAD_KEY[e.key], e.is_on, e.orientation, e.field_master, e.scale_multipoles, e.aperture_type,
            e.aperture_at, e.has_mag_poles, e.has_elec_poles]
    out = struct.pack("<9i", *ints) + value.astype("<f8").tobytes() + struct.pack("<d", e.ref_time)
    for arr in (e.a, e.b, e.a_elec, e.b_elec):
        out += np.array(arr[:], dtype="<f8").tobytes()
    return out
