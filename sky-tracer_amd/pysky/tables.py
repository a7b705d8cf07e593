"""Tabulated medium (the reference's SceneData tables) for the C ABI's
sky_medium_tables, and a loader for the reference's data directory
(crates/sky-core/src/data.rs:15-152):

    bands.csv              index, center_nm, lower_nm, upper_nm, solar_irradiance_w_m2, ozone_cross_section_cm2
    atmosphere_profile.csv altitude_km, temperature_k, air_cm3, ozone_cm3
    aerosol_profile.csv    altitude_km, mass_g_m3 of inso, waso, soot, suso
    aerosol_optics.csv     species, band, scattering, absorption (km^-1 per g/m^3)

The reference also reads mie_phase.csv, which its checkout does not contain;
the aerosol phase here is Henyey-Greenstein with AtmosphereParams.aerosol_g.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L

SPECIES = ("inso", "waso", "soot", "suso")  # atmosphere.rs:8
BAND_COUNT = 41  # spectrum.rs:5
GROUND_ALBEDO = 0.18  # atmosphere.rs:9
SOLAR_ANGULAR_RADIUS_RAD = 0.00465047  # atmosphere.rs:10
REFERENCE_ATMOSPHERE_RADIUS_KM = 6480.0  # geometry.rs:15


@dataclass
class MediumTables:
    atmosphere: np.ndarray       # [n, 3]: altitude_km, air_cm3, ozone_cm3 (ascending altitude)
    aerosol: np.ndarray          # [n, 5]: altitude_km, mass_g_m3 x 4
    aerosol_optics: np.ndarray   # [bands, 4, 2]: scattering, absorption
    ozone_cross_section_cm2: np.ndarray
    rayleigh_cross_section_m2: Optional[np.ndarray] = None
    solar_irradiance_w_m2: Optional[np.ndarray] = None

    def to_c(self) -> L.MediumTablesC:
        arrs = {k: np.ascontiguousarray(getattr(self, k), dtype=np.float64)
                for k in ("atmosphere", "aerosol", "aerosol_optics", "ozone_cross_section_cm2")}
        for k in ("rayleigh_cross_section_m2", "solar_irradiance_w_m2"):
            v = getattr(self, k)
            arrs[k] = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        t = L.MediumTablesC()
        t.atmosphere_points = arrs["atmosphere"].shape[0]
        t.aerosol_points = arrs["aerosol"].shape[0]
        for k, a in arrs.items():
            setattr(t, k, None if a is None else a.ctypes.data_as(C.POINTER(C.c_double)))
        t._keep = arrs
        return t


def _rows(path):
    """Comma-separated rows, skipping blanks and '#' comments (data.rs:154-170)."""
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line and not line.startswith("#"):
                out.append([c.strip() for c in line.split(",")])
    return out


def load_tables(data_dir: str):
    """(band centres nm, MediumTables) from the reference's data directory."""
    bands = _rows(os.path.join(data_dir, "bands.csv"))
    if any(len(r) != 6 for r in bands):
        raise ValueError("bands.csv: expected 6 columns")
    if len(bands) != BAND_COUNT:
        raise ValueError(f"expected {BAND_COUNT} bands, found {len(bands)}")
    centers = np.array([float(r[1]) for r in bands])
    atm = np.array([[float(v) for v in r] for r in _rows(os.path.join(data_dir, "atmosphere_profile.csv"))])
    if atm.ndim != 2 or atm.shape[1] != 4:
        raise ValueError("atmosphere_profile.csv: expected 4 columns")
    aer = np.array([[float(v) for v in r] for r in _rows(os.path.join(data_dir, "aerosol_profile.csv"))])
    if aer.ndim != 2 or aer.shape[1] != 5:
        raise ValueError("aerosol_profile.csv: expected 5 columns")
    optics = np.zeros((len(bands), 4, 2))
    seen = np.zeros((len(bands), 4), bool)
    for r in _rows(os.path.join(data_dir, "aerosol_optics.csv")):
        if len(r) != 4:
            raise ValueError("aerosol_optics.csv: expected 4 columns")
        sp, b = SPECIES.index(r[0]), int(r[1])
        optics[b, sp] = (float(r[2]), float(r[3]))
        seen[b, sp] = True
    if not seen.all():
        raise ValueError("aerosol_optics.csv: missing entries")
    atm = atm[np.argsort(atm[:, 0], kind="stable")]  # sorted by altitude (data.rs:88)
    aer = aer[np.argsort(aer[:, 0], kind="stable")]
    t = MediumTables(atmosphere=atm[:, [0, 2, 3]], aerosol=aer, aerosol_optics=optics,
                     ozone_cross_section_cm2=np.array([float(r[5]) for r in bands]),
                     solar_irradiance_w_m2=np.array([float(r[4]) for r in bands]))
    return centers, t


def load_reference_scene(data_dir: str, sun_elevation_deg: float = 0.0, sun_azimuth_deg: float = 0.0,
                         aerosol_g: float = 0.8):
    """load_scene_data (data.rs:15-51) over the tabulated medium: the
    reference's planet (6360 / 6480 km), sun radius, ground albedo and band
    set.  Film values are band-integrated radiances (W m^-2 sr^-1), so the
    XYZ weights are the colour-matching functions alone (weight 1 per band,
    spectrum.rs:226-241)."""
    from .api import AtmosphereParams, SceneData

    centers, tables = load_tables(data_dir)
    atm = AtmosphereParams(aerosol_g=aerosol_g, ground_albedo=GROUND_ALBEDO,
                           sun_angular_radius_rad=SOLAR_ANGULAR_RADIUS_RAD, tables=tables)
    return SceneData(atmosphere_radius_km=REFERENCE_ATMOSPHERE_RADIUS_KM, atmosphere=atm,
                     sun_elevation_deg=sun_elevation_deg, sun_azimuth_deg=sun_azimuth_deg,
                     wavelengths_nm=list(centers), weights_nm=[1.0] * len(centers))
