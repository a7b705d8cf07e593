"""Copies the reference's scene tables (data/bands.csv, atmosphere_profile.csv,
aerosol_profile.csv, aerosol_optics.csv: band set with solar irradiance and
ozone cross sections, the AFGL atmosphere profile, the OPAC aerosol profile
and optics) into tests/golden/reference_data/, so that the tabulated-medium
tests run where the reference checkout is absent.  data/mie_phase.csv, which
the reference's loader also expects, is not in its checkout.

    python tests/golden/make_reference_data.py /root/reference
"""
import os
import shutil
import sys

FILES = ("bands.csv", "atmosphere_profile.csv", "aerosol_profile.csv", "aerosol_optics.csv")


def main(ref_root):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_data")
    os.makedirs(out, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(ref_root, "data", f), os.path.join(out, f))
        print("copied", f)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
