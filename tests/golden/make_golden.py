"""Writes the committed fixtures of tests/golden/.

  rng_vectors.json  pcg4d / rng_u32 / rng_f32 golden vectors from an
                    independent Python-integer restatement of
                    crates/sky-spectral-path-tracer/src/sky_trace.wgsl:156-189.
  cie_1931.json     the reference's CIE 1931 2-degree table
                    (data/CIE_xyz_1931_2deg.csv, the file spectrum.rs:255
                    includes) at every 5 nm, plus the table's linear
                    interpolation (spectrum.rs:196-219) at the 16- and
                    64-wavelength midpoint grids.

Run from the repository root in the build container (reads the reference
checkout):  python tests/golden/make_golden.py
"""
import csv
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
REF_CSV = "/root/reference/data/CIE_xyz_1931_2deg.csv"
M32 = 0xFFFFFFFF


def pcg4d_py(v):
    """sky_trace.wgsl:156-168, vec4u arithmetic modulo 2^32."""
    v = [(x * 1664525 + 1013904223) & M32 for x in v]
    x, y, z, w = v
    x = (x + y * w) & M32
    y = (y + z * x) & M32
    z = (z + x * y) & M32
    w = (w + y * z) & M32
    x ^= x >> 16
    y ^= y >> 16
    z ^= z >> 16
    w ^= w >> 16
    x = (x + y * w) & M32
    y = (y + z * x) & M32
    z = (z + x * y) & M32
    w = (w + y * z) & M32
    return [x, y, z, w]


def rng_u32_py(pixel, sample, band, dim, seed):
    """sky_trace.wgsl:174-184: counter (pixel ^ seed_lo, sample ^ seed_hi, band, dim)."""
    v = pcg4d_py([pixel ^ (seed & M32), sample ^ (seed >> 32), band, dim])
    return v[0] ^ v[1] ^ v[2] ^ v[3]


def rng_f32_py(*key):
    """sky_trace.wgsl:186-189: (u >> 8) * 2^-24, exact in binary32."""
    return (rng_u32_py(*key) >> 8) / 16777216.0


def rng_fixture():
    keys = [
        (0, 0, 0, 0, 0),
        (0, 0, 0, 1, 1),
        (1, 2, 3, 4, 1),
        (65535, 1023, 63, 17, 1),
        (524287, 1023, 63, 99, 0x5EC720260430),
        (123456, 7, 40, 5, 0x5EC720260430),
        (M32, M32, M32, M32, 0xFFFFFFFFFFFFFFFF),
        (8191, 63, 15, 2, 7),
    ]
    raw = [[0, 0, 0, 0], [1, 2, 3, 4], [M32, 0, M32, 0], [0xDEADBEEF, 0x01234567, 0x89ABCDEF, 42]]
    return {
        "source": "independent Python-integer restatement of sky_trace.wgsl:156-189",
        "pcg4d": [{"in": r, "out": pcg4d_py(r)} for r in raw],
        "rng": [{"pixel": k[0], "sample": k[1], "band": k[2], "dim": k[3], "seed": k[4],
                 "u32": rng_u32_py(*k), "f32": rng_f32_py(*k)} for k in keys],
    }


def cie_fixture():
    rows = []
    with open(REF_CSV) as f:
        for r in csv.reader(f):
            if not r or r[0].strip().startswith("#"):
                continue
            try:
                rows.append([float(x) for x in r[:4]])
            except ValueError:
                continue
    table = {int(round(r[0])): r[1:] for r in rows}

    def lerp_table(lam):
        lo = min(table)
        hi = max(table)
        if lam <= lo:
            return table[lo]
        if lam >= hi:
            return table[hi]
        a = int(lam // 1)
        b = a + 1
        t = lam - a
        return [table[a][i] * (1 - t) + table[b][i] * t for i in range(3)]

    def grid(n):
        step = 400.0 / n
        return [380.0 + step * (i + 0.5) for i in range(n)]

    return {
        "source": "data/CIE_xyz_1931_2deg.csv of the reference (CIE 1931 2-degree observer, 1 nm)",
        "every_5nm": [[lam] + table[lam] for lam in range(360, 831, 5)],
        "grid16": [[l] + lerp_table(l) for l in grid(16)],
        "grid64": [[l] + lerp_table(l) for l in grid(64)],
        "y_555": table[555][1],
    }


def main():
    with open(os.path.join(HERE, "rng_vectors.json"), "w") as f:
        json.dump(rng_fixture(), f, indent=1)
    with open(os.path.join(HERE, "cie_1931.json"), "w") as f:
        json.dump(cie_fixture(), f, indent=1)
    print("wrote rng_vectors.json, cie_1931.json")


if __name__ == "__main__":
    main()
