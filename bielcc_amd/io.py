"""Writers in the reference's output formats (io/io_vtk.h, io/io_aereco.h),
for results computed on the device and brought back as numpy arrays.

Numbers are written as C++ streams write a double by default (printf "%g",
6 significant digits), so files match the reference's byte for byte.
"""
from __future__ import annotations

import math

import numpy as np


def _g(x: float) -> str:
    return "%g" % x


def _row(v) -> str:
    return " ".join(_g(float(x)) for x in v)


def write_field_vtk_c(path: str, points, vals) -> None:
    """IO_VTK::WritePntsWithField_C (io/io_vtk.cpp:482-551): POLYDATA points with
    the real part, imaginary part, per-component magnitude and phase of a
    complex 3-vector per point.  points [n][3], vals [n][3] complex."""
    points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    vals = np.asarray(vals, dtype=np.complex128).reshape(-1, 3)
    if points.shape[0] != vals.shape[0]:
        raise ValueError("Points and field vectors must have same size")
    out = ["# vtk DataFile Version 3.0", "Points with complex vector field", "ASCII", "DATASET POLYDATA",
           f"POINTS {points.shape[0]} double"]
    out += [_row(p) for p in points]
    out.append(f"POINT_DATA {vals.shape[0]}")
    out.append("VECTORS real_part double")
    out += [_row(v.real) for v in vals]
    out.append("VECTORS image_part double")
    out += [_row(v.imag) for v in vals]
    out.append("VECTORS magnitude double")
    out += [_row([abs(c) for c in v]) for v in vals]
    for g, name in enumerate("xyz"):
        out += [f"SCALARS phase_{name} double 1", "LOOKUP_TABLE default"]
        out += [_g(math.atan2(v[g].imag, v[g].real)) for v in vals]
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_aereco_vecs_c(path_real: str, path_imag: str, vals) -> None:
    """IO_AERECO::WriteVecs_C (io/io_aereco.cpp:98-128): header "2 n/2", one
    vector per line, real and imaginary parts in two files."""
    vals = np.asarray(vals, dtype=np.complex128).reshape(-1, 3)
    head = f"2 {vals.shape[0] // 2}"
    with open(path_real, "w") as f:
        f.write("\n".join([head] + [_row(v.real) for v in vals]) + "\n")
    with open(path_imag, "w") as f:
        f.write("\n".join([head] + [_row(v.imag) for v in vals]) + "\n")


def write_aereco_field_c(path_real: str, path_imag: str, path_abs: str, vals) -> None:
    """IO_AERECO::WriteField_C (io/io_aereco.cpp:135-173): header "n1 n1" with
    n1 = n/2 and the first n1^2 vectors (real, imaginary, |v|).  The reference
    reads past the end of its input when n1^2 > n; that case raises here."""
    vals = np.asarray(vals, dtype=np.complex128).reshape(-1, 3)
    n1 = vals.shape[0] // 2
    if n1 * n1 > vals.shape[0]:
        raise ValueError("(n/2)^2 rows exceed the values given")
    head = f"{n1} {n1}"
    rows = vals[: n1 * n1]
    mags = [math.sqrt(abs(v[0]) * abs(v[0]) + abs(v[1]) * abs(v[1]) + abs(v[2]) * abs(v[2])) for v in rows]
    for path, lines in ((path_real, [_row(v.real) for v in rows]), (path_imag, [_row(v.imag) for v in rows]),
                        (path_abs, [_g(m) for m in mags])):
        with open(path, "w") as f:
            f.write("\n".join([head] + lines) + "\n")
