"""Particle dump / layer CSV format of the reference.

Reading follows Input.cpp:203-236 (columns looked up by name: x, y[, z],
itype, vx, vy[, vz], rho; p is not read).  Writing follows Output.cpp:81-167:
header order x, y[, z], itype, vx, vy[, vz], rho, p, values printed with
std::to_string (six decimals).  ``write_dump(..., lossless=True)`` prints
%.17g instead so that a dump round-trips bit-exactly.
"""
from __future__ import annotations

import csv
import os
from typing import Dict, Optional

import numpy as np


def read_dump(path: str, dim: int = 2, ntotal: Optional[int] = None) -> Dict[str, np.ndarray]:
    with open(path, newline="") as f:
        reader = csv.reader(f)
        header = [h.strip() for h in next(reader)]
        rows = [row for row in reader if row]
    col = {name: k for k, name in enumerate(header)}
    n = len(rows)
    if ntotal is not None and n != ntotal:
        raise RuntimeError("dump corrupted: dump ntotal doesn't match params.ntotal!")
    names = ["x", "y", "z"][:dim]
    vnames = ["vx", "vy", "vz"][:dim]
    data = np.array(rows, dtype=object)
    r = np.empty((n, dim), dtype=np.float64)
    v = np.empty((n, dim), dtype=np.float64)
    for k, nm in enumerate(names):
        r[:, k] = data[:, col[nm]].astype(np.float64)
    for k, nm in enumerate(vnames):
        v[:, k] = data[:, col[nm]].astype(np.float64)
    rho = data[:, col["rho"]].astype(np.float64)
    # itype is parsed as an integer (row["itype"].get<rr_int>())
    itype = np.array([int(float(x)) for x in data[:, col["itype"]]], dtype=np.int32)
    out = {"r": r, "v": v, "rho": rho, "itype": itype}
    if "p" in col:
        out["p"] = data[:, col["p"]].astype(np.float64)
    return out


def _fmt6(x: float) -> str:
    return "%f" % x


def write_dump(path: str, r: np.ndarray, itype: np.ndarray, v: Optional[np.ndarray] = None,
               rho: Optional[np.ndarray] = None, p: Optional[np.ndarray] = None, lossless: bool = False) -> None:
    """Output.cpp:81-167 (make_csv_header + print_dump)."""
    dim = r.shape[1]
    header = ["x", "y", "z"][:dim] + ["itype"]
    if v is not None:
        header += ["vx", "vy", "vz"][:dim]
    if rho is not None:
        header += ["rho"]
    if p is not None:
        header += ["p"]
    fmt = (lambda x: repr(float(x))) if lossless else _fmt6
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f, lineterminator="\n")
        w.writerow(header)
        for i in range(r.shape[0]):
            row = [fmt(x) for x in r[i]] + [str(int(itype[i]))]
            if v is not None:
                row += [fmt(x) for x in v[i]]
            if rho is not None:
                row.append(fmt(rho[i]))
            if p is not None:
                row.append(fmt(p[i]))
            w.writerow(row)
