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


def check_loaded_state(params, state) -> bool:
    """The checks fileInput runs on the loaded particles (Input.cpp:292-318):
    check_finite on r, v, rho (ConsistencyCheck.cpp:78-119; p is not in a
    dump and starts at 0) and check_particles_are_within_boundaries
    (:146-196, strictly inside the geometry box).  With consistency_treatment
    STOP the first failing check raises the reference's message; otherwise
    it warns and the state is kept as loaded.  True when both pass."""
    import warnings

    from . import params as P

    r, v, rho = state["r"], state["v"], state["rho"]
    dim = r.shape[1]
    lo = np.array([params.x_mingeom, params.y_mingeom, params.z_mingeom][:dim])
    hi = np.array([params.x_maxgeom, params.y_maxgeom, params.z_maxgeom][:dim])
    finite = np.all(np.isfinite(r), axis=1) & np.all(np.isfinite(v), axis=1) & np.isfinite(rho)
    inside = np.all((r > lo) & (r < hi), axis=1) & np.all(np.isfinite(r), axis=1)
    ok = True
    for good, what in ((finite, "Encounter not finite value"), (inside, "Encounter particle outside boundaries")):
        bad = np.flatnonzero(~good)
        if len(bad) == 0:
            continue
        ok = False
        k = int(bad[-1])
        message = (f"{what}\n\t type: {int(state['itype'][k])}\n\t r: {tuple(r[k])}\n\t k: {k}\n"
                   f"\t total_error_particles: {len(bad)}")
        if params.consistency_treatment == P.CONSISTENCY_STOP:
            raise RuntimeError(message)
        warnings.warn(what)
    return ok


def file_input(initial_dump_path: str, experiment_directory: str, specific_version=(0, 0, 0, "RRSPH_HIP"),
               write_params: bool = True):
    """fileInput (Input.cpp:238-331): the experiment's ParticleParams and
    ModelParams, the start time from the dump's file stem, the derived
    constants, the dump itself (its row count must be ntotal), the depth of
    its fluid, the load-time consistency checks, the post-fill constants,
    and, as the reference does, ComputingParams.json / ModelParams.json /
    ParticleParams.json rewritten into the experiment directory.  Returns
    (params, state) with state = r, v, rho, itype (and p when the dump has
    it)."""
    from . import params as P

    pp = P.load_particle_params(experiment_directory)
    raw_model = P._read_json(os.path.join(experiment_directory, "ModelParams.json"))
    mp = P.parse_model_params(raw_model)
    prm = P.make_params(mp, pp, start_time=P.start_time_from_dump_name(initial_dump_path))
    state = read_dump(initial_dump_path, dim=prm.dim, ntotal=prm.ntotal)
    prm.depth = P.find_depth(prm, state["r"])
    pp["depth"] = prm.depth
    check_loaded_state(prm, state)
    P.post_fill_in_model_params(prm)
    if write_params:
        P.write_experiment_params(experiment_directory, mp, pp, prm, specific_version, raw_model)
    return prm, state
