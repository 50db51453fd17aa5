"""Particle snapshots in the reference's ASCII format.

FDPS's writeParticleAscii writes FileHeader::writeAscii (src/class.h:27-30:
"%e\\n" time, "%d\\n" particle count) followed by one RealPtcl::writeAscii line
per particle (src/class.h:177-178): id, tag, mass, pos[3], vel[3], dens, eng,
pres, pot, ent, temp, tab-separated, "%lld" for the integers and "%lf" for the
rest.  read_ascii accepts any whitespace, as fscanf does (src/class.h:187-188).
"""
from __future__ import annotations

import numpy as np

COLUMNS = ["id", "tag", "mass", "x", "y", "z", "vx", "vy", "vz", "dens", "eng", "pres", "pot", "ent", "temp"]


def read_ascii(path: str) -> dict:
    with open(path) as f:
        time = float(f.readline())
        n = int(f.readline())
        data = np.loadtxt(f, ndmin=2, max_rows=n)
    if data.shape[0] != n or data.shape[1] < 15:
        raise ValueError(f"{path}: expected {n} rows of 15 columns, got {data.shape}")
    snap = {"time": time, "id": data[:, 0].astype(np.int64), "tag": data[:, 1].astype(np.int64),
            "mass": data[:, 2].copy(), "pos": data[:, 3:6].copy(), "vel": data[:, 6:9].copy()}
    for k, c in zip(["dens", "eng", "pres", "pot", "ent", "temp"], range(9, 15)):
        snap[k] = data[:, c].copy()
    return snap


def write_ascii(path: str, snap: dict) -> None:
    n = len(snap["mass"])
    z = np.zeros(n)
    with open(path, "w") as f:
        f.write("%e\n" % snap.get("time", 0.0))
        f.write("%d\n" % n)
        for i in range(n):
            vals = [snap["mass"][i], *snap["pos"][i], *snap["vel"][i]] + [
                snap.get(k, z)[i] for k in ("dens", "eng", "pres", "pot", "ent", "temp")]
            f.write("%d\t%d\t" % (snap["id"][i], snap["tag"][i]) + "\t".join("%f" % v for v in vals) + "\n")


def material_from_tag(tag: np.ndarray) -> np.ndarray:
    """GI_universal::setEoS (src/GI.h:447-456): even tag -> granite table
    (material 0), odd tag -> iron table (material 1)."""
    return np.where(np.asarray(tag) % 2 == 0, 0, 1).astype(np.int32)


def to_ic(snap: dict) -> dict:
    """Initial conditions for Simulation from a snapshot (mode-1 resume,
    src/main.cpp:63-66: positions, velocities, mass, density, energy)."""
    return {"pos": snap["pos"], "vel": snap["vel"], "mass": snap["mass"], "dens": snap["dens"], "eng": snap["eng"],
            "mat": material_from_tag(snap["tag"]), "id": snap["id"]}
