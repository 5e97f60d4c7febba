"""Extracts small golden fixtures from the reference's own output frames.

The reference (kaiqideng/this_studio, DEMSPH/) ships ASCII VTU frames that
its CUDA runs wrote: elasticBounce/ball_*.vtu, damBreak/SPH_*.vtu and
iceCylinderInteraction/*.vtu.  These are data files; this script parses them
as text and stores inputs/outputs as compact .npz/.json fixtures in this
directory.  It reads /root/reference only when run by hand; tests read the
fixtures, never the reference.

    python tests/golden/make_fixtures.py          # every fixture
    python tests/golden/make_fixtures.py vtu      # vtu_frames.json only
"""
from __future__ import annotations

import json
import re
from pathlib import Path

import numpy as np

REF = Path("/root/reference/DEMSPH")
OUT = Path(__file__).resolve().parent

_ARR = re.compile(r'<DataArray([^>]*)>(.*?)</DataArray>', re.S)


def read_vtu(path: Path) -> dict:
    txt = path.read_text()
    out = {}
    for attrs, body in _ARR.findall(txt):
        m = re.search(r'Name="([^"]+)"', attrs)
        name = m.group(1) if m else None
        if name is None:
            name = "Points" if "NumberOfComponents=\"3\"" in attrs and "Points" not in out else None
        if name is None:
            continue
        typ = re.search(r'type="([^"]+)"', attrs).group(1)
        vals = np.array(body.split(), dtype=np.int64 if typ.startswith(("Int", "UInt")) else np.float64)
        nc = re.search(r'NumberOfComponents="(\d+)"', attrs)
        if nc and int(nc.group(1)) > 1:
            vals = vals.reshape(-1, int(nc.group(1)))
        out[name] = vals
    return out


def elastic_bounce():
    rows = []
    for f in sorted((REF / "elasticBounce").glob("ball_*.vtu")):
        d = read_vtu(f)
        rows.append([float(d["TIME"][0]), int(d["STEP"][0]), *d["Points"][0].tolist(),
                     *d["velocity"][0].tolist(), *d["angularVelocity"][0].tolist()])
    arr = np.array(rows)
    np.savez_compressed(OUT / "elastic_bounce_frames.npz", frames=arr,
                        columns=np.array(["t", "step", "x", "y", "z", "vx", "vy", "vz", "wx", "wy", "wz"]))
    return len(rows)


def dam_break():
    frames = sorted((REF / "damBreak").glob("SPH_*.vtu"))
    f0 = read_vtu(frames[0])
    f1 = read_vtu(frames[1])
    f2 = read_vtu(frames[2])
    np.savez_compressed(OUT / "dam_break_frames.npz",
                        pos0=f0["Points"].astype(np.float32),
                        pos1=f1["Points"].astype(np.float32), vel1=f1["velocity"].astype(np.float32),
                        rho1=f1["density"].astype(np.float32), p1=f1["pressure"].astype(np.float32),
                        pos2=f2["Points"].astype(np.float32),
                        time=np.array([float(f0["TIME"][0]), float(f1["TIME"][0]), float(f2["TIME"][0])]),
                        step=np.array([int(f0["STEP"][0]), int(f1["STEP"][0]), int(f2["STEP"][0])]))
    # Physics-level series over all frames: front (max x) and column height.
    series = []
    for f in frames:
        d = read_vtu(f)
        p = d["Points"]
        series.append([float(d["TIME"][0]), int(d["STEP"][0]), float(p[:, 0].max()), float(p[:, 2].max()),
                       float(d["pressure"].max()), float(d["density"].mean())])
    np.savez_compressed(OUT / "dam_break_series.npz", series=np.array(series),
                        columns=np.array(["t", "step", "front_x", "max_z", "max_p", "mean_rho"]))
    return len(frames)


def ice_cylinder():
    d = REF / "iceCylinderInteraction"
    w0 = read_vtu(d / "triangleWall_0000.vtu")
    b0 = read_vtu(d / "ball_0000.vtu")
    b1 = read_vtu(d / "ball_0001.vtu")
    bt1 = read_vtu(d / "ballTriangleInteraction_0001.vtu")
    np.savez_compressed(OUT / "ice_cylinder_frames.npz",
                        wall_vertices=w0["Points"].astype(np.float64),
                        wall_connectivity=w0["connectivity"].astype(np.int32),
                        ball_pos0=b0["Points"].astype(np.float64), ball_vel0=b0["velocity"].astype(np.float64),
                        ball_pos1=b1["Points"].astype(np.float32), ball_vel1=b1["velocity"].astype(np.float32),
                        ball_time=np.array([float(b0["TIME"][0]), float(b1["TIME"][0])]),
                        ball_step=np.array([int(b0["STEP"][0]), int(b1["STEP"][0])]),
                        contacts1_json=np.array(json.dumps({k: np.asarray(v).tolist() for k, v in bt1.items()})))
    return b0["Points"].shape[0]


def vtu_frames():
    """Size and SHA-256 of the reference's frame-0 VTU files, whose content is
    the scenes' initial state: a writer in the same layout must reproduce
    them byte for byte.  Also the array layout (name, type, components) of
    every file kind, from frame 1."""
    import hashlib
    files = ["elasticBounce/ball_0000.vtu", "elasticBounce/triangleWall_0000.vtu", "damBreak/SPH_0000.vtu",
             "iceCylinderInteraction/ball_0000.vtu", "iceCylinderInteraction/triangleWall_0000.vtu",
             "iceCylinderInteraction/ballTriangleInteraction_0000.vtu"]
    out = {"frame0": {}, "layout": {}}
    for rel in files:
        b = (REF / rel).read_bytes()
        out["frame0"][rel] = {"bytes": len(b), "sha256": hashlib.sha256(b).hexdigest(),
                              "lines": b.count(b"\n")}
    arr = re.compile(r'<DataArray([^>]*)>')
    for rel in ["elasticBounce/ball_0001.vtu", "elasticBounce/triangleWall_0001.vtu", "damBreak/SPH_0001.vtu",
                "iceCylinderInteraction/ballTriangleInteraction_0001.vtu"]:
        txt = (REF / rel).read_text()
        rows = []
        for attrs in arr.findall(txt):
            name = re.search(r'Name="([^"]+)"', attrs)
            comp = re.search(r'NumberOfComponents="(\d+)"', attrs)
            rows.append([name.group(1) if name else "", re.search(r'type="([^"]+)"', attrs).group(1),
                         int(comp.group(1)) if comp else 1])
        out["layout"][rel.split("/")[1].rsplit("_", 1)[0]] = rows
    (OUT / "vtu_frames.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    return len(files)


if __name__ == "__main__":
    import sys
    if sys.argv[1:] == ["vtu"]:
        print("VTU frame-0 files:", vtu_frames())
    else:
        print("elasticBounce frames:", elastic_bounce())
        print("damBreak frames:", dam_break())
        print("iceCylinder balls:", ice_cylinder())
        print("VTU frame-0 files:", vtu_frames())
