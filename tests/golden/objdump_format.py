"""Reader of the reference loader's dump (oracle/ref_objdump.cpp output)."""
import numpy as np


def _vertices(lines, n):
    a = np.array([[float.fromhex(x) for x in lines[i].split()] for i in range(n)], np.float32).reshape(n, 8)
    return a[:, 0:3].copy(), a[:, 3:6].copy(), a[:, 6:8].copy()


def _tris(lines, n):
    return np.array([[int(x) for x in lines[i].split()] for i in range(n)], np.uint32).reshape(n, 3)


def parse(text: str) -> dict:
    lines = text.splitlines()
    V, T, G = (int(x) for x in lines[0].split())
    k = 1
    out = {}
    out["pos"], out["nrm"], out["uv"] = _vertices(lines[k:], V)
    k += V
    out["idx"] = _tris(lines[k:], T)
    k += T
    groups = []
    for g in range(G):
        w = lines[k].split()
        k += 1
        nv, nt = (int(x) for x in lines[k].split())
        k += 1
        gp, gn, gu = _vertices(lines[k:], nv)
        k += nv
        gx = _tris(lines[k:], nt)
        k += nt
        groups.append(dict(first=int(w[0]), count=int(w[1]), alpha=float.fromhex(w[2]),
                           shininess=float.fromhex(w[3]), diffuse=[float.fromhex(x) for x in w[4:7]],
                           color_map="" if w[7] == "-" else w[7], pos=gp, nrm=gn, uv=gu, idx=gx))
    out["groups"] = groups
    return out


def to_npz(d: dict, path):
    arrs = {k: d[k] for k in ("pos", "nrm", "uv", "idx")}
    meta = []
    for g, gr in enumerate(d["groups"]):
        meta.append([gr["first"], gr["count"], gr["alpha"], gr["shininess"], *gr["diffuse"]])
        for k in ("pos", "nrm", "uv", "idx"):
            arrs[f"g{g}_{k}"] = gr[k]
    arrs["groups"] = np.array(meta, np.float64).reshape(len(meta), 7)
    arrs["color_maps"] = np.array([gr["color_map"] for gr in d["groups"]] or [""])
    np.savez_compressed(path, **arrs)


def from_npz(path) -> dict:
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in ("pos", "nrm", "uv", "idx")}
    groups = []
    for g, row in enumerate(z["groups"]):
        groups.append(dict(first=int(row[0]), count=int(row[1]), alpha=float(row[2]), shininess=float(row[3]),
                           diffuse=[float(x) for x in row[4:7]], color_map=str(z["color_maps"][g]),
                           **{k: z[f"g{g}_{k}"] for k in ("pos", "nrm", "uv", "idx")}))
    d["groups"] = groups
    return d
