"""Point-cloud export of simulation frames for an external viewer.

The reference hands its particle buffers to an OpenGL renderer
(fluids/Simulator.cpp:16-36 maps the five GL buffers around every step,
fluids/SSFRenderer.cpp:113-146 draws them as screen-space fluid).  Here the
state lives in device arrays owned by the caller; a frame is exported as a
binary little-endian PLY point cloud (ParaView, MeshLab, Blender and Open3D
read it), one vertex per particle with its position, speed and identity.

A step reorders the particles by cell (the sort permutes the state and iid
carries identity), so frames are written in identity order by default: vertex
k of every frame is the same particle, which lets a viewer animate a sequence
or colour by a per-particle attribute.
"""
from typing import Optional, Sequence

import numpy as np

_FIELDS = (("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("speed", "<f4"), ("iid", "<u4"))
_PLY_TYPES = {"<f4": "float", "<u4": "uint"}


def _host(a) -> np.ndarray:
    """A host numpy view of a numpy array or a (device) torch tensor."""
    if hasattr(a, "detach"):
        a = a.detach().to("cpu").numpy()
    return np.asarray(a)


def frame_vertices(pos, vel, iid, nparticle: Optional[int] = None,
                   identity_order: bool = True) -> np.ndarray:
    """The structured vertex array of one frame.

    pos, vel: (n, 3) or (n, 4) float32 (the library's layouts); iid: (n,)
    uint32 identities.  With identity_order the vertices are sorted by iid.
    """
    p, v, ids = _host(pos), _host(vel), _host(iid)
    n = p.shape[0] if nparticle is None else int(nparticle)
    if p.ndim != 2 or p.shape[1] not in (3, 4) or v.shape != p.shape or p.shape[0] < n:
        raise ValueError("pos and vel must be (n, 3) or (n, 4) arrays of the same shape")
    if ids.shape[0] < n:
        raise ValueError("iid holds fewer than nparticle entries")
    p = p[:n, :3].astype(np.float32, copy=False)
    v = v[:n, :3].astype(np.float32, copy=False)
    ids = ids[:n].astype(np.uint32, copy=False)
    out = np.empty(n, dtype=np.dtype(list(_FIELDS)))
    out["x"], out["y"], out["z"] = p[:, 0], p[:, 1], p[:, 2]
    out["speed"] = np.sqrt(np.einsum("ij,ij->i", v.astype(np.float64), v.astype(np.float64)))
    out["iid"] = ids
    if identity_order:
        out = out[np.argsort(ids, kind="stable")]
    return out


def write_ply(path: str, pos, vel, iid, nparticle: Optional[int] = None,
              identity_order: bool = True, comments: Sequence[str] = ()) -> None:
    """Writes one frame as a binary little-endian PLY point cloud with vertex
    properties x, y, z, speed (float) and iid (uint)."""
    verts = frame_vertices(pos, vel, iid, nparticle, identity_order)
    lines = ["ply", "format binary_little_endian 1.0"]
    lines += [f"comment {c}" for c in comments]
    lines.append(f"element vertex {len(verts)}")
    lines += [f"property {_PLY_TYPES[t]} {name}" for name, t in _FIELDS]
    lines.append("end_header")
    with open(path, "wb") as f:
        f.write(("\n".join(lines) + "\n").encode("ascii"))
        f.write(verts.tobytes())


def read_ply(path: str) -> np.ndarray:
    """Reads a file written by write_ply back into its structured vertex array."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.find(b"end_header\n")
    if not data.startswith(b"ply\n") or end < 0:
        raise ValueError("not a PLY file")
    header = data[:end].decode("ascii").split("\n")
    if "format binary_little_endian 1.0" not in header:
        raise ValueError("only binary little-endian PLY is read")
    n = None
    props = []
    for line in header:
        w = line.split()
        if w[:2] == ["element", "vertex"]:
            n = int(w[2])
        elif w[:1] == ["property"]:
            props.append((w[2], {"float": "<f4", "uint": "<u4"}[w[1]]))
    if n is None or tuple(props) != _FIELDS:
        raise ValueError("unexpected PLY vertex layout")
    body = data[end + len(b"end_header\n"):]
    dt = np.dtype(list(_FIELDS))
    if len(body) != n * dt.itemsize:
        raise ValueError(f"PLY body has {len(body)} bytes, expected {n * dt.itemsize}")
    return np.frombuffer(body, dtype=dt, count=n).copy()
