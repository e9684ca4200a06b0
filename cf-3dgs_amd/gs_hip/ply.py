"""The Gaussian PLY files of CFGaussianModel.save_ply / load_ply
(scene/gaussian_model_cf.py:398-432, 459-515).

One ``vertex`` element, every property a 4-byte float, in the order of
``construct_list_of_attributes`` (:398-410): x y z nx ny nz (normals are
zeros), f_dc_0..2, f_rest_0..3(K-1)-1, opacity, scale_0..2, rot_0..3.  The
SH coefficients are written channel-major -- ``features.transpose(1, 2)
.flatten(1)`` -- so f_rest_0..K-2 hold the red channel of coefficients
1..K-1, then green, then blue.  Values are the raw (pre-activation)
parameters.  The reference writes through ``plyfile`` in binary little-endian
format, which is not installed here; this module writes that format
directly (the PLY header, then one little-endian float32 record per vertex)
and reads binary (either byte order) and ASCII vertex data.
"""
from __future__ import annotations

import os
from typing import Dict, List

import numpy as np

_PLY_TYPES = {
    "char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2", "ushort": "u2",
    "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4",
    "double": "f8", "float64": "f8",
}


def attribute_names(n_rest_coeffs: int) -> List[str]:
    """construct_list_of_attributes for 1 DC and ``n_rest_coeffs`` further
    coefficients per channel."""
    names = ["x", "y", "z", "nx", "ny", "nz"]
    names += [f"f_dc_{i}" for i in range(3)]
    names += [f"f_rest_{i}" for i in range(3 * n_rest_coeffs)]
    names += ["opacity"]
    names += [f"scale_{i}" for i in range(3)]
    names += [f"rot_{i}" for i in range(4)]
    return names


def save_ply(path: str, xyz, features_dc, features_rest, opacity, scaling, rotation) -> None:
    """Write the raw parameters (torch tensors or arrays): xyz [P,3],
    features_dc [P,1,3], features_rest [P,K-1,3], opacity [P,1],
    scaling [P,3], rotation [P,4]."""
    def host(t):
        t = t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
        return np.asarray(t, dtype=np.float32)

    xyz = host(xyz)
    P = xyz.shape[0]
    f_dc = np.ascontiguousarray(host(features_dc).transpose(0, 2, 1)).reshape(P, -1)
    f_rest = np.ascontiguousarray(host(features_rest).transpose(0, 2, 1)).reshape(P, -1)
    cols = np.concatenate([xyz, np.zeros_like(xyz), f_dc, f_rest, host(opacity).reshape(P, 1),
                           host(scaling).reshape(P, 3), host(rotation).reshape(P, 4)], axis=1)
    names = attribute_names(f_rest.shape[1] // 3)
    assert cols.shape[1] == len(names)
    header = "ply\nformat binary_little_endian 1.0\nelement vertex %d\n" % P
    header += "".join(f"property float {n}\n" for n in names) + "end_header\n"
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(header.encode("ascii"))
        f.write(np.ascontiguousarray(cols, dtype="<f4").tobytes())


def read_vertices(path: str) -> np.ndarray:
    """The ``vertex`` element of a PLY file as a numpy structured array."""
    with open(path, "rb") as f:
        data = f.read()
    end = data.find(b"end_header")
    if not data.startswith(b"ply") or end < 0:
        raise ValueError(f"{path}: not a PLY file")
    nl = data.find(b"\n", end)
    body = data[nl + 1:]
    fmt, elements, cur = None, [], None
    for raw in data[:end].decode("ascii").splitlines():
        tok = raw.split()
        if not tok or tok[0] in ("ply", "comment", "obj_info"):
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            cur = {"name": tok[1], "count": int(tok[2]), "props": []}
            elements.append(cur)
        elif tok[0] == "property":
            if tok[1] == "list":
                raise ValueError(f"{path}: list properties are not supported")
            cur["props"].append((tok[2], _PLY_TYPES[tok[1]]))
    if not elements or elements[0]["name"] != "vertex":
        raise ValueError(f"{path}: the first element is not 'vertex'")
    v = elements[0]
    if fmt == "ascii":
        rows = body.decode("ascii").split("\n")[: v["count"]]
        arr = np.array([[float(x) for x in r.split()] for r in rows], dtype=np.float64)
        out = np.empty(v["count"], dtype=[(n, t) for n, t in v["props"]])
        for k, (n, _) in enumerate(v["props"]):
            out[n] = arr[:, k]
        return out
    order = {"binary_little_endian": "<", "binary_big_endian": ">"}.get(fmt)
    if order is None:
        raise ValueError(f"{path}: unknown PLY format {fmt}")
    dt = np.dtype([(n, order + t) for n, t in v["props"]])
    return np.frombuffer(body, dtype=dt, count=v["count"])


def load_ply(path: str, max_sh_degree: int = 3, num_gauss: int = -1) -> Dict[str, np.ndarray]:
    """The raw parameters of a Gaussian PLY file as load_ply reads them:
    xyz [P,3], features_dc [P,1,3], features_rest [P,K-1,3], opacity [P,1],
    scaling [P,3], rotation [P,4] (float32), the first ``num_gauss`` rows
    (all when -1)."""
    el = read_vertices(path)
    names = el.dtype.names
    P = el.shape[0]
    xyz = np.stack([np.asarray(el["x"]), np.asarray(el["y"]), np.asarray(el["z"])], axis=1)
    opacity = np.asarray(el["opacity"])[..., np.newaxis]
    f_dc = np.zeros((P, 3, 1))
    for c in range(3):
        f_dc[:, c, 0] = np.asarray(el[f"f_dc_{c}"])
    extra = sorted((n for n in names if n.startswith("f_rest_")), key=lambda x: int(x.split("_")[-1]))
    if len(extra) != 3 * (max_sh_degree + 1) ** 2 - 3:
        raise ValueError(f"{path}: {len(extra)} f_rest properties, expected {3 * (max_sh_degree + 1) ** 2 - 3}")
    f_rest = np.zeros((P, len(extra)))
    for k, n in enumerate(extra):
        f_rest[:, k] = np.asarray(el[n])
    f_rest = f_rest.reshape(P, 3, (max_sh_degree + 1) ** 2 - 1)

    def by_index(prefix):
        cols = sorted((n for n in names if n.startswith(prefix)), key=lambda x: int(x.split("_")[-1]))
        return np.stack([np.asarray(el[n]) for n in cols], axis=1)

    scaling, rotation = by_index("scale_"), by_index("rot")
    n = P if num_gauss == -1 else num_gauss
    f32 = lambda a: np.ascontiguousarray(a[:n], dtype=np.float32)  # noqa: E731
    return dict(xyz=f32(xyz), features_dc=f32(f_dc.transpose(0, 2, 1)), features_rest=f32(f_rest.transpose(0, 2, 1)),
                opacity=f32(opacity), scaling=f32(scaling), rotation=f32(rotation))
