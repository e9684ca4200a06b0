"""Gaussian PLY files (scene/gaussian_model_cf.py:398-432 save_ply,
459-515 load_ply) on the CPU.

The reference writes them through plyfile, which is not installed here, so
the format is pinned by the PLY specification and the reference's attribute
list: the header text, then each vertex as consecutive little-endian float32
values in the order of construct_list_of_attributes, with the SH
coefficients channel-major (features.transpose(1, 2).flatten(1)).  The
expected bytes below are assembled with struct.pack from that description,
independently of the writer.
"""
import struct

import numpy as np
import pytest

from gs_hip import ply


def _model(P=7, K=16, seed=0):
    rng = np.random.default_rng(seed)
    f = lambda *s: rng.standard_normal(s).astype(np.float32)  # noqa: E731
    return dict(xyz=f(P, 3), features_dc=f(P, 1, 3), features_rest=f(P, K - 1, 3), opacity=f(P, 1),
                scaling=f(P, 3), rotation=f(P, 4))


def test_attribute_list_is_the_references():
    names = ply.attribute_names(15)
    assert names[:9] == ["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"]
    assert names[9:54] == [f"f_rest_{i}" for i in range(45)]
    assert names[54:] == ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    assert len(names) == 62


def test_bytes_follow_the_format(tmp_path):
    m = _model()
    path = str(tmp_path / "a" / "g.ply")  # the directory is created, as mkdir_p does
    ply.save_ply(path, **m)
    data = open(path, "rb").read()
    names = ply.attribute_names(15)
    header = ("ply\nformat binary_little_endian 1.0\nelement vertex 7\n"
              + "".join("property float %s\n" % n for n in names) + "end_header\n").encode()
    assert data[:len(header)] == header
    body = b""
    for i in range(7):
        vals = list(m["xyz"][i]) + [0.0, 0.0, 0.0] + list(m["features_dc"][i, 0])
        for c in range(3):  # channel-major: all red coefficients, then green, then blue
            vals += [m["features_rest"][i, k, c] for k in range(15)]
        vals += list(m["opacity"][i]) + list(m["scaling"][i]) + list(m["rotation"][i])
        body += struct.pack("<62f", *[float(v) for v in vals])
    assert data[len(header):] == body


def test_round_trip_and_truncation(tmp_path):
    m = _model(P=33, seed=1)
    path = str(tmp_path / "g.ply")
    ply.save_ply(path, **m)
    r = ply.load_ply(path)
    for k, v in m.items():
        assert r[k].dtype == np.float32 and np.array_equal(r[k], v), k
    r5 = ply.load_ply(path, num_gauss=5)
    assert all(np.array_equal(r5[k], v[:5]) for k, v in m.items())
    with pytest.raises(ValueError):
        ply.load_ply(path, max_sh_degree=2)  # 45 f_rest properties, 24 expected


def test_degree_one_model_and_other_encodings(tmp_path):
    m = _model(P=4, K=4, seed=2)
    names = ply.attribute_names(3)
    rows = []
    for i in range(4):
        vals = list(m["xyz"][i]) + [0.0] * 3 + list(m["features_dc"][i, 0])
        for c in range(3):
            vals += [m["features_rest"][i, k, c] for k in range(3)]
        vals += list(m["opacity"][i]) + list(m["scaling"][i]) + list(m["rotation"][i])
        rows.append(vals)
    # ASCII, with a comment line, and big-endian binary, both read back
    head = "ply\nformat {}\ncomment made by a test\nelement vertex 4\n" + \
        "".join("property float %s\n" % n for n in names) + "end_header\n"
    asc = head.format("ascii 1.0") + "".join(" ".join(repr(float(v)) for v in r) + "\n" for r in rows)
    (tmp_path / "a.ply").write_text(asc)
    big = head.format("binary_big_endian 1.0").encode() + b"".join(struct.pack(">%df" % len(r), *r) for r in rows)
    (tmp_path / "b.ply").write_bytes(big)
    for name in ("a.ply", "b.ply"):
        r = ply.load_ply(str(tmp_path / name), max_sh_degree=1)
        for k, v in m.items():
            assert np.array_equal(r[k], v), (name, k)
