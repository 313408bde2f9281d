"""Minimal PLY writer for the reader tests (numpy structured arrays; ascii or
binary of either byte order) and the 3DGS attribute layout of
gaussian_model.py:173-218."""
import numpy as np

PLY_TYPES = {"f4": "float", "f8": "double", "i1": "char", "u1": "uchar", "i2": "short", "u2": "ushort",
             "i4": "int", "u4": "uint"}


def write_ply(path, vertex, fmt="binary_little_endian", pre_elements=()):
    """vertex: list of (name, numpy dtype str, column array).  pre_elements:
    (name, count, [(prop, kind, ...)], body_bytes_or_lines) written before the
    vertex element."""
    n = len(vertex[0][2]) if vertex else 0
    hdr = ["ply", f"format {fmt} 1.0", "comment gs-evt reader test"]
    for name, count, props, _ in pre_elements:
        hdr.append(f"element {name} {count}")
        hdr.extend(props)
    hdr.append(f"element vertex {n}")
    for name, dt, _ in vertex:
        hdr.append(f"property {PLY_TYPES[dt]} {name}")
    hdr.append("end_header")
    with open(path, "wb") as f:
        f.write(("\n".join(hdr) + "\n").encode())
        for _, _, _, body in pre_elements:
            f.write(body)
        if fmt == "ascii":
            cols = [np.asarray(c) for _, _, c in vertex]
            for i in range(n):
                f.write((" ".join(repr(float(c[i])) if c.dtype.kind == "f" else str(int(c[i])) for c in cols)
                         + "\n").encode())
        else:
            bo = "<" if fmt == "binary_little_endian" else ">"
            rec = np.empty(n, dtype=[(name, bo + dt) for name, dt, _ in vertex])
            for name, _, c in vertex:
                rec[name] = c
            f.write(rec.tobytes())


def gaussian_vertex(raw, dtype="f4"):
    """3DGS attributes of a RawScene in the save_ply order."""
    P = raw.xyz.shape[0]
    f_dc = raw.features_dc.transpose(0, 2, 1).reshape(P, -1)
    f_rest = raw.features_rest.transpose(0, 2, 1).reshape(P, -1)
    cols = [("x", raw.xyz[:, 0]), ("y", raw.xyz[:, 1]), ("z", raw.xyz[:, 2]),
            ("nx", np.zeros(P)), ("ny", np.zeros(P)), ("nz", np.zeros(P))]
    cols += [(f"f_dc_{i}", f_dc[:, i]) for i in range(f_dc.shape[1])]
    cols += [(f"f_rest_{i}", f_rest[:, i]) for i in range(f_rest.shape[1])]
    cols += [("opacity", raw.opacity[:, 0])]
    cols += [(f"scale_{i}", raw.scaling[:, i]) for i in range(raw.scaling.shape[1])]
    cols += [(f"rot_{i}", raw.rotation[:, i]) for i in range(raw.rotation.shape[1])]
    return [(name, dtype, np.asarray(c, np.float32)) for name, c in cols]
