"""SFEM raw mesh folders: meta.yaml plus one binary file per coordinate and per element node.

Layout (python/sfem/mesh/box_mesh.py:770-900, the reference's writer):
  meta.yaml        element_type (TET4, HEX8, ...), spatial_dimension, rpath, n_points,
                   points: [{x: x.raw}, {y: y.raw}, {z: z.raw}], elem_num_nodes, n_elements,
                   elements: [{i0: i0.raw}, ...], optional boundary_nodes: [{name: file}]
                   (files under boundary_nodes/) and boundary_faces: [{name: surface/name}]
  x.raw, y.raw, z.raw   float32 coordinates [n_points]
  i<k>.raw              int32 node ids of element node k [n_elements] (the SoA connectivity)
  boundary_nodes/<name>.int32.raw   int32 node sets
Only safe YAML loading is used.
"""
import os

import numpy as np
import yaml

GEOM_T = np.float32
IDX_T = np.int32


def _entries(lst):
    """[{key: value}, ...] -> [(key, value), ...] in file order."""
    out = []
    for item in lst or []:
        for k, v in item.items():
            out.append((k, v))
    return out


def read_meta(folder):
    with open(os.path.join(folder, "meta.yaml")) as f:
        return yaml.safe_load(f)


def read_mesh(folder, with_boundary=True):
    """Returns dict(element_type, elements [nv][E] int32, points [dim][N] float32,
    boundary_nodes {name: int32 array}, surfaces {name: (element_type, elements [nv][F])})."""
    meta = read_meta(folder)
    n_points = int(meta["n_points"])
    n_elements = int(meta["n_elements"])
    pts = []
    for _, fname in _entries(meta["points"]):
        a = np.fromfile(os.path.join(folder, fname), dtype=GEOM_T)
        if a.size != n_points:
            raise ValueError(f"{fname}: {a.size} values, meta.yaml says n_points = {n_points}")
        pts.append(a)
    els = []
    for _, fname in _entries(meta["elements"]):
        a = np.fromfile(os.path.join(folder, fname), dtype=IDX_T)
        if a.size != n_elements:
            raise ValueError(f"{fname}: {a.size} values, meta.yaml says n_elements = {n_elements}")
        els.append(a)
    if len(els) != int(meta["elem_num_nodes"]):
        raise ValueError("elements list length differs from elem_num_nodes")
    out = {"element_type": meta["element_type"], "elements": np.stack(els), "points": np.stack(pts),
           "boundary_nodes": {}, "surfaces": {}, "meta": meta}
    if with_boundary:
        for name, fname in _entries(meta.get("boundary_nodes")):
            out["boundary_nodes"][name] = np.fromfile(os.path.join(folder, "boundary_nodes", fname), dtype=IDX_T)
        for name, sub in _entries(meta.get("boundary_faces")):
            sm = read_meta(os.path.join(folder, sub))
            faces = [np.fromfile(os.path.join(folder, sub, f), dtype=IDX_T) for _, f in _entries(sm["elements"])]
            out["surfaces"][name] = (sm["element_type"], np.stack(faces))
    return out


def write_mesh(folder, element_type, elements, points, boundary_nodes=None):
    """Writes a mesh folder in the same layout (meta.yaml keys, file names, dtypes)."""
    elements = np.asarray(elements)
    points = np.asarray(points)
    os.makedirs(folder, exist_ok=True)
    prefix = ["x", "y", "z"]
    meta = {"element_type": element_type, "spatial_dimension": int(points.shape[0]), "rpath": True,
            "n_points": int(points.shape[1]), "points": [], "elem_num_nodes": int(elements.shape[0]),
            "n_elements": int(elements.shape[1]), "elements": []}
    for d in range(points.shape[0]):
        meta["points"].append({prefix[d]: f"{prefix[d]}.raw"})
        np.ascontiguousarray(points[d], dtype=GEOM_T).tofile(os.path.join(folder, f"{prefix[d]}.raw"))
    for k in range(elements.shape[0]):
        meta["elements"].append({f"i{k}": f"i{k}.raw"})
        np.ascontiguousarray(elements[k], dtype=IDX_T).tofile(os.path.join(folder, f"i{k}.raw"))
    if boundary_nodes:
        os.makedirs(os.path.join(folder, "boundary_nodes"), exist_ok=True)
        meta["boundary_nodes"] = []
        for name, nodes in boundary_nodes.items():
            fname = f"{name}.int32.raw"
            meta["boundary_nodes"].append({name: fname})
            np.ascontiguousarray(nodes, dtype=IDX_T).tofile(os.path.join(folder, "boundary_nodes", fname))
    with open(os.path.join(folder, "meta.yaml"), "w") as f:
        f.write("# SFEM mesh meta file\n")
        yaml.safe_dump(meta, f)
    return meta
