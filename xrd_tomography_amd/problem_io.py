"""On-disk format of a forward/MLEM problem: a directory with ``problem.json``
and one raw little-endian array file per input.

problem.json::

    {"format": "xrd-tomography-problem", "version": 1, "name": "...",
     "geometry": {<every xrd_geometry field>},
     "discretization": {<every xrd_discretization field>},
     "lut": [q_min, q_max, a_min, a_max],
     "meta": {"photons_per_view": ..., "mlem_iterations": ..., "voxel_mm": ...,
              "compton_scale": ... (optional)},
     "arrays": {"<name>": {"file": "<name>.bin", "dtype": "<f4" | "<i4" | "<f8",
                           "shape": [...], "order": "first index fastest"}}}

Arrays are stored in the memory order xrd_forward takes them (first index
fastest, the MATLAB column-major layout of Forward_XRD.cu:80-119):

  spectrum      <f4 [NRG, NRG]     spectrum[e_src + NRG*e_det]
  ifine         <f4 [QF, MAT]      fine form factor, q fastest
  segmentation  <i4 [NX, NY]       material id, 0 = air
  a1, a2        <f4 [NX, NY]       attenuation basis maps, 1/voxel
  s2_table      <f4 [NQ, NA]       Compton LUT values 0..255
  x_true        <f8 [NWT, MAT]     form factor at the coarse q centres (optional)
  counts        <f4 [rows]         measured counts y, row v + V*(det + DET*e) (optional)
  background    <f4 [rows]         Compton/background estimate c (optional)

compton_scale is the factor that turns the projector's Compton output C into
photons (XS_grid_edges(6)/255 in main.m:68); photons_per_view = 0 means the
spectrum is already in absolute photons, so the counts are N + compton_scale * C.

Loading checks every array's size against the descriptors.
"""
from __future__ import annotations

import json
import os
from typing import Optional

import numpy as np

from .descriptors import Discretization, Geometry
from .synthetic import Problem

FORMAT = "xrd-tomography-problem"
VERSION = 1
_DTYPES = {"<f4": np.float32, "<i4": np.int32, "<f8": np.float64}


def _shapes(g: Geometry, d: Discretization) -> dict:
    rows = g.n_views * g.n_cols * g.n_rows * d.n_energy
    return {"spectrum": [d.n_energy, d.n_energy], "ifine": [d.n_qf, d.n_mat],
            "segmentation": [g.nx, g.ny], "a1": [g.nx, g.ny], "a2": [g.nx, g.ny],
            "s2_table": [d.lut_nq, d.lut_na], "x_true": [d.n_q, d.n_mat],
            "counts": [rows], "background": [rows]}


_KIND = {"spectrum": "<f4", "ifine": "<f4", "segmentation": "<i4", "a1": "<f4", "a2": "<f4",
         "s2_table": "<f4", "x_true": "<f8", "counts": "<f4", "background": "<f4"}


def save_problem(prob: Problem, path: str, counts: Optional[np.ndarray] = None,
                 background: Optional[np.ndarray] = None) -> None:
    """Write prob (and optional measured counts / background) under directory path."""
    os.makedirs(path, exist_ok=True)
    g, d = prob.geometry, prob.disc
    shapes = _shapes(g, d)
    arrays = {}
    data = {"spectrum": prob.spectrum, "ifine": prob.ifine, "segmentation": prob.segmentation,
            "a1": prob.a1, "a2": prob.a2, "s2_table": prob.s2_table, "x_true": prob.x_true,
            "counts": counts, "background": background}
    for name, a in data.items():
        if a is None:
            continue
        a = np.asarray(a)
        if a.size != int(np.prod(shapes[name])):
            raise ValueError(f"{name}: {a.size} elements, expected shape {shapes[name]}")
        flat = (a.reshape(-1, order="F") if a.ndim > 1 else a.reshape(-1)).astype(_KIND[name])
        fname = f"{name}.bin"
        flat.tofile(os.path.join(path, fname))
        arrays[name] = {"file": fname, "dtype": _KIND[name], "shape": shapes[name],
                        "order": "first index fastest"}
    doc = {"format": FORMAT, "version": VERSION, "name": prob.name,
           "geometry": g.as_dict(), "discretization": d.as_dict(), "lut": list(prob.lut),
           "meta": {"photons_per_view": prob.photons_per_view,
                    "mlem_iterations": prob.mlem_iterations, "voxel_mm": prob.voxel_mm},
           "arrays": arrays}
    if prob.compton_scale:
        doc["meta"]["compton_scale"] = float(prob.compton_scale)
    with open(os.path.join(path, "problem.json"), "w") as f:
        json.dump(doc, f, indent=1)


def load_problem(path: str):
    """(Problem, extras) from directory path; extras holds the optional 'counts'
    and 'background' arrays (1-D float32) when present."""
    with open(os.path.join(path, "problem.json")) as f:
        doc = json.load(f)
    if doc.get("format") != FORMAT or doc.get("version") != VERSION:
        raise ValueError(f"{path}: not a {FORMAT} v{VERSION} directory")
    g = Geometry(**doc["geometry"])
    d = Discretization(**doc["discretization"])
    shapes = _shapes(g, d)
    out = {}
    for name, spec in doc["arrays"].items():
        if name not in shapes:
            raise ValueError(f"{path}: unknown array {name!r}")
        if spec["dtype"] != _KIND[name] or list(spec["shape"]) != shapes[name]:
            raise ValueError(f"{path}: {name} is {spec['dtype']} {spec['shape']}, the "
                             f"descriptors need {_KIND[name]} {shapes[name]}")
        fn = os.path.join(path, spec["file"])
        n = int(np.prod(shapes[name]))
        if os.path.getsize(fn) != n * np.dtype(_DTYPES[spec["dtype"]]).itemsize:
            raise ValueError(f"{fn}: size does not match {shapes[name]} {spec['dtype']}")
        flat = np.fromfile(fn, dtype=_DTYPES[spec["dtype"]], count=n)
        out[name] = flat.reshape(shapes[name], order="F") if len(shapes[name]) > 1 else flat
    for req in ("spectrum", "ifine", "segmentation", "a1", "a2", "s2_table"):
        if req not in out:
            raise ValueError(f"{path}: required array {req!r} missing")
    meta = doc.get("meta", {})
    prob = Problem(name=doc.get("name", os.path.basename(path)), geometry=g, disc=d,
                   spectrum=out["spectrum"], ifine=out["ifine"],
                   segmentation=out["segmentation"], a1=out["a1"], a2=out["a2"],
                   s2_table=out["s2_table"], lut=tuple(doc["lut"]),
                   x_true=out.get("x_true", np.zeros(shapes["x_true"])),
                   photons_per_view=float(meta.get("photons_per_view", 0.0)),
                   mlem_iterations=int(meta.get("mlem_iterations", 0)),
                   voxel_mm=float(meta.get("voxel_mm", 1.0)),
                   compton_scale=float(meta.get("compton_scale", 0.0)))
    extras = {k: out[k] for k in ("counts", "background") if k in out}
    return prob, extras
