"""Model input from the reference's ``print -json`` output.

``load_json(model, doc)`` defines, on a ``Model`` (GPU) or an oracle model,
everything of a reference model that lies on the GPU path, from the JSON
document the reference writes with ``print -json`` (printDomain,
SRC/runtime/commands/modeling/printing.cpp:135-295):

* geometry.nodes: ``{"name", "ndf", "crd", "mass"}`` (Node.cpp:1422-1437);
* geometry.constraints: single-point constraints ``{"name", "node", "dof",
  "ref_value"}`` (SP_Constraint.cpp:251-256) become ``fix`` flags;
* properties.nDMaterials: J2Plasticity ``{"G", "K", "Fy", "Fs", "Hsat",
  "Hiso", "eta", "density"}`` (J2Plasticity.cpp:178-191);
* properties.uniaxialMaterials: Steel02 (Steel02.cpp:382-397), Concrete02
  (Concrete02.cpp:404-415), ElasticMaterial (ElasticMaterial.cpp:284-291);
* properties.sections: FiberSection2d ``{"fibers": [{"coord": [y, 0.0],
  "area", "material"}]}`` (FiberSection2d.cpp:631-645) and FiberSection3d
  with ``"torsion"`` (FiberSection3d.cpp:951-966);
* properties.crdTransformations: LinearCrdTransf2d / LinearCrdTransf3d
  ``{"vecInLocXZPlane"}`` (LinearCrdTransf2d.cpp:1180-1190,
  LinearCrdTransf3d.cpp:1536-1548);
* geometry.elements: Brick (Brick.cpp:224-237), FourNodeQuad
  (FourNodeQuad.cpp:776-797, plane strain), DispBeamColumn2d/3d
  (DispBeamColumn2d.cpp:1064-1077, DispBeamColumn3d.cpp:1149-1162),
  ForceBeamColumn2d/3d with Lobatto integration (ForceBeamColumn2d.cpp
  Print, the same fields).

Anything else raises ``XaraError`` naming the object, rather than being
skipped.  The reference's JSON carries no load values (NodalLoad.cpp:210-215
prints an empty list), so loads, patterns and the analysis are defined by
the caller as usual.
"""
import json
import math

import numpy as np

from .model import XaraError


def _num(x):
    return float(x)


def _tag(x):
    return int(x)  # names are printed as numbers or as quoted numbers


def _parse(doc):
    if isinstance(doc, dict):
        return doc
    text = doc.read() if hasattr(doc, "read") else doc
    consts = {"inf": math.inf, "-inf": -math.inf, "nan": math.nan, "-nan": math.nan}
    return json.loads(text, parse_constant=lambda c: consts.get(c.lower(), math.nan))


def _err(msg):
    return XaraError(-1, f"print -json model: {msg}")


def load_json(model, doc):
    """Define the model described by a reference ``print -json`` document
    (a path-free string, file object or parsed dict) on `model`."""
    d = _parse(doc)
    if "StructuralAnalysisModel" in d:
        d = d["StructuralAnalysisModel"]
    props, geom = d.get("properties", {}), d.get("geometry", {})
    ndm, ndf = model.ndm, model.ndf

    # ---- nodes ----
    nodes = geom.get("nodes", [])
    if not nodes:
        raise _err("no nodes")
    tags = np.array([_tag(n["name"]) for n in nodes], np.int32)
    crds = np.zeros((len(nodes), ndm))
    mass = np.zeros((len(nodes), ndf))
    has_mass = False
    for k, n in enumerate(nodes):
        if int(n.get("ndf", ndf)) != ndf:
            raise _err(f"node {n['name']} has ndf {n.get('ndf')}, the model {ndf}")
        c = n["crd"]
        if len(c) != ndm:
            raise _err(f"node {n['name']} has {len(c)} coordinates, the model ndm {ndm}")
        crds[k] = c
        if "mass" in n:
            mass[k] = [_num(v) for v in n["mass"]]
            has_mass = True
    model.nodes(tags, crds)
    if has_mass:
        sel = np.any(mass != 0.0, axis=1)
        if sel.any():
            model.masses(tags[sel], mass[sel])

    # ---- single-point constraints ----
    fix = {}
    for c in geom.get("constraints", []):
        if "node" not in c or "dof" not in c:
            raise _err(f"constraint {c.get('name')} is not a single-point constraint")
        if _num(c.get("ref_value", 0.0)) != 0.0:
            raise _err(f"constraint {c.get('name')}: imposed nonzero displacement")
        flags = fix.setdefault(_tag(c["node"]), [0] * ndf)
        dof = int(c["dof"])
        if not 1 <= dof <= ndf:
            raise _err(f"constraint {c.get('name')}: dof {dof}")
        flags[dof - 1] = 1
    if fix:
        ft = np.array(sorted(fix), np.int32)
        model.fix_many(ft, np.array([fix[t] for t in ft], np.int32))

    # ---- materials ----
    # (older printers wrote the key as "ndMaterials")
    for m in props.get("nDMaterials", props.get("ndMaterials", [])):
        t = m.get("type")
        if t != "J2Plasticity":
            raise _err(f"nDMaterial {m.get('name')} of type {t} is not on the GPU path")
        if _num(m.get("density", 0.0)) != 0.0:
            raise _err(f"nDMaterial {m.get('name')}: density is not supported")
        model.nDMaterial("J2Plasticity", _tag(m["name"]), _num(m["K"]), _num(m["G"]), _num(m["Fy"]),
                         _num(m["Fs"]), _num(m["Hsat"]), _num(m["Hiso"]), _num(m.get("eta", 0.0)))
    for m in props.get("uniaxialMaterials", []):
        t, tag = m.get("type"), _tag(m["name"])
        if t == "Steel02":
            model.uniaxialMaterial("Steel02", tag, _num(m["fy"]), _num(m["E"]), _num(m["b"]), _num(m["R0"]),
                                   _num(m["cR1"]), _num(m["cR2"]), _num(m["a1"]), _num(m["a2"]), _num(m["a3"]),
                                   _num(m["a4"]), _num(m.get("sigini", 0.0)))
        elif t == "Concrete02":
            model.uniaxialMaterial("Concrete02", tag, _num(m["fc"]), _num(m["epsc"]), _num(m["fcu"]),
                                   _num(m["epscu"]), _num(m["ratio"]), _num(m["ft"]), _num(m["Ets"]))
        elif t in ("ElasticMaterial", "Elastic"):
            if _num(m.get("Eneg", m["Epos"])) != _num(m["Epos"]) or _num(m.get("eta", 0.0)) != 0.0:
                raise _err(f"uniaxialMaterial {tag}: only a symmetric, rate-independent Elastic is supported")
            model.uniaxialMaterial("Elastic", tag, _num(m["Epos"]))
        else:
            raise _err(f"uniaxialMaterial {tag} of type {t} is not on the GPU path")

    # ---- sections ----
    for s in props.get("sections", []):
        t, tag = s.get("type"), _tag(s["name"])
        if t not in ("FiberSection2d", "FiberSection3d"):
            raise _err(f"section {tag} of type {t} is not on the GPU path")
        rows = [(_num(f["coord"][0]), _num(f["coord"][1]) if t == "FiberSection3d" else 0.0, _num(f["area"]),
                 _tag(f["material"])) for f in s["fibers"]]
        gj = _num(s["torsion"]) if (t == "FiberSection3d" and "torsion" in s) else None
        model.section("Fiber", tag, rows, GJ=gj)

    # ---- coordinate transformations ----
    for c in props.get("crdTransformations", []):
        t, tag = c.get("type"), _tag(c["name"])
        if "iOffset" in c or "jOffset" in c:
            raise _err(f"geomTransf {tag}: joint offsets are not supported")
        if t == "LinearCrdTransf2d":
            model.geomTransf("Linear", tag)
        elif t == "LinearCrdTransf3d":
            model.geomTransf("Linear", tag, *[_num(v) for v in c["vecInLocXZPlane"]])
        else:
            raise _err(f"geomTransf {tag} of type {t} is not on the GPU path")

    # ---- elements, grouped by type and parameters for the bulk calls ----
    groups = {}
    for e in geom.get("elements", []):
        t, tag = e.get("type"), _tag(e["name"])
        nds = tuple(_tag(n) for n in e["nodes"])
        if t == "Brick":
            mats = {_tag(x) for x in e["material"]}
            if len(mats) != 1:
                raise _err(f"element {tag}: one material for all Gauss points is required")
            b = tuple(_num(x) for x in e.get("bodyForces", [0, 0, 0]))
            key = ("stdBrick", mats.pop()) + b
        elif t == "FourNodeQuad":
            mats = {_tag(x) for x in e["materials"]}
            if len(mats) != 1:
                raise _err(f"element {tag}: one material for all Gauss points is required")
            b = tuple(_num(x) for x in e.get("bodyForces", [0, 0]))
            key = ("quad", _num(e["thickness"]), mats.pop(), _num(e.get("surfacePressure", 0.0)),
                   _num(e.get("density", 0.0))) + b
        elif t in ("DispBeamColumn2d", "DispBeamColumn3d"):
            secs = [_tag(x) for x in e["sections"]]
            if len(set(secs)) != 1:
                raise _err(f"element {tag}: one section at every integration point is required")
            integ = e.get("integration", {}).get("type", "Legendre")
            if integ != "Legendre":
                raise _err(f"element {tag}: {integ} integration is not on the GPU path")
            key = ("dispBeamColumn", _tag(e["crdTransformation"]), secs[0], len(secs),
                   _num(e.get("massperlength", 0.0)))
        elif t in ("ForceBeamColumn2d", "ForceBeamColumn3d"):
            secs = [_tag(x) for x in e["sections"]]
            if len(set(secs)) != 1:
                raise _err(f"element {tag}: one section at every integration point is required")
            integ = e.get("integration", {}).get("type", "Lobatto")
            if integ != "Lobatto":
                raise _err(f"element {tag}: {integ} integration is not on the GPU path")
            if _num(e.get("massperlength", 0.0)) != 0.0:
                raise _err(f"element {tag}: forceBeamColumn mass is not supported")
            # maxIters and tol are not printed: the command's defaults (10, 1e-12)
            key = ("forceBeamColumn", _tag(e["crdTransformation"]), secs[0], len(secs))
        else:
            raise _err(f"element {tag} of type {t} is not on the GPU path")
        groups.setdefault(key, ([], []))
        groups[key][0].append(tag)
        groups[key][1].append(nds)
    for key, (etags, conn) in groups.items():
        kind, args = key[0], key[1:]
        model.elements(kind, etags, conn, *args)
    return model
