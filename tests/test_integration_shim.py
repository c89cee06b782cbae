"""The Julia binding shown in INTEGRATION.md reads only what the reference's
types hold. Every `x.field` (and `x.f1.f2`) read of a typed argument `x::T` in
the shim is checked against the struct definitions of the reference's sources
(tests/golden/reference_struct_fields.json, written from the reference's text
by tests/golden/gen_reference_struct_fields.py), against the fields the
reference itself reads from third-party shape and target types, and against
the shim's own POD structs. Module constants and medium accessors the shim
calls must exist in the reference. The POD structs must match the C header's
sizes and field order."""
import json
import os
import re

import pytest

from conftest import GOLDEN, ROOT

FIX = json.load(open(os.path.join(GOLDEN, "reference_struct_fields.json")))


def shim_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```julia\n(.*?)```", text, re.S)
    assert len(blocks) == 1
    return blocks[0]


def shim_structs(src):
    out = {}
    for m in re.finditer(r"^(?:mutable )?struct (\w+)(.*?)^end", src, re.S | re.M):
        fields = re.findall(r"(\w+)::", m.group(2))
        out[m.group(1)] = fields
    return out


def base(t):
    return re.sub(r"\{.*\}$", "", t.strip())


def fields_of(typ, own):
    typ = base(typ)
    if typ in own:
        return {f: None for f in own[typ]}
    if typ in FIX["structs"]:
        return {f: t for f, t in FIX["structs"][typ]["fields"]}
    if typ in FIX["used_fields"]:
        return {f: None for f in FIX["used_fields"][typ]}
    return None


def methods(src):
    """(name, {arg: type}, body) of every definition with typed arguments."""
    lines = src.split("\n")
    heads = [i for i, l in enumerate(lines)
             if re.match(r"^(function\s+)?[\w.]+\(.*::", l) and not l.startswith(" ")]
    out = []
    for k, i in enumerate(heads):
        end = heads[k + 1] if k + 1 < len(heads) else len(lines)
        if lines[i].startswith("function"):
            end = next(j for j in range(i + 1, len(lines)) if lines[j] == "end") + 1
        head = lines[i]
        j = i
        while head.count("(") > head.count(")"):
            j += 1
            head += lines[j]
        args = dict((a, t) for a, t in re.findall(r"[(,]\s*(\w+)::([\w.{}<: ]+?)[,)]", head))
        out.append((lines[i].split("(")[0], args, "\n".join(lines[i:end])))
    return out


def test_every_field_read_exists_in_the_reference_types():
    src = shim_source()
    own = shim_structs(src)
    checked = 0
    for name, args, body in methods(src):
        for arg, typ in args.items():
            if fields_of(typ, own) is None:
                continue  # an untyped or abstract argument (Integer, Type{T} ...)
            for m in re.finditer(r"\b%s((?:\.\w+)+)" % re.escape(arg), body):
                cur = typ
                for f in m.group(1).strip(".").split("."):
                    fs = fields_of(cur, own)
                    assert fs is not None, f"{name}: {arg}.{m.group(1)}: type {cur} unknown"
                    assert f in fs, f"{name}: {arg}::{typ} reads .{f}, not a field of {cur}"
                    cur = fs[f] or ""
                    checked += 1
                    if fields_of(cur, own) is None:
                        break
    assert checked >= 40, checked


def test_shim_lowers_every_reference_emitter_and_medium():
    src = shim_source()
    lowered = set(re.findall(r"^(?:function\s+)?to_pod\(\w+::(\w+)\)", src, re.M))
    emitters = {n for n in FIX["structs"] if re.search(r"Emitter$", n)}
    assert emitters <= lowered, emitters - lowered
    assert {"CascadiaMediumProperties", "SimpleMediumProperties",
            "HomogenousIceProperties"} <= lowered
    for shape in ("Spherical", "Rectangular", "Circular"):
        assert re.search(r"^shape_pod\(s::%s," % shape, src, re.M)
    assert re.search(r"^target_pods\(d::DetectorLines\)", src, re.M)


def test_constants_and_accessors_exist_in_the_reference():
    src = shim_source()
    for c in re.findall(r"PM\.(\w+)", src):
        assert c in FIX["constants"], c
    meth = FIX["methods"]
    for acc in re.findall(r"CherenkovMediumBase\.(\w+)\(m\b", src):
        defined = any(k.startswith(f"CherenkovMediumBase.{acc}(") for k in meth)
        called = acc in ("phase_refractive_index", "scattering_length")  # called by the reference
        assert defined or called, acc
    assert "make_cuda_spectral_dist(SpectralDist)" in meth


def test_pod_layouts_match_the_header():
    src = shim_source()
    own = shim_structs(src)
    hdr = open(os.path.join(ROOT, "include", "photonprop_hip.h")).read()
    for jl, c in (("PPSource", "pp_source"), ("PPMedium", "pp_medium"),
                  ("PPTarget", "pp_target"), ("PPConfig", "pp_config")):
        body = re.search(r"typedef struct %s \{(.*?)\} %s;" % (c, c), hdr, re.S).group(1)
        cf = re.findall(r"^\s*\w+\s+(\w+)(?:\[\d+\])?;", body, re.M)
        assert own[jl] == cf, (jl, own[jl], cf)


@pytest.mark.parametrize("kind", ["lines", "shapes"])
def test_python_mirror_flattens_detector_lines(P, kind):
    import numpy as np
    lines = [P.DetectorLine(x, 0.0, 1.0, [P.Spherical(np.float32([x, 0, z]), 0.3)
                                          for z in (0.0, 10.0, 20.0)]) for x in (0.0, 50.0)]
    t = P.DetectorLines(lines) if kind == "lines" else [s for ln in lines for s in ln.targets]
    pods = P.targets_to_pod(t)
    assert [(p.position[0], p.position[2], p.module_id) for p in pods] == [
        (0.0, 0.0, 1), (0.0, 10.0, 2), (0.0, 20.0, 3), (50.0, 0.0, 4), (50.0, 10.0, 5),
        (50.0, 20.0, 6)]
