"""Writes tests/golden/reference_struct_fields.json: the fields of the reference
types a Julia binding of the library reads, taken from the reference's source
text (/root/reference/src/*.jl, read as text; nothing is run or imported).

* "structs": every `struct` of src/*.jl with its fields in order (name, type)
  and the file:line of its definition.
* "used_fields": fields of third-party types (NeutrinoTelescopeBase shapes and
  targets) that the reference's own code reads, with the file:line of a use;
  their definitions are not in the reference tree.
* "constants": the Cascadia constants of src/water_properties.jl.
* "methods": the functions of the reference's medium accessors and spectrum
  helpers a binding may call, with the file:line of their definition.

Run from the repository root: python tests/golden/gen_reference_struct_fields.py
"""
import json
import os
import re

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_struct_fields.json")


def structs(path, rel):
    out = {}
    lines = open(path).read().split("\n")
    i = 0
    while i < len(lines):
        m = re.match(r"^(?:(?:Base\.)?@kwdef\s+)?(?:mutable\s+)?struct\s+(\w+)", lines[i])
        if m and not lines[i].startswith(" "):
            name, fields, j = m.group(1), [], i + 1
            while j < len(lines) and lines[j].strip() != "end":
                t = lines[j].split("#")[0].strip()
                fm = re.match(r"^(\w+)::(.+)$", t)
                if fm:
                    fields.append([fm.group(1), fm.group(2).strip()])
                elif re.match(r"^\w+$", t):
                    fields.append([t, "Any"])
                elif t.startswith("function") or re.match(r"^\w+\(", t):
                    break  # inner constructor: the fields are above it
                j += 1
            out[name] = {"fields": fields, "at": f"{rel}:{i + 1}"}
            i = j
        i += 1
    return out


def uses(pattern, files):
    found = {}
    for rel in files:
        for k, line in enumerate(open(os.path.join(REF, rel)).read().split("\n"), 1):
            for m in re.finditer(pattern, line):
                found.setdefault(m.group(1), f"{rel}:{k}")
    return found


def shape_uses(rel):
    """target_shape.<field> inside each check_intersection(target_shape::<Shape>, ...)."""
    out, cur = {}, None
    for k, line in enumerate(open(os.path.join(REF, rel)).read().split("\n"), 1):
        m = re.match(r"^function\s+(\w+)\(", line)
        if m:
            sm = re.match(r"^function\s+check_intersection\(target_shape::(\w+)", line)
            cur = sm.group(1) if sm else None
        if cur:
            for fm in re.finditer(r"target_shape\.(\w+)", line):
                out.setdefault(cur, {}).setdefault(fm.group(1), f"{rel}:{k}")
    return out


def main():
    src = sorted(f for f in os.listdir(os.path.join(REF, "src")) if f.endswith(".jl"))
    all_structs = {}
    for f in src:
        all_structs.update(structs(os.path.join(REF, "src", f), f"src/{f}"))
    shapes = shape_uses("src/photon_prop_cuda.jl")
    target = uses(r"targ\.(\w+)", ["src/processing.jl"])
    consts = {}
    for k, line in enumerate(open(os.path.join(REF, "src/water_properties.jl")), 1):
        m = re.match(r"^const (CASCADIA_\w+)\s*=\s*([^#\n]+)", line)
        if m:
            consts[m.group(1)] = {"value": m.group(2).strip(), "at": f"src/water_properties.jl:{k}"}
    methods = {}
    for f in src:
        for k, line in enumerate(open(os.path.join(REF, "src", f)), 1):
            m = re.match(r"^(?:function\s+)?((?:CherenkovMediumBase\.)?\w+)\((::|medium::|spec::)"
                         r"(\w+)", line)
            if m and m.group(3) in ("CascadiaMediumProperties", "SpectralDist",
                                    "HomogenousIceProperties", "SimpleMediumProperties",
                                    "WaterProperties"):
                methods.setdefault(f"{m.group(1)}({m.group(3)})", f"src/{f}:{k}")
    json.dump({"structs": all_structs,
               "used_fields": {**shapes, "PhotonTarget": target},
               "constants": consts, "methods": methods},
              open(OUT, "w"), indent=1, sort_keys=True)
    print(f"wrote {OUT}: {len(all_structs)} structs, {len(consts)} constants, "
          f"{len(methods)} methods")


if __name__ == "__main__":
    main()
