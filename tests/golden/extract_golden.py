"""Transcribe the reference's Mathematica known-answer arrays into JSON fixtures.

Reads the numeric arrays (and only those) out of the reference unit tests
  unit_test/tstAPIC.hpp    checkGridVelocity/checkParticleVelocity/
                           checkGridMomentum/checkGridMass, orders 1, 2, 3
  unit_test/tstPolyPIC.hpp the same four checks for order 1
and writes them, with the test set-up constants, to ``apic_golden.json`` and
``polypic_golden.json`` next to this script.  The arrays are data produced by
the Picasso authors' notebooks (doc/apic/*.nb, doc/polypic/*.nb; BSD-3-Clause,
Copyright (c) 2021 by the Picasso authors).

Run once in a container that has the reference sources:
    python tests/golden/extract_golden.py /root/reference
The fixtures are committed; nothing at test time reads the reference.
"""
import json
import os
import re
import sys

FUNC_RE = re.compile(
    r"void\s+(check\w+|set\w+)\s*\(\s*std::integral_constant<int,\s*(\d)>")
ARR_RE = re.compile(r"double\s+(\w+)((?:\[\d+\])+)\s*=\s*(\{.*?\});", re.S)
NUM_RE = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def parse_file(path):
    """Return {(function, order): {array_name: nested list}} for a test file."""
    text = open(path).read()
    starts = [(m.start(), m.group(1), int(m.group(2))) for m in FUNC_RE.finditer(text)]
    out = {}
    for idx, (pos, name, order) in enumerate(starts):
        end = starts[idx + 1][0] if idx + 1 < len(starts) else len(text)
        body = text[pos:end]
        arrays = {}
        for am in ARR_RE.finditer(body):
            aname, dims, init = am.group(1), am.group(2), am.group(3)
            shape = [int(d) for d in re.findall(r"\[(\d+)\]", dims)]
            vals = [float(v) for v in NUM_RE.findall(init)]
            n = 1
            for s in shape:
                n *= s
            assert len(vals) == n, (path, name, order, aname, len(vals), shape)
            arrays[aname] = reshape(vals, shape)
        out[(name, order)] = arrays
    return out


def reshape(vals, shape):
    if len(shape) == 1:
        return vals
    step = len(vals) // shape[0]
    return [reshape(vals[i * step:(i + 1) * step], shape[1:]) for i in range(shape[0])]


# Set-up constants of both drivers (tstAPIC.hpp:540-567,581-589,698-716;
# tstPolyPIC.hpp:196-226,351-369; unit_test/inputs/polypic_test.json).
SETUP = {
    "near_eps": 1.0e-11,
    "global_box": [-50.0, -50.0, -50.0, 50.0, 50.0, 50.0],
    "cell_size": 0.5,
    "halo_cell_width": 2,
    "particle_mass": 0.134,
    "particle_x": [9.31, -8.28, -3.34],
    "particle_cell_local": [120, 85, 95],
    "polypic_dt": 0.0001,
    "grid_function": "F(i,j,k,d) = 1e-10*(d+1)*ic^5 - 1.2e-9*((d+1)+jc*ic)^3"
                     " + 1e-10*(ic*jc*kc)^2 + (d+1)^2, ic=i-2, jc=j-2, kc=k-2",
}


def main(ref_root):
    here = os.path.dirname(os.path.abspath(__file__))
    apic = parse_file(os.path.join(ref_root, "unit_test", "tstAPIC.hpp"))
    poly = parse_file(os.path.join(ref_root, "unit_test", "tstPolyPIC.hpp"))

    apic_out = {"setup": SETUP, "source": "unit_test/tstAPIC.hpp", "orders": {}}
    for order in (1, 2, 3):
        apic_out["orders"][str(order)] = {
            "grid_velocity": apic[("checkGridVelocity", order)]["mathematica"],
            "particle_u": apic[("checkParticleVelocity", order)]["mathematica_u"],
            "particle_b": apic[("checkParticleVelocity", order)]["mathematica_b"],
            "grid_momentum": apic[("checkGridMomentum", order)]["mathematica"],
            "grid_mass": apic[("checkGridMass", order)]["mathematica"],
        }
    poly_out = {"setup": SETUP, "source": "unit_test/tstPolyPIC.hpp", "orders": {
        "1": {
            "grid_velocity": poly[("checkGridVelocity", 1)]["mathematica"],
            "particle_modes": poly[("checkParticleVelocity", 1)]["mathematica"],
            "set_particle_modes": poly[("setParticleVelocity", 1)]["mathematica"],
            "grid_momentum": poly[("checkGridMomentum", 1)]["mathematica"],
            "grid_mass": poly[("checkGridMass", 1)]["mathematica"],
        }}}
    with open(os.path.join(here, "apic_golden.json"), "w") as f:
        json.dump(apic_out, f, indent=1)
    with open(os.path.join(here, "polypic_golden.json"), "w") as f:
        json.dump(poly_out, f, indent=1)
    print("wrote apic_golden.json, polypic_golden.json")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
