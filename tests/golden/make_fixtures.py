"""Write tests/golden/reference_outputs.json from the reference tree.

Run in a container that has /root/reference (it reads the shipped OM dumps and
runs the shipped pre-generated C++ built by oracle/Makefile into oracle/_ref).
The JSON holds only data: digests of the OM dumps and of the state the shipped
generated code produces on fixed inputs, plus a few sample values.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference"


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def dump_digest(path):
    txt = open(path).read().rstrip("\n")
    return {"sha256": sha(txt.encode()), "lines": txt.count("\n") + 1}


def legacy_life_run():
    from oracle.runner import reference_machine
    from paraiso_amd.generator.native import Language, Setup
    from paraiso_amd.generator.plan import translate
    from paraiso_amd.schemes.life import life_om_legacy
    plan = translate(Setup((128, 128), Language.CPlusPlus), life_om_legacy())
    m = reference_machine("life", plan.om.setup.static_values, plan.memory_size(), ["init", "proceed"])
    m.run("init")
    out = {}
    for g in (0, 50, 300):
        while m.get("generation") < g:
            m.run("proceed")
        cell = m.get("cell")
        out[str(g)] = {"cell_sha256": sha(cell.tobytes()), "population": m.get("population"),
                       "live": int(cell.sum())}
    return out


def legacy_hydro_run(steps=2):
    from oracle.runner import reference_machine
    from paraiso_amd.generator.native import Language, Setup
    from paraiso_amd.generator.plan import translate
    from paraiso_amd.schemes.hydro import hydro_om_legacy
    plan = translate(Setup((1024, 1024), Language.CPlusPlus), hydro_om_legacy())
    m = reference_machine("hydro", plan.om.setup.static_values, plan.memory_size(), ["init", "proceed"])
    for k, v in LEGACY_HYDRO_SCALARS.items():
        m.set(k, v)
    m.run("init")
    out = {}
    for s in range(steps + 1):
        if s:
            m.run("proceed")
        rec = {"time": float(m.get("time"))}
        for f in ("density", "velocity0", "velocity1", "pressure"):
            a = m.get(f)
            rec[f + "_sha256"] = sha(a.tobytes())
            rec[f + "_sum"] = float(a.astype(np.float64).sum())
        out[str(s)] = rec
    return out


LEGACY_HYDRO_SCALARS = {"dR0": 1.0 / 1024, "dR1": 1.0 / 1024, "extent0": 1.0, "extent1": 1.0,
                        "cfl": 0.5, "time": 0.0, "generation": 0}


def main():
    out = {
        "om_dump": {
            "life_legacy": dump_digest(os.path.join(REF, "examples-old/Life-exampled/output/OM.txt")),
            "hydro_legacy": dump_digest(os.path.join(REF, "examples-old/Hydro-exampled/output/OM.txt")),
        },
        "legacy_life_128_open": legacy_life_run(),
        "legacy_hydro_1024_open_f32": legacy_hydro_run(),
        "legacy_hydro_scalars": LEGACY_HYDRO_SCALARS,
    }
    with open(os.path.join(HERE, "reference_outputs.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "reference_outputs.json"))


if __name__ == "__main__":
    main()
