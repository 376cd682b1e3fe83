"""Write the golden fixtures used by the parity tests.

Reads the reference's own fixture files (tests/data/*.json of hf-rs, produced
there by PySCF 2.13/2.14 and mpmath) and keeps the values the Fock-build tests
need, so that the tests run without the reference tree.  Data only: inputs and
expected outputs.

Usage:  python tests/golden/make_golden.py /root/reference/tests/data
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(src):
    load = lambda f: json.load(open(os.path.join(src, f)))

    boys = load("boys_reference.json")
    json.dump({"source": boys["source"], "rows": boys["rows"]}, open(os.path.join(HERE, "boys.json"), "w"))

    sph = load("spherical_reference.json")
    spdf = sph["spdf_integrals"]
    meta = {k: spdf[k] for k in ("source", "symbols", "geometry_bohr", "basis_gaussian94", "shell_ranges",
                                 "cartesian_order", "quartet_shape")}
    json.dump(meta, open(os.path.join(HERE, "spdf_meta.json"), "w"), indent=1)
    np.savez_compressed(os.path.join(HERE, "spdf_eri.npz"), eri=np.array(spdf["eri"], dtype=np.float64))

    scf = load("scf_reference.json")
    systems = {}
    for name, v in scf["systems"].items():
        res = {b: {"reference": r["reference"], "energy": r["energy"], "n_ao": r["n_ao"], "gradient": r["gradient"]}
               for b, r in v["results"].items()}
        systems[name] = {"geometry_bohr": v["geometry_bohr"], "charge": v["charge"], "spin_2s": v["spin_2s"],
                         "results": res}
    w = sph["systems"]["water_ccpvdz"]
    systems["water_ccpvdz_spherical"] = {
        "geometry_bohr": "O 0 0 0; H 1.43 0.2 1.1; H -1.2 -0.3 1.3", "charge": 0, "spin_2s": 0,
        "results": {"cc-pvdz": {"reference": "rhf", "energy": float(w["energy"]), "n_ao": int(w["n_ao"]),
                                "spherical": True}}}
    json.dump({"source": scf["source"], "systems": systems}, open(os.path.join(HERE, "scf_energies.json"), "w"),
              indent=1)

    df = load("df_reference.json")
    json.dump({k: df[k] for k in ("source", "n_ao", "n_aux", "metric", "three")},
              open(os.path.join(HERE, "df_reference.json"), "w"))
    dfhf = load("df_hf_reference.json")
    water = [c for c in dfhf["cases"] if c["name"] == "water"][0]
    json.dump({"source": dfhf["source"], "water": water}, open(os.path.join(HERE, "df_hf_water.json"), "w"))
    sad = load("sad_reference.json")
    json.dump(sad, open(os.path.join(HERE, "sad_reference.json"), "w"))
    cd = load("cd_hf_reference.json")
    json.dump(cd, open(os.path.join(HERE, "cd_hf_reference.json"), "w"), indent=1)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/data")
