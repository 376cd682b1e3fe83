"""Write the Gaussian-94 basis-set data this engine needs into ``*.g94`` files.

The numbers are Basis Set Exchange exports (BSD 3-Clause, see LICENSE.BSE) that
hf-rs embeds as string constants in src/basis_data.rs (STO-3G :15, 6-31G :2398,
6-31G* :3344, def2-SVP :6334, cc-pVDZ :19496, cc-pVTZ :23095,
def2-universal-JKFIT :42267, def2-SVP-RIFIT :51585).  Only the elements H..Ne
are kept (H..Ar for 6-31G, whose Cholesky fixtures include HCl): every
workload and fixture used here is made of them.

Usage:  python extract_basis.py /root/reference/src/basis_data.rs
"""
import os
import re
import sys

SETS = {
    "STO_3G_GBS": "sto-3g",
    "BASIS_6_31G_GBS": "6-31g",
    "BASIS_6_31G_STAR_GBS": "6-31g_star",
    "BASIS_DEF2_SVP_GBS": "def2-svp",
    "BASIS_CC_PVDZ_GBS": "cc-pvdz",
    "BASIS_CC_PVTZ_GBS": "cc-pvtz",
    "AUX_DEF2_UNIVERSAL_JKFIT_GBS": "def2-universal-jkfit",
    "AUX_DEF2_SVP_RIFIT_GBS": "def2-svp-rifit",
}
KEEP = {"H", "HE", "LI", "BE", "B", "C", "N", "O", "F", "NE"}
KEEP_THIRD_ROW = KEEP | {"NA", "MG", "AL", "SI", "P", "S", "CL", "AR"}


def blocks(text):
    """Split G94 text into per-element blocks terminated by '****'."""
    out, cur = [], []
    for line in text.splitlines():
        if not line.strip():
            continue
        cur.append(line)
        if line.strip() == "****":
            out.append(cur)
            cur = []
    return out


def main(path):
    src = open(path).read()
    here = os.path.dirname(os.path.abspath(__file__))
    for const, name in SETS.items():
        m = re.search(r"pub const %s: &str = r#\"(.*?)\"#;" % const, src, re.S)
        if not m:
            raise SystemExit("constant %s not found" % const)
        keep = KEEP_THIRD_ROW if name == "6-31g" else KEEP
        kept = [b for b in blocks(m.group(1)) if b[0].split()[0].upper() in keep]
        with open(os.path.join(here, name + ".g94"), "w") as f:
            f.write("! Basis Set Exchange export (BSD 3-Clause, LICENSE.BSE); elements %s\n"
                    % ("H..Ar" if keep is KEEP_THIRD_ROW else "H..Ne"))
            for b in kept:
                f.write("\n".join(b) + "\n")
        print(name, len(kept), "elements")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/basis_data.rs")
