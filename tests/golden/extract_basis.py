"""Extract the element blocks H..Ne of the basis sets this package uses from the
Gaussian-format basis-set files that ship with polymer (basisset/<name>), so
they travel with the package.  The files keep the original format (element
header after a '****' separator, 'S/P/SP/D/F nprim scale' shell lines,
Fortran 'D' exponents), which is what the basis reader parses
(src/pread.f90:611-1037 is the reference reader of that format).

Run once in a checkout that has the reference tree:
    python tests/golden/extract_basis.py /root/reference/basisset polymer_amd/basis
"""
import os
import sys

ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"]
BASES = ["STO-3G", "6-31G", "6-31Gss", "cc-pVDZ"]


def extract(src_dir, name, out_dir):
    with open(os.path.join(src_dir, name)) as f:
        lines = f.read().splitlines()
    blocks = []
    i = 0
    while i < len(lines):
        if lines[i].startswith("**") and i + 1 < len(lines):
            head = lines[i + 1].split()
            if head and head[0] in ELEMENTS:
                j = i + 2
                while j < len(lines) and not lines[j].startswith("**"):
                    j += 1
                blocks.append(lines[i + 1:j])
                i = j
                continue
        i += 1
    with open(os.path.join(out_dir, name), "w") as f:
        f.write("****\n")
        for b in blocks:
            f.write("\n".join(b) + "\n****\n")
    return len(blocks)


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/basisset"
    here = os.path.dirname(os.path.abspath(__file__))
    out = sys.argv[2] if len(sys.argv) > 2 else os.path.join(here, "..", "..", "polymer_amd", "basis")
    for b in BASES:
        n = extract(src, b, out)
        print(b, n, "elements")
