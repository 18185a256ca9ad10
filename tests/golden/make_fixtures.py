"""Extract the golden fixtures used by the tests from the reference's data files.

Run once in a container that has the reference checkout mounted:

    python tests/golden/make_fixtures.py /root/reference

Only data is copied: inputs (potential.in, ga.in, train.in gzip-compressed),
the reference's own outputs (energy/force/virial.out) and a subsample of each
ga.out (every 50th generation plus the last two lines).  For the two
wurtzite cases also the DFT data their train.in was converted from (the
converter's input.txt and the data files it names, bundled into
dft_inputs.txt.gz: a "#list" line, then input.txt, then a "#file <name>"
line before each file).  No source file of the reference is read or copied.

Cases (reference path -> fixture directory):
  tests/si_diamond            -> si_diamond   minimal Tersoff, one species
  examples/si_diamond         -> si_example   same train.in, another GA run
  develop/hfo2                -> hfo2         RI, every parameter fixed
  develop/nn2b                -> nn2b         NN2B, 10 neurons, dimers
  develop/wurzite/GaAs        -> gaas         minimal Tersoff, two species
  develop/wurzite/Si          -> si_type2     minimal Tersoff, two species
"""
import gzip
import os
import shutil
import sys

CASES = [
    ("tests/si_diamond", "si_diamond", {"ga1.out": "ga.out", "energy1.out": "energy.out",
                                         "force1.out": "force.out", "virial1.out": "virial.out"}),
    ("examples/si_diamond", "si_example", None),
    ("develop/hfo2", "hfo2", None),
    ("develop/nn2b", "nn2b", None),
    ("develop/wurzite/GaAs", "gaas", None),
    ("develop/wurzite/Si", "si_type2", None),
]
DEFAULT = {"ga.out": "ga.out", "energy.out": "energy.out", "force.out": "force.out",
           "virial.out": "virial.out"}


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    for src, dst, outs in CASES:
        s = os.path.join(ref, src)
        d = os.path.join(here, dst)
        os.makedirs(d, exist_ok=True)
        for name in ("potential.in", "ga.in"):
            shutil.copyfile(os.path.join(s, name), os.path.join(d, name))
        with open(os.path.join(s, "train.in"), "rb") as fi, \
                gzip.GzipFile(os.path.join(d, "train.in.gz"), "wb", 9, mtime=0) as fo:
            fo.write(fi.read())
        for name_in, name_out in (outs or DEFAULT).items():
            text = open(os.path.join(s, name_in)).read()
            if name_out == "ga.out":
                lines = text.strip("\n").split("\n")
                keep = [l for k, l in enumerate(lines) if k % 50 == 0 or k >= len(lines) - 2]
                text = "\n".join(keep) + "\n"
            with open(os.path.join(d, name_out), "w") as fo:
                fo.write(text)
        if src.startswith("develop/wurzite/"):
            bundle_dft(s, d)
        print("wrote", d)


def bundle_dft(s, d):
    listing = open(os.path.join(s, "input.txt")).read().replace("\r\n", "\n")
    parts = ["#list\n", listing.rstrip("\n") + "\n"]
    for name in listing.split()[2:]:
        text = open(os.path.join(s, name)).read().replace("\r\n", "\n")
        parts += ["#file %s\n" % name, text.rstrip("\n") + "\n"]
    with gzip.GzipFile(os.path.join(d, "dft_inputs.txt.gz"), "wb", 9, mtime=0) as fo:
        fo.write("".join(parts).encode())


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
