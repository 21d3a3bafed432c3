"""Copies the Limbo DEF parser test designs the reference ships into tests/golden/lefdef/.

Source: third_party/Limbo/test/parsers/def/benchmarks/{simple,simple_w}.def of the
reference tree (Limbo, MIT license, (c) 2016 Yibo Lin, University of Texas at
Austin).  The files are data (DEF text); a provenance comment is prepended.
Run in a tree where the reference is present:
    python tests/golden/make_lefdef_fixtures.py /root/reference
"""
import os
import sys

SRC = "third_party/Limbo/test/parsers/def/benchmarks"
HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    out = os.path.join(HERE, "lefdef")
    os.makedirs(out, exist_ok=True)
    for name in ("simple.def", "simple_w.def"):
        text = open(os.path.join(ref, SRC, name)).read()
        with open(os.path.join(out, name), "w") as f:
            f.write(f"# {SRC}/{name} (Limbo, MIT license, (c) 2016 Yibo Lin)\n")
            f.write(text)
        print("wrote", os.path.join(out, name))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
