"""Copies the reference's own test data (input decks and the one golden
output set) into tests/golden/ so that tests never read the reference tree at
run time.  These are data files (decks: key/value text; golden: printed
field values), copied unchanged.

Provenance: test/<deck>/<deck>.pnt and test/sedov/sedov.{xy,geo,case,zr,ze,zp} of
StanfordLegion/pennant-legion.
"""
import shutil
import sys
from pathlib import Path

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
HERE = Path(__file__).resolve().parent

DECKS = ["sedovsmall", "sedov", "sedovbig", "noh", "nohsmall", "nohsquare", "leblanc",
         "leblancbig"]


def main():
    (HERE / "decks").mkdir(exist_ok=True)
    for d in DECKS:
        shutil.copyfile(REF / "test" / d / f"{d}.pnt", HERE / "decks" / f"{d}.pnt")
    (HERE / "sedov").mkdir(exist_ok=True)
    for ext in ("xy", "geo", "case", "zr", "ze", "zp"):
        shutil.copyfile(REF / "test" / "sedov" / f"sedov.{ext}", HERE / "sedov" / f"sedov.{ext}")


if __name__ == "__main__":
    main()
