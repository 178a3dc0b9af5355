"""Copies the PNG data files of the reference's renderer test
(src/tests/*.png: five material textures and six 64 x 64 sky faces) into
ref_textures/ as test fixtures.  Data only; run where the reference tree is
present."""
import shutil
import sys
from pathlib import Path

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference") / "src" / "tests"
DST = Path(__file__).resolve().parent / "ref_textures"

if __name__ == "__main__":
    DST.mkdir(exist_ok=True)
    for png in sorted(SRC.glob("*.png")):
        shutil.copyfile(png, DST / png.name)
        print(png.name)
