"""Copy the reference's committed OpenCV fixture data into tests/golden/.

The fixtures are data generated by OpenCV 4.10.0 / NumPy 2.1.3 (AutoRIFT.jl,
test/fixtures/manifest.json; MIT licence): flat C-order .bin arrays plus meta.json.
OpenCV is absent here, so they cannot be regenerated; they are copied verbatim.
Groups kept: ZNCC matchTemplate surfaces, pyrUp steps and cascades, minMaxLoc peaks and
exact L2 distance transforms.
"""
import os
import shutil
import sys

SRC = "/root/reference/test/fixtures"
DST = os.path.dirname(os.path.abspath(__file__))


def wanted(group, case):
    if group == "matchtemplate":
        return case.startswith("zncc_")
    return group in ("pyrup", "peak", "disttransform")


def main():
    if not os.path.isdir(SRC):
        sys.exit(f"{SRC} not found")
    n = 0
    for group in sorted(os.listdir(SRC)):
        gdir = os.path.join(SRC, group)
        if not os.path.isdir(gdir):
            continue
        for case in sorted(os.listdir(gdir)):
            if not wanted(group, case):
                continue
            out = os.path.join(DST, group, case)
            os.makedirs(out, exist_ok=True)
            for f in os.listdir(os.path.join(gdir, case)):
                shutil.copyfile(os.path.join(gdir, case, f), os.path.join(out, f))
            n += 1
    shutil.copyfile(os.path.join(SRC, "manifest.json"), os.path.join(DST, "manifest.json"))
    print(f"copied {n} cases")


if __name__ == "__main__":
    main()
