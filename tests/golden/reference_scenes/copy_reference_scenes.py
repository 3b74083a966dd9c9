"""Copies the reference renderer's own scene files that tests/test_reference_scenes.py and
tests/test_gpu_xml.py load verbatim, with the IOR spectra the gold conductor of testCbox.xml names,
into this directory. They are data (scene descriptions and tabulated spectra), kept byte for byte:
  cbox_fog.xml, testCbox.xml, veach-MIS.xml   RaytracerNew2/Scenes/
  data/IOR/Au.eta.spd, data/IOR/Au.k.spd      RaytracerNew2/data/IOR/
The meshes those files name (./data/cbox/*.obj, ./data/mitsuba2.obj, ...) are not part of the
reference's repository; the tests write stand-ins for them next to a copy of the scene file.

Run: python tests/golden/reference_scenes/copy_reference_scenes.py [reference root]
"""
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
FILES = [("RaytracerNew2/Scenes/cbox_fog.xml", "cbox_fog.xml"),
         ("RaytracerNew2/Scenes/testCbox.xml", "testCbox.xml"),
         ("RaytracerNew2/Scenes/veach-MIS.xml", "veach-MIS.xml"),
         ("RaytracerNew2/data/IOR/Au.eta.spd", "data/IOR/Au.eta.spd"),
         ("RaytracerNew2/data/IOR/Au.k.spd", "data/IOR/Au.k.spd")]


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    for src, dst in FILES:
        out = os.path.join(HERE, dst)
        os.makedirs(os.path.dirname(out), exist_ok=True)
        shutil.copyfile(os.path.join(root, src), out)
        print(f"{src} -> {os.path.relpath(out, HERE)}")


if __name__ == "__main__":
    main()
