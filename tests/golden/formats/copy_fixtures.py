"""Copy the reference's own data files used by tests/test_formats.py.

The state file (peachgk-md/sta_peachgk-2.0.dat, gzip-compressed), the energy
and pressure outputs (peachgk-md/out_ene.dat, out_pre.dat) and the control
script (peachgk-md/peachgk.ini) are data the reference ships; they are stored
here unchanged.

    python tests/golden/formats/copy_fixtures.py [REFDIR]
"""
import gzip
import os
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/peachgk-md"

with open(os.path.join(REF, "sta_peachgk-2.0.dat"), "rb") as f, \
        gzip.open(os.path.join(HERE, "sta_peachgk-2.0.dat.gz"), "wb", compresslevel=9) as g:
    shutil.copyfileobj(f, g)
for name in ("out_ene.dat", "out_pre.dat", "peachgk.ini"):
    shutil.copyfile(os.path.join(REF, name), os.path.join(HERE, name))
