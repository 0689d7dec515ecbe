"""Every field of the ctypes mirrors of the C structs sits where the C
compiler puts it: the library's rrsph_params_c, rrsph_gauge_spec and
rrsph_diagnostics (include/rrsph_hip.h) and the oracle's oracle_params
(oracle/sph_oracle.h).  A field added, dropped or reordered on one side
only would otherwise hand the other side shifted values without an error.
Compiled with gcc on the host; no device is needed."""
import ctypes
import os
import subprocess
import tempfile

import pytest

import oracle
from sph2d_amd import hip as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CASES = [
    ("rrsph_params_c", H.RrsphParams, "rrsph_hip.h", os.path.join(ROOT, "include")),
    ("rrsph_gauge_spec", H.GaugeSpec, "rrsph_hip.h", os.path.join(ROOT, "include")),
    ("rrsph_diagnostics", H.Diagnostics, "rrsph_hip.h", os.path.join(ROOT, "include")),
    ("oracle_params", oracle.OracleParams, "sph_oracle.h", os.path.join(ROOT, "oracle")),
]


@pytest.mark.parametrize("cname,cls,header,incdir", CASES, ids=[c[0] for c in CASES])
def test_every_field_offset_matches_c(cname, cls, header, incdir):
    names = [f[0] for f in cls._fields_]
    lines = [f'  printf("%zu\\n", sizeof({cname}));']
    lines += [f'  printf("%zu\\n", offsetof({cname}, {n}));' for n in names]
    src = "#include <stddef.h>\n#include <stdio.h>\n#include \"%s\"\nint main(void) {\n%s\n  return 0;\n}\n" % (
        header, "\n".join(lines))
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", incdir, "-o", exe, c], check=True)
        out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(cls), (cname, out[0], ctypes.sizeof(cls))
    for n, off in zip(names, out[1:]):
        assert off == getattr(cls, n).offset, (cname, n, off, getattr(cls, n).offset)
