"""CPU-only checks of the drop-in boundary: the C-ABI library loads, exports
every entry point include/rrsph_hip.h declares, and the ctypes mirror of
rrsph_params_c has the C layout.  No compute call is made (no GPU here)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from sph2d_amd import hip as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rrsph_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w]+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n.startswith(("rrsph_hip_", "RRSPH_HIP_"))))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(H.LIB_PATH):
        pytest.fail("librrsph_hip.so not built (python -m sph2d_amd.build)")
    lib = ctypes.CDLL(H.LIB_PATH)
    decl = declared_functions()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(lib, name), name
    assert sorted(H.EXPORTED) == decl


def test_version_hooks_without_device():
    L = H.load_library()
    assert L.RRSPH_HIP_GetSpecificVersionName().decode() == "RRSPH_HIP"
    assert L.RRSPH_HIP_GetSpecificVersionMajor() == 1


def test_params_struct_layout_matches_c():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "rrsph_hip.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(rrsph_params_c), offsetof(rrsph_params_c, piston_a),
         offsetof(rrsph_params_c, precision), offsetof(rrsph_params_c, start_time), sizeof(rrsph_diagnostics),
         sizeof(rrsph_gauge_spec), offsetof(rrsph_gauge_spec, x), offsetof(rrsph_params_c, piston_seg_zmax),
         offsetof(rrsph_gauge_spec, z));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    size, off_a, off_prec, off_t, dsize, gsize, goff, off_seg, goffz = map(int, out)
    assert gsize == ctypes.sizeof(H.GaugeSpec) and goff == H.GaugeSpec.x.offset and goffz == H.GaugeSpec.z.offset
    assert off_seg == H.RrsphParams.piston_seg_zmax.offset
    assert size == ctypes.sizeof(H.RrsphParams)
    assert off_a == H.RrsphParams.piston_a.offset
    assert off_prec == H.RrsphParams.precision.offset
    assert off_t == H.RrsphParams.start_time.offset
    assert dsize == ctypes.sizeof(H.Diagnostics)


def test_create_without_device_reports_error():
    """With no HIP device the library must refuse (no silent CPU fallback)."""
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a device is present")
    except ImportError:
        pass
    from sph2d_amd import generators as G

    lay = G.dam_break()
    with pytest.raises(H.RrsphError) as e:
        H.HipStep(lay.params)
    assert e.value.status == 8


def _no_device():
    try:
        import torch

        return not torch.cuda.is_available()
    except ImportError:
        return True


@pytest.mark.parametrize(
    "change,status,text",
    [
        ({"dim": 4}, 1, "dim must be 2 or 3"),
        ({"dim": 3, "density_treatment": 2}, 7, "3D supports cubic kernels"),
        ({"ntotal": 0}, 1, "invalid sizes"),
        ({"max_neighbours": 1}, 1, "invalid sizes"),
        ({"x_maxgeom": 1.0e9}, 6, "can't make grid with so many cells"),
        ({"nfluid": 10**7}, 1, "nfluid > ntotal"),
        ({"piston_segments": 2, "piston_seg_zmax": 1.0}, 1, "a segmented piston needs dim 3"),
        ({"piston_segments": 9}, 1, "piston_segments at most 8"),
    ],
)
def test_create_validates_params_on_host(change, status, text):
    """Parameter errors are reported before any device is touched, with the
    reference's message where it has one (Input.cpp:64-70 for the grid)."""
    import copy

    from sph2d_amd import generators as G

    prm = copy.copy(G.dam_break().params)
    for k, v in change.items():
        setattr(prm, k, v)
    with pytest.raises(H.RrsphError) as e:
        H.HipStep(prm)
    assert e.value.status == status
    assert text in str(e.value)


def test_dropin_driver_reports_errors_like_the_reference(tmp_path):
    """The C++ drop-in (include/rrsph_hip.hpp, driven by tests/cpp/dropin_driver)
    turns a library error into a std::exception caught in main, printed as
    'Catch exception: <message>' (sources/SPH.cpp:76-85)."""
    import dropin

    from sph2d_amd import generators as G

    if not os.path.exists(dropin.DRIVER):
        pytest.fail("dropin_driver not built (python -m sph2d_amd.build)")
    lay = G.dam_break()
    prm = lay.params
    prm.max_neighbours = 1
    res = dropin.run(tmp_path, prm, lay, nsteps=1)
    assert res.returncode == 1
    assert "Catch exception: RRSPH_HIP: invalid sizes" in res.stdout
    if _no_device():
        prm.max_neighbours = 64
        res = dropin.run(tmp_path, prm, lay, nsteps=1)
        assert res.returncode == 1
        assert "Catch exception: RRSPH_HIP: no HIP device" in res.stdout
