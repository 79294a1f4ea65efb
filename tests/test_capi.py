"""The C-ABI library: it loads, exports every entry point include/climage.h
declares, and its structs have the layout the Python front end assumes.
No compute call is made (this runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import climage_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "climage.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(clim_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = ca.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(ca.exported_symbols()) == sorted(n for n in ca.exported_symbols() if n in names)


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text('#include "climage.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(clim_image), '
                   'sizeof(clim_denoise_params), sizeof(clim_rgb_conversion_params), sizeof(clim_noise_model), '
                   'sizeof(clim_demosaic_params), offsetof(clim_demosaic_params, levels)); return 0;}\n')
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(ca.Image), C.sizeof(ca.DenoiseParameters), C.sizeof(ca.RGBConversionParameters),
            C.sizeof(ca.NoiseModel), C.sizeof(ca.DemosaicParameters), ca.DemosaicParameters.levels.offset]
    assert got == want


def test_host_math_without_device():
    # pure host helpers of the ABI work without a GPU
    m = ca.cam_ycbcr([[1, 0, 0], [0, 1, 0], [0, 0, 1]])
    assert abs(m[0, 0] - 0.2126729) < 1e-7
    p = ca.default_params(ca.BGGR, 64.0, 16383.0)
    assert p.bayerPattern == ca.BGGR and p.levels == 5 and abs(p.scale_mul[0] - 65535.0 / (16383.0 - 64.0)) < 1e-4
    assert p.rgbConversionParameters.contrast == C.c_float(1.05).value
