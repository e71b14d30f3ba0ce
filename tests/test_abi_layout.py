"""The ctypes mirrors of the descriptor structs match the C header's layout
(compiled here with gcc: no device work)."""
import ctypes as C
import os
import subprocess

from conftest import ROOT


def _c_sizes(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "plenvdb_hip.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu\\n", sizeof(pvdb_step_desc),'
                   ' offsetof(pvdb_step_desc, adam_scalars), sizeof(pvdb_render_desc),'
                   ' offsetof(pvdb_render_desc, capacity), offsetof(pvdb_step_desc, launch),'
                   ' offsetof(pvdb_step_desc, weight_rgbper), offsetof(pvdb_step_desc, color_master));'
                   ' return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    return [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True,
                                           text=True).stdout.split()]


def test_descriptor_layouts_match_header(tmp_path):
    from plenvdb_amd import fused, renderer

    step_size, step_off, render_size, render_off, launch_off, rgbper_off, master_off = \
        _c_sizes(tmp_path)
    assert C.sizeof(fused.StepDesc) == step_size
    assert fused.StepDesc.adam_scalars.offset == step_off
    assert fused.StepDesc.launch.offset == launch_off
    assert fused.StepDesc.weight_rgbper.offset == rgbper_off
    assert fused.StepDesc.color_master.offset == master_off
    assert C.sizeof(renderer.RenderDesc) == render_size
    assert renderer.RenderDesc.capacity.offset == render_off
