"""The C ABI: struct layouts seen from Python equal the header's, and the
built library loads and exports every entry point the header declares
(no compute call: this runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from xmd_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xmd_hip.h")


def _c_offsets(tmp_path, struct, fields):
    src = tmp_path / "off.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){",
             f'printf("%zu\\n", sizeof({struct}));']
    for f in fields:
        lines.append(f'printf("%zu\\n", offsetof({struct}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return [int(x) for x in out]


@pytest.mark.parametrize("struct,cls", [("xmd_params", _abi.XmdParams), ("xmd_topology", _abi.XmdTopology),
                                        ("xmd_qa_contact", _abi.XmdQaContact)])
def test_struct_layout_matches_header(tmp_path, struct, cls):
    names = [f[0] for f in cls._fields_]
    got = _c_offsets(tmp_path, struct, names)
    assert got[0] == C.sizeof(cls)
    for name, off in zip(names, got[1:]):
        assert getattr(cls, name).offset == off, name


def declared_entry_points():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(xmd_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    from xmd_amd import engine
    lib = engine.lib()
    names = declared_entry_points()
    assert "xmd_engine_create" in names and "xmd_step" in names
    for name in names:
        assert hasattr(lib, name), name
    assert lib.xmd_version().decode().startswith("xmd-hip")


def test_create_without_device_fails_loudly():
    """Without a GPU the engine refuses to run (no CPU fallback)."""
    import glob
    if glob.glob("/dev/kfd"):
        pytest.skip("a GPU is visible")
    from xmd_amd import synth
    from xmd_amd._host import XmdError
    from xmd_amd.engine import Engine
    with pytest.raises(XmdError) as ei:
        Engine(synth.config_params(0), synth.config_system(0))
    assert ei.value.status == _abi.ERR_NO_DEVICE
