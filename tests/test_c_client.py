"""The plain-C client of the C ABI (tests/c_client/xmd_c_client.c) links
the engine library without Python and gets every malformed input refused
with XMD_ERR_ARG and a message (no device needed: the checks run first)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLIENT = os.path.join(ROOT, "tests", "c_client", "xmd_c_client")


def client():
    if not os.path.exists(CLIENT):
        subprocess.run(["make", "-s", "-C", os.path.dirname(CLIENT)], check=True)
    return CLIENT


def test_c_client_errors():
    out = subprocess.run([client(), "errors"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("xmd-hip")
    assert "errors ok" in out.stdout
    for what in ("dt", "chain offsets", "native order", "atype"):
        assert f"{what} refused:" in out.stdout
