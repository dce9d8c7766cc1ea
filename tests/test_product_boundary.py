"""The product path has no CPU fallback: without libobit_hip.so its
drivers raise, and no module of the package imports the oracle."""
import os
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]


def test_missing_library_raises():
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from obit_amd import _lib\n"
        "_lib.LIB_PATH = '/nonexistent/libobit_hip.so'\n"
        "try:\n"
        "    _lib.load()\n"
        "except Exception as e:\n"
        "    print('raised', type(e).__name__)\n"
        "else:\n"
        "    print('loaded')\n" % str(ROOT))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert "raised" in out.stdout, out.stdout + out.stderr


def test_package_never_imports_the_oracle():
    for p in (ROOT / "obit_amd").glob("*.py"):
        text = p.read_text()
        assert "import oracle" not in text and "from oracle" not in text, p.name
