"""The host code of the library and the oracle built with AddressSanitizer
and UndefinedBehaviorSanitizer (tests/sanitize/build_san.py) and run over
their CPU-only paths: parsing and printing, the planners and emulators, the
batched form's kernel code on host threads, and every oracle kernel.  Any
sanitizer report fails the test."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SCRIPT = os.path.join(HERE, "sanitize", "build_san.py")


@pytest.mark.skipif(not os.path.exists(SCRIPT) or shutil.which("gcc") is None
                    or not os.path.exists("/opt/rocm/llvm/bin/clang++"), reason="needs the sanitizer build tools")
def test_host_code_is_clean_under_address_and_undefined_behaviour_sanitizers():
    r = subprocess.run([sys.executable, SCRIPT], capture_output=True, text=True, timeout=1500)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    assert "oracle driver: ok" in r.stdout and "library driver: ok" in r.stdout
