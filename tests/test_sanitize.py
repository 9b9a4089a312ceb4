"""The host-only parts of the library (the .imt diagnostics-text reader and
writer, the MV-data wire format, the geometry rationalisation, the host
lambda maps, the synthetic generator) built with AddressSanitizer and UndefinedBehaviorSanitizer
(tests/sanitize/Makefile) and driven by tests/sanitize/host_fuzz.cpp: a
reference-written .imt file must round-trip byte for byte, and thousands
of mutated copies of it, random parameter sets and random sizes must run
without a single sanitizer report (any report aborts the driver)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
@pytest.mark.parametrize("seed", [7, 1234567])
def test_host_code_under_address_and_undefined_behaviour_sanitizers(seed):
    subprocess.run(["make", "-C", SAN], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(SAN, "_build", "host_fuzz"), os.path.join(HERE, "golden", "me_cif.imt"), "1500",
                        str(seed)], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "records round-tripped" in r.stdout and "MV-data fields round-tripped" in r.stdout
