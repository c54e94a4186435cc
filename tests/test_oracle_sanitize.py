"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: the
driver tests/sanitize/oracle_driver.c includes oracle/owl_oracle.c and runs
every entry point over random small geometries (2-D / 3-D, strides, rates,
SAME / VALID, the GEBP branch cut into several blocks, pooling windows with
NaN and -inf, argmax, clamped upsampling), checking the FP32 Owl algorithm
against the FP64 loops and argmax against the pooled values.  Any sanitizer
report aborts the run (-fno-sanitize-recover=all)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_clean_under_asan_ubsan(tmp_path):
    exe = tmp_path / "oracle_driver"
    subprocess.run(["gcc", "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-ffp-contract=off",
                    os.path.join(HERE, "sanitize", "oracle_driver.c"), "-o", str(exe), "-lm"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "oracle driver: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
