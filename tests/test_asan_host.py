"""Host AddressSanitizer run of the kernels' per-site Wilson and force
arithmetic (cpucheck.cpp + asan_driver.cpp, built with -fsanitize=address
for the host only): every (ndim, Nc, representation, hermitian, boundary,
mode) variant on buffers of exactly the sizes the kernels assume must run
clean, and a planted one-site-short input must be reported, which shows the
sanitizer is active."""
import os
import subprocess

import pytest

from tfleonardym_amd import build

SOURCES = [os.path.join(build.CSRC, f) for f in ("cpucheck.cpp", "asan_driver.cpp", "wilson_site.h",
                                                 "force_site.h", "lym_types.h")]


@pytest.fixture(scope="module")
def driver():
    b = build.ASAN_BIN
    if not os.path.exists(b) or os.path.getmtime(b) < max(os.path.getmtime(s) for s in SOURCES):
        build.build_asan_check()
    return b


def _run(args):
    env = dict(os.environ, ASAN_OPTIONS="verify_asan_link_order=0:detect_leaks=0")
    return subprocess.run(args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600, env=env)


def test_site_arithmetic_is_clean_under_asan(driver):
    p = _run([driver])
    assert p.returncode == 0, p.stderr[-2000:]
    assert "asan ok: 192 cases" in p.stdout


def test_asan_reports_a_planted_overflow(driver):
    p = _run([driver, "short"])
    assert p.returncode != 0
    assert "heap-buffer-overflow" in p.stderr
