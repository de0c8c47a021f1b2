"""Runs the sanitizer build of the host-only code (tests/cpp/host_check.cpp, built by
tests/cpp/Makefile with g++ -fsanitize=address,undefined) and fails on any sanitizer report
or crash.  The program prints the C-ABI status of each call ("rc=N"), returned here."""

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "_build", "host_check_san")


def build():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-C", os.path.join(HERE, "cpp"), "sanitize"], check=True,
                       stdout=subprocess.DEVNULL)
    return BIN


def run(*args, timeout=300):
    env = dict(os.environ)
    # every report is fatal; the ASan runtime need not come first among preloaded libraries
    env["ASAN_OPTIONS"] = "detect_leaks=1:abort_on_error=0:verify_asan_link_order=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env.setdefault("OMP_NUM_THREADS", "2")
    p = subprocess.run([build(), *map(str, args)], capture_output=True, text=True,
                       timeout=timeout, env=env)
    report = p.stderr
    assert p.returncode == 0, f"host_check {args} exit {p.returncode}:\n{report[-4000:]}"
    assert "Sanitizer" not in report and "runtime error" not in report, report[-4000:]
    return p.stdout
