"""The CPU engines under AddressSanitizer and UndefinedBehaviorSanitizer: the
packed oracle, the bit-parallel engines (faithful and group-major), the SIMD
host engine and the generated AOT engine, rebuilt with -fsanitize=address,
undefined (oracle/Makefile's sanitize target; AOT sources compiled with the
same flags) and driven by the oracle pin tests in a child process that
preloads the ASan runtime.  Any invalid memory access or undefined
operation aborts the child (leak checking is off: the child is Python)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpu_engines_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()
    pre = os.environ.get("LD_PRELOAD")
    env = dict(os.environ, RRTL_ORACLE_SANITIZE="1", LD_PRELOAD=asan + (":" + pre if pre else ""),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    probe = ("import oracle\n"
             "from rrtl_amd import designs, lower\n"
             "p = lower.lower_module(designs.counter())\n"
             "oracle.PackedRef(p, 8).tick()\n"
             "maps = open('/proc/self/maps').read()\n"
             "assert 'libasan' in maps and '_build/san/libpacked_ref.so' in maps, 'sanitized engines not loaded'\n")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "tests/test_oracle_pins.py", "tests/test_riscv.py", "tests/test_faultsim.py"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
