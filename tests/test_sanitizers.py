"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the CPU side.

* The FP64 oracle is rebuilt with both sanitizers (``make -C oracle
  sanitize``) and its whole test suite runs against that build, with libasan
  preloaded into a child interpreter; any report aborts the child.
* The C-ABI layer (api.cpp: argument checks, gradient tables, shells, the
  progress hook) is rebuilt with host-only sanitizers and linked with the
  regular kernel objects; the ABI tests, which make no device call, run
  against it the same way.
* The sanitizer flags are shown to catch a planted heap overflow and a
  planted signed overflow, so a clean run means something.
"""

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle"
SAN_ORACLE = ORACLE / "_san" / "liboracle_shore.so"
ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=1",
       "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def _gcc_lib(name):
    return subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True,
                          check=True).stdout.strip()


def _make_var(name):
    r = subprocess.run(["make", "-s", "-C", str(ORACLE), "--eval", f"print-var: ; @echo $({name})", "print-var"],
                       capture_output=True, text=True, check=True)
    return r.stdout.split()


@pytest.fixture(scope="module")
def san_oracle():
    subprocess.run(["make", "-s", "-C", str(ORACLE), "sanitize"], check=True)
    return SAN_ORACLE


PLANTED = {
    "heap-buffer-overflow": "#include <cstdio>\n#include <cstring>\nint main(int c, char**) { "
                            "char* a = new char[4]; std::memset(a, 1, 3 + c + 1); "
                            "std::printf(\"%d\\n\", a[0]); delete[] a; return 0; }\n",
    "signed integer overflow": "#include <climits>\n#include <cstdio>\nint main(int c, char**) { "
                               "int x = INT_MAX - 1 + c; x += c; std::printf(\"%d\\n\", x); return 0; }\n",
}


@pytest.mark.parametrize("kind", sorted(PLANTED))
def test_sanitizer_flags_catch_planted_bugs(tmp_path, kind):
    src = tmp_path / "planted.cpp"
    src.write_text(PLANTED[kind])
    exe = tmp_path / "planted"
    flags = [f for f in _make_var("SANFLAGS") if f not in ("-fPIC",)]
    subprocess.run(["g++", *flags, str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, env={**os.environ, **ENV})
    assert r.returncode != 0
    assert kind in r.stderr


def _preload(rt):
    # the sanitizer runtime goes first; whatever the environment preloads stays
    return ":".join(p for p in (str(rt), os.environ.get("LD_PRELOAD", "")) if p)


def _run_child(code, env, timeout=600):
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=timeout,
                       env={**os.environ, **ENV, **env})
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    return out


def test_oracle_suite_under_sanitizers(san_oracle):
    tests = sorted(str(p.relative_to(ROOT)) for p in (ROOT / "tests").glob("test_oracle_*.py"))
    assert len(tests) >= 5
    code = ("import sys, pytest, oracle\n"
            "oracle.load()\n"
            f"assert {str(san_oracle)!r} in open('/proc/self/maps').read(), 'sanitized oracle not mapped'\n"
            f"sys.exit(pytest.main({tests!r} + ['-q', '-x', '-p', 'no:cacheprovider', '-m', 'not gpu']))\n")
    out = _run_child(code, {"ORACLE_LIB": str(san_oracle), "LD_PRELOAD": _preload(_gcc_lib("libasan.so"))})
    assert " passed" in out


def test_abi_layer_under_host_sanitizers():
    from cs_dmri_amd import build as B
    lib = B.build_host_sanitized()
    code = ("import sys, pytest\n"
            "from cs_dmri_amd import _lib as L\n"
            "L.lib()\n"
            f"assert {str(lib)!r} in open('/proc/self/maps').read(), 'sanitized library not mapped'\n"
            "sys.exit(pytest.main(['tests/test_lib_abi.py', '-q', '-x', '-p', 'no:cacheprovider', "
            "'-m', 'not gpu']))\n")
    out = _run_child(code, {"SHOREFIT_LIB": str(lib), "LD_PRELOAD": _preload(B.asan_runtime())})
    assert " passed" in out
