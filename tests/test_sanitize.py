"""AddressSanitizer + UndefinedBehaviorSanitizer builds (CPU, no GPU needed):

- the oracle (nfft_oracle.c) with a driver over every transform kind and
  operation, 1-D and N-D, in and out of place, and the rejected sizes
  (oracle/sanitize_check.c; gcc; leak checking on);
- the library with its host code instrumented (each -fsanitize= behind
  -Xarch_host; device code unchanged) and a driver over the argument and
  error paths of the extended C ABI and of both drop-in libraries
  (notorious-fft_amd/csrc/sanitize_args.cpp, sanitize_dropin.cpp).  Leak
  checking is off there: the HIP runtime the library loads is not
  instrumented.
Either driver exits non-zero on a failed expectation, and the sanitizers
abort on the first memory or undefined-behaviour error."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _make(*args):
    r = subprocess.run(["make", *args], cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-4000:]


def _run(cmd, env_extra):
    env = dict(os.environ)
    env.update(env_extra)
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                       timeout=300)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-4000:]
    assert "AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout


def test_oracle_under_sanitizers():
    _make("-C", "oracle", "sanitize")
    _run([os.path.join(ROOT, "oracle", "_san", "oracle_check")],
         {"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "print_stacktrace=1"})


def test_library_host_paths_under_sanitizers():
    _make("-C", "notorious-fft_amd/csrc", "sanitize", "-j8")
    san = os.path.join(ROOT, "notorious-fft_amd", "_san")
    _run([os.path.join(san, "sanitize_args"), os.path.join(san, "libnotorious_fft_hip_f64.so"),
          os.path.join(san, "libnotorious_fft_hip_f32.so")],
         {"ASAN_OPTIONS": "detect_leaks=0", "UBSAN_OPTIONS": "print_stacktrace=1"})
