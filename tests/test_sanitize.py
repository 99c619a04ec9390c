"""The engine's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

`lapis-silo_amd/csrc/sanitize_check` is the engine's host sources built with both sanitizers
(the counterpart of the reference's Debug build with ASAN, CMakeLists.txt:11) and linked with a
driver of the C ABI.  Its host mode runs the entry points that need no device: the diff of aligned
sequences, the saved-state scanners and bitmap decoder on whole, truncated and corrupted buffers,
and the NDJSON and
Arrow IPC writers at every capacity below the one needed.  Any sanitizer finding ends the process
with a report and a non-zero status."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "lapis-silo_amd", "csrc", "sanitize_check")


def run_checks(mode: str, timeout: int) -> str:
    assert os.path.exists(BINARY), "sanitize_check is not built (make -C lapis-silo_amd/csrc sanitize_check)"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([BINARY, mode], capture_output=True, text=True, timeout=timeout, env=env)
    out = p.stdout + p.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert p.returncode == 0, out[-4000:]
    assert f"{mode} checks: 0 failed" in out, out[-4000:]
    return out


def test_host_entry_points_under_sanitizers():
    run_checks("host", 120)
