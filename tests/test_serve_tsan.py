"""The serving scheduler's host code under ThreadSanitizer, on the CPU.

tests/tsan/build.py compiles csrc/serve.hip and tests/tsan/server_stub.hip
host-only with -Xarch_host -fsanitize=thread.  The stub stands in for the
HIP calls the server makes (an event is a release/acquire flag, the order
a real event gives the completion thread) and for the searches behind its
batch dispatch (a deterministic answer after a short sleep).  Four threads
submit 240 mixed requests to a three-slot server under a small admission
budget while a fifth polls its counters: no race may be reported and every
answer must equal the directly computed one."""
import importlib.util
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "tsan", "_build", "server_stub")


def _build():
    spec = importlib.util.spec_from_file_location("tsan_build", os.path.join(HERE, "tsan", "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.build()


def test_server_scheduler_has_no_data_race():
    exe = EXE if os.path.exists(EXE) else _build()
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = "halt_on_error=1 exitcode=66"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=env)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "failed 0, differing 0" in r.stdout, r.stdout
