"""The request server's scheduling core (rtmuq2d_amd/csrc/rtm_sched.h, the
bookkeeping rtm_serve.hip runs over its HIP backend) on the CPU, over a
stand-in backend (tests/cpu_serve_harness.cpp) whose allocator refuses or
fails requests and whose launches fail a chosen launch group, work-list
build, pass switch or download, or everything (a sticky device error).
Each failure retires only the requests it touches, with the reason; every
other request completes with exactly its own steps; memory is released and
lists are rebuilt only after in-flight work has drained; and the server
takes new requests afterwards.  Submission from four threads while a fifth
ticks runs under ThreadSanitizer.  Host compilers only."""
from __future__ import annotations

import os
import subprocess

import pytest

import common

CSRC = os.path.join(common.REPO, "rtmuq2d_amd", "csrc")
INC = os.path.join(common.REPO, "include")
SRC = os.path.join(common.REPO, "tests", "cpu_serve_harness.cpp")
SCENARIOS = ["basic", "group_failure", "upload_failure", "no_fit", "prepare_failure", "pass_failures",
             "sticky", "threads", "pass_groups"]


def _build(tmp, flags, name):
    exe = str(tmp / name)
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", *flags, "-I", INC, "-I", CSRC, SRC, "-o", exe, "-lpthread"],
                   check=True)
    return exe


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("serve_core")
    return {"asan": _build(tmp, ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"], "h_asan"),
            "tsan": _build(tmp, ["-fsanitize=thread"], "h_tsan")}


def _run(exe, scenario):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([exe, scenario], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout[-2000:] + r.stderr[-4000:]


@pytest.mark.parametrize("scenario", SCENARIOS)
def test_scheduling_core(harness, scenario):
    _run(harness["asan"], scenario)


def test_concurrent_submission_is_race_free(harness):
    _run(harness["tsan"], "threads")
