"""CLI argument handling of the drop-in (paths that exit before any GPU work):
BLQH_without_spin/init.cpp:95-117 and BLQH_with_spin/init.cpp:108-157."""
import os
import subprocess

import pytest

import blqh_amd

pytestmark = pytest.mark.skipif(not os.path.exists(blqh_amd.CLI_PATH), reason="CLI not built")


def run(args, exe=None):
    exe = exe or blqh_amd.CLI_PATH
    return subprocess.run([exe] + args, capture_output=True, text=True, timeout=60)


def test_nel_up_larger_than_nel():
    r = run(["-e", "8", "-u", "9"])
    assert r.returncode == 2
    assert r.stdout.startswith("-u: nel_up < nel !\nUsage: ")


def test_help_needs_argument_and_prints_usage():
    r = run(["-h", "x"])
    assert r.returncode == 0
    assert "Default: (l,n,u,d,lambda) = (4,4,2,1,200)" in r.stdout


def test_spinful_validation():
    exe = blqh_amd.CLI_PATH + "_spin"
    r = run(["-n", "8", "-e", "8", "-U", "8", "-u", "4"], exe)
    assert r.returncode == 2
    assert r.stdout.startswith("-U: No. of upper-layer electrons should be in [0,nel_up] !")
    r = run(["-n", "8", "-e", "40", "-U", "4", "-u", "4"], exe)
    assert r.returncode == 2 and r.stdout.startswith("-e: Total No. of electrons should be in (0,4*nphi)")
    r = run(["-n", "24", "-e", "20", "-U", "4", "-u", "4"], exe)
    assert r.returncode == 2 and r.stdout.startswith("Hilbert space too large !")
