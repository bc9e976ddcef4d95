"""The headless driver nbody_run without a device: it parses its options, prints its usage,
rejects unknown options and reports the missing device as a JSON error line."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "nbodysimulation_amd", "nbody_run")


@pytest.fixture(scope="module")
def exe(nbl):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "nbodysimulation_amd"), "nbody_run"], check=True,
                   capture_output=True)
    return EXE


def test_usage_and_bad_options(exe):
    r = subprocess.run([exe, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--bodies" in r.stdout and "--resume" in r.stdout
    for bad in (["--bogus"], ["--algorithm", "tree"], ["--steps", "-1"], ["--bodies"]):
        r = subprocess.run([exe, *bad], capture_output=True, text=True)
        assert r.returncode == 2, bad


def test_missing_device_is_reported(exe):
    r = subprocess.run([exe, "--bodies", "100", "--steps", "1"], capture_output=True, text=True)
    assert r.returncode == 1
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["status"] == 2 and "device" in line["error"]
