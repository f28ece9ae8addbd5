"""The C++ host interface (include/pimc_actions.hpp) compiled standalone with g++ and
driven the way a PIMC++ stage drives its actions, against the oracle library."""

import json
import os
import subprocess

import pytest

from tests.helpers import oracle_cdll

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def mirror_binary(tmp_path_factory):
    lib = oracle_cdll()._name
    out = str(tmp_path_factory.mktemp("mirror") / "stage_mirror")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "stage_mirror.cc"), lib,
                    f"-Wl,-rpath,{os.path.dirname(lib)}", "-o", out], check=True)
    return out


def test_stage_round_trip_through_the_mirror(mirror_binary, oracle):
    r = json.loads(subprocess.run([mirror_binary], check=True, capture_output=True, text=True).stdout)
    assert r["s_old"] != r["s_new"]
    assert r["s_new_after_reject"] == r["s_old"]          # RejectCopy restored NEW from OLD
    assert r["s_old_after_accept"] == r["s_new"]          # AcceptCopy copied NEW into OLD
    assert r["empty_request_throws"] is True
    # KineticClass::d_dBeta over the whole path: 3/(2 tau) per link minus the spring terms
    assert 0.0 < r["kin_dbeta"] < 8 * 16 * 1.5 / 0.05 + 1e-9
    assert r["dbeta_needs_tables"] is True
    assert r["time_spent"] >= 0.0


@pytest.mark.gpu
def test_stage_round_trip_gpu_matches_oracle(mirror_binary, tmp_path):
    from pimcplusplus_amd import native
    out = str(tmp_path / "stage_mirror_gpu")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "stage_mirror.cc"), native.LIB_PATH,
                    f"-Wl,-rpath,{os.path.dirname(native.LIB_PATH)}", "-o", out], check=True)
    g = json.loads(subprocess.run([out], check=True, capture_output=True, text=True, timeout=120).stdout)
    o = json.loads(subprocess.run([mirror_binary], check=True, capture_output=True, text=True).stdout)
    for key in ("s_old", "s_new", "s_new_after_reject", "s_old_after_accept", "kin_dbeta"):
        assert abs(g[key] - o[key]) <= 1e-12 * (1 + abs(o[key])), key
    assert g["s_new_after_reject"] == g["s_old"] and g["s_old_after_accept"] == g["s_new"]
    assert g["empty_request_throws"] is True and g["dbeta_needs_tables"] is True
