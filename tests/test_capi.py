"""The C ABI without a GPU: the library loads, exports every symbol the
header declares, and its host-side helpers agree with the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

from swfwi_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "swfwi_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "defined", "sizeof"}
    return sorted({n for n in names if n not in skip and (n.startswith("swfwi_") or n.startswith("fd4t10s"))})


def test_header_symbols_exported():
    lib = capi.LIB_PATH
    assert os.path.exists(lib), "build the library first"
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    # and the ctypes table covers the same set
    assert set(declared_symbols()) == set(capi.SIGNATURES)


def test_library_loads_and_reports():
    L = capi.load()
    assert b"gfx950" in L.swfwi_version()
    assert capi.device_count() >= 0


def test_expand_domain_matches_oracle(oracle):
    rng = np.random.default_rng(0)
    v = rng.uniform(1500, 4500, (31, 23)).astype(np.float32)
    a = capi.expand_domain(v, 9, 12.5, 0.001)
    b = oracle.expand_domain(v, 9, 12.5, 0.001)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_ricker_matches_oracle(oracle):
    for fm in (10.0, 15.0):
        a = capi.ricker(4000, fm, 0.001, 1.0)
        b = oracle.ricker(4000, fm, 0.001, 1.0)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_random_codes_match_oracle(oracle):
    for seed in (1, 2):
        assert np.array_equal(capi.random_codes(seed, 60), oracle.gen_codes(seed, 60))
        assert np.array_equal(capi.random_codes(seed, 60, skip=120), oracle.gen_codes(seed, 60, skip=120))


def test_errors_are_reported_not_fatal():
    with pytest.raises(capi.SwfwiError) as e:
        capi.expand_domain(np.zeros((0, 3), np.float32), 2, 1.0, 1.0)
    assert e.value.code == 1


@pytest.mark.parametrize("kw,code,needle", [
    (dict(nx=50000, nz=50000), 5, "32-bit"),               # padded grid beyond 2^31 points
    (dict(nb=4), 1, "nb must be >= 6"),                     # the band must lie in the sponge
    (dict(sx=[80]), 1, "sources exceeds"),                  # Params::check of the tools
    (dict(gx=[0], gz=[40]), 1, "geophones exceeds"),
    (dict(dt=0.0), 1, "positive"),
    (dict(precision=7), 1, "precision"),
])
def test_solver_arguments_are_checked_before_any_device_call(kw, code, needle):
    """Solver creation validates its geometry on the host first, so a bad
    argument reports SWFWI_ERR_ARG / SWFWI_ERR_RANGE (with the reference's
    messages) rather than a device error, with or without a GPU."""
    a = dict(nx=60, nz=40, nb=8, dt=0.001, dx=10.0, fm=15.0, nt=100, sx=[5], sz=[1], gx=[0, 1, 2], gz=[0, 0, 0],
             precision=0)
    a.update(kw)
    i32 = lambda v: np.array(v, np.int32)  # noqa: E731
    with pytest.raises(capi.SwfwiError) as e:
        capi.Solver(a["nx"], a["nz"], a["nb"], a["dt"], a["dx"], a["fm"], a["nt"], i32(a["sx"]), i32(a["sz"]),
                    i32(a["gx"]), i32(a["gz"]), a["precision"])
    assert e.value.code == code, e.value
    assert needle in str(e.value)
