"""The allocation-free forward modelling entry point
(swfwi_forward_modeling_noalloc) through the C ABI, against the allocating
swfwi_forward_modeling on the same handle: bit-identical seismograms on two
calls in a row, the second on a smaller geometry (the fixed per-GPU
workspace keeps the first call's contents), for an encoded supershot and a
single shot; a geometry beyond the workspace is refused with
SWFWI_ERR_RANGE before any device work."""
import numpy as np
import pytest

from swfwi_amd import capi, workloads

pytestmark = pytest.mark.gpu

WORKSPACE_BYTES = 64 << 20


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if capi.device_count() < 1:
        pytest.skip("no HIP device")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def need_bytes(case):
    r = lambda n: (n + 63) // 64 * 64  # noqa: E731
    return 4 * (2 * r(case.nxpad * case.nzpad) + r(case.nt * case.ng) + r(case.nt * case.ns) + r(2))


def solver_for(case):
    vtrue, _ = workloads.models(case)
    vt = capi.expand_domain(vtrue, case.nb, case.dx, case.dt)
    s = capi.Solver(case.nx, case.nz, case.nb, case.dt, case.dx, case.fm, case.nt, case.sx, case.sz, case.gx,
                    case.gz)
    s.bind_velocity(vt)
    return s


def encoded(case, seed):
    wlt = capi.ricker(case.nt, case.fm, case.dt, case.amp)
    codes = capi.random_codes(seed, case.ns).astype(np.float32)
    return np.ascontiguousarray(wlt[:, None] * codes[None, :]), wlt


def test_noalloc_forward_matches_allocating_entry_point():
    big = workloads.ess_301x921(nt=600)
    small = workloads.small_case(nx=48, nz=32, nb=8, nt=360, ns=3)
    assert need_bytes(small) < need_bytes(big) <= WORKSPACE_BYTES
    for case, seed in ((big, 1), (small, 2)):
        print(f"noalloc vs forward_modeling on {case.name} nt={case.nt}", flush=True)
        s = solver_for(case)
        enc, wlt = encoded(case, seed)
        ref = s.forward_modeling(enc, shot=-1)
        got = s.forward_modeling_noalloc(enc, shot=-1)
        assert np.array_equal(bits(got), bits(ref)), case.name
        assert np.abs(ref).max() > 0
        shot = case.ns // 2
        ref1 = s.forward_modeling(wlt, shot=shot)
        got1 = s.forward_modeling_noalloc(wlt, shot=shot)
        assert np.array_equal(bits(got1), bits(ref1)), (case.name, shot)
        s.close()


def test_noalloc_forward_refuses_geometry_beyond_workspace():
    # nt*ng alone is 20.5 M floats, beyond the 16 M of the workspace
    case = workloads.small_case(nx=512, nz=40, nb=8, nt=40000, ns=1)
    assert need_bytes(case) > WORKSPACE_BYTES
    print(f"noalloc refusal on {case.name} nt={case.nt}", flush=True)
    s = solver_for(case)
    with pytest.raises(capi.SwfwiError) as e:
        s.forward_modeling_noalloc(np.zeros((case.nt, case.ns), np.float32), shot=-1)
    assert e.value.code == capi.ERR_RANGE and "workspace" in str(e.value)
    s.close()
