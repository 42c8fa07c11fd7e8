"""CPU-side checks of the product boundary: the HIP library loads and exports
every entry point include/pxhip.h declares (no device calls)."""
import os
import re

import physx_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "pxhip.h")).read()
    return sorted(set(re.findall(r"\b(pxh_[a-z_0-9]+)\s*\(", text)))


def test_header_matches_exports_list():
    assert declared() == sorted(P.EXPORTED)


def test_library_exports_every_symbol():
    lib = P.load()
    for name in declared():
        assert hasattr(lib, name), name


def test_default_desc_matches_reference_defaults():
    d = P.default_desc()
    assert d.solver_type == P.PXH_SOLVER_TGS
    assert abs(d.contact_offset - 0.02) < 1e-9            # 0.02 * PxTolerancesScale::length
    assert abs(d.bounce_threshold_velocity - 2.0) < 1e-9  # 0.2 * scale.speed (PxSceneDesc.h:1044)
    assert abs(d.friction_offset_threshold - 0.04) < 1e-9
    assert abs(d.angular_damping - 0.05) < 1e-9           # ScBodyCore.cpp:54
    assert abs(d.max_angular_velocity_sq - 1e4) < 1e-3


def test_gpu_plugin_exports_and_sdk_harness():
    """The reference SDK built with its GPU support on links the HIP GPU
    plug-in: the plug-in's exports (physxgpu/include/PxPhysXGpu.h:203-241) and
    the public GPU entry points of the static plug-in mode (include/gpu/
    PxGpu.h:70-144) are there, and the same SDK's CPU broadphase runs the
    incremental scenario of the GPU test (no device calls)."""
    import pytest
    import oracle_libs as L
    import plugin_scenario
    if not L.have_refgpu():
        pytest.skip("GPU-enabled reference SDK not built (oracle/_ref/gpu)")
    lib = L.refgpu()
    for name in ("PxCreatePhysXGpu", "PxCreateCudaContextManager", "PxSetPhysXGpuProfilerCallback",
                 "PxSetPhysXGpuFoundationInstance", "PxGetSuggestedCudaDeviceOrdinal", "PxGpuCudaRegisterFunction",
                 "PxGpuCudaRegisterFatBinary", "PxGpuGetCudaFunctionTable", "PxGpuGetCudaFunctionTableSize",
                 "PxGpuGetCudaModuleTable", "PxGpuGetCudaModuleTableSize", "PxGpuCreatePhysicsGpu",
                 "PxCudaRegisterFunction", "PxCudaRegisterFatBinary", "PxGetCudaModuleTable",
                 "PxGetCudaModuleTableSize", "PxGetCudaFunctionTable", "PxGetCudaFunctionTableSize",
                 "PxGetPhysicsGpu"):
        assert hasattr(lib, name), name
    abp = L.SdkBroadPhase(L.SdkBroadPhase.ABP)
    n = 0
    for c, u, r, b, g, d in plugin_scenario.updates(lib, cap=2000, n0=1500, rounds=3):
        cr, dl = abp.update(c, u, r, b, g, d)
        n += len(cr)
    assert n > 100
    abp.close()
