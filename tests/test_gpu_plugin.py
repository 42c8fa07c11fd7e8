"""The HIP GPU plug-in behind the reference SDK: PxCreateBroadPhase with
PxBroadPhaseType::eGPU in the reference SDK built with its GPU support on
(static plug-in mode) reaches PxCreatePhysXGpu -> createGpuBroadPhase, our
Bp::BroadPhase over libpxhip, through the SDK's own ImmGPUBP
(immediatemode/src/NpImmediateMode.cpp:2083-2128). Its created and deleted
pairs must equal those of the same SDK's ABP broadphase (eABP) bit for bit
over incremental updates with statics, kinematics, shared groups, moves,
removals and reused handles."""
import numpy as np
import pytest

import oracle_libs as L
import plugin_scenario

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not L.have_refgpu(), reason="GPU-enabled reference SDK not built (oracle/_ref/gpu)")]


def _set(p):
    return set(map(tuple, np.sort(p, axis=1).tolist()))


def test_sdk_gpu_broadphase_through_plugin_matches_abp(cuda):
    import ctypes
    lib = L.refgpu()
    name = ctypes.create_string_buffer(256)
    assert lib.pxrefgpu_plugin_info(name, 256) == 1
    print("plug-in context manager device:", name.value.decode(), flush=True)
    gpu = L.SdkBroadPhase(L.SdkBroadPhase.GPU)
    abp = L.SdkBroadPhase(L.SdkBroadPhase.ABP)
    total_c = total_d = 0
    for k, (c, u, r, b, g, d) in enumerate(plugin_scenario.updates(lib)):
        gc, gd = gpu.update(c, u, r, b, g, d)
        ac, ad = abp.update(c, u, r, b, g, d)
        print(f"update {k}: created {len(ac)} deleted {len(ad)} (gpu {len(gc)} / {len(gd)})", flush=True)
        assert _set(gc) == _set(ac), k
        assert _set(gd) == _set(ad), k
        assert len(gc) == len(ac) and len(gd) == len(ad)
        total_c += len(ac)
        total_d += len(ad)
    assert total_c > 1000 and total_d > 100, (total_c, total_d)
    gpu.close()
    abp.close()
