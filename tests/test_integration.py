"""The drop-in, built: the reference solver's own sources compiled against
this engine through the reference-side adapter
(optimized-gpu-accelerator_amd/integration/pdhg_hip_adapter.cpp), as
oracle/_ref/libpramana_hip.so (its C API) and oracle/_ref/pramana_hip (its
CLI).  CPU checks: the adapter compiles against the reference's headers
(src/pdhg/pdhg.h:73-80), the reference's C API reports this engine's device
information, and without a device the reference's solve() keeps its own CPU
PDHG for PDHG requests, as its backend choice does (src/pdhg/pdhg.cpp:491-500)."""
import pytest

import integration_lib as I
import pdhg_amd as P

pytestmark = pytest.mark.skipif(not I.available(), reason="integration build absent (needs the reference tree)")


def test_adapter_serves_the_reference_c_api_device_info():
    info = I.gpu_info()  # pramana_gpu_info (src/api/c_api.cpp:281) -> gpuInfoJson -> pdhg_gpu_info_json
    for key in ("available", "arch", "compute_units", "memory_bytes", "peak_bandwidth_GBs"):
        assert key in info, key
    assert "nvrtc" not in info and "compute_capability" not in info  # not the CUDA driver's record
    assert info["available"] == P.gpu_available()


def test_reference_pipeline_without_device_uses_reference_cpu_pdhg():
    if P.gpu_available():
        pytest.skip("a GPU is present: the GPU integration tests cover this path")
    opt = I.netlib_optima()["afiro"]
    st, obj, cert, j = I.solve(I.netlib_path("afiro"), algorithm="pdhg-gpu")
    assert st == 1 and cert == 1
    assert j["engine"] == "pdhg-cpu+crossover"
    assert j["telemetry"]["pdhg"]["backend"] == "cpu"
    assert abs(obj - opt) <= 1e-8 * (1 + abs(opt))


def test_reference_cli_runs_over_the_adapter():
    rc, out, err = I.cli("info")
    assert rc == 0 and '"compute_units"' in out, (rc, out, err)
