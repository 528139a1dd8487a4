"""Register budget of the solve kernels, checked at compile time (hipcc cross-compiles gfx950 without
a GPU). The kernels run one 512-thread workgroup per CU with all 256 VGPRs; a change that pushes
more live values into spill scratch slows every PCG iteration (an extra 160 B/lane of spills once
cost 14% of the 192x192 solve), so the per-lane scratch of each kernel is held to its budget."""
import os
import re
import shutil
import subprocess

import pytest

from ding_nsch_core_solver_amd import build as B

# bytes of spill scratch per lane allowed for each kernel family
BUDGET = {"fp64_solve_kernel": 32, "mixed_solve_kernel": 96, "correction_faces_kernel": 0,
          "correction_history_kernel": 0, "correction_finish_kernel": 0, "pressure_ghost_kernel": 0,
          "bicg_": 0, "momentum_": 0, "ch_pcg_kernel": 0, "ch_factor_seq_kernel": 0, "ch_diag_": 0}


def _usage(source, tmp_path):
    try:
        hipcc = B.hipcc()
    except FileNotFoundError:
        pytest.skip("hipcc not available")
    flags = [f for f in B.HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    cmd = [hipcc, *flags, f"-I{B.INCLUDE}", f"-I{B.CSRC}", "-c", os.path.join(B.CSRC, source),
           "-o", str(tmp_path / "k.o"), "--offload-device-only", "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    usage, name = {}, None
    for line in out.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and name:
            usage[name] = int(m.group(1))
    return usage


@pytest.mark.parametrize("source", ["pp_fp64.hip", "pp_correct.hip", "pp_bicgstab.hip", "pp_momentum.hip", "pp_ch.hip"])
def test_kernel_spill_scratch_within_budget(tmp_path, source):
    if shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("hipcc not available")
    usage = _usage(source, tmp_path)
    checked = 0
    for mangled, scratch in usage.items():
        for family, budget in BUDGET.items():
            if family in mangled:
                assert scratch <= budget, f"{mangled}: {scratch} B/lane of scratch (budget {budget})"
                checked += 1
    assert checked > 0, f"no kernel of {source} matched the budget table: {sorted(usage)}"
