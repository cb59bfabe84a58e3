"""The P1 contraction of signing, T_j[r] = sum_c P1[r][c] V_j[c], in the three
forms the contraction measurement program times (the bitsliced VALU column
loop, the int8-MFMA GF(2)-plane product and the shared-key LDS tables of
multiples): all give T bit for bit equal to
the program's CPU restatement of the reference's loops (sign.cu:409-418) for
16 signatures of a 65,536-signature, 64-key batch."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_contraction_forms_match_cpu():
    from cumayo_amd import build
    path = build.contraction_path()
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    print("[contraction] running the three forms", flush=True)
    res = subprocess.run([path], capture_output=True, text=True, timeout=120)
    line = res.stdout.strip().splitlines()[-1]
    out = json.loads(line)
    print(f"[contraction] {out}", flush=True)
    assert res.returncode == 0, res.stderr[-2000:]
    assert out["hip_error"] == 0
    assert out["valu_mismatches"] == 0 and out["mfma_mismatches"] == 0 and out["lds_table_mismatches"] == 0
    assert out["valu_ms"] > 0 and out["mfma_ms"] > 0 and out["lds_table_ms"] > 0
