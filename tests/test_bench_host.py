"""Host-side pieces of bench.py, on the CPU: the one-GPU refusal of --gpus N > 1 and the
reading of rocprofv3 counter tables that the traffic and VALU figures of the bench line
come from (a table in rocprofv3's counter_collection.csv column layout)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

HEADER = ('"Correlation_Id","Dispatch_Id","Agent_Id","Queue_Id","Process_Id","Thread_Id",'
          '"Grid_Size","Kernel_Id","Kernel_Name","Workgroup_Size","LDS_Block_Size",'
          '"Scratch_Size","VGPR_Count","Accum_VGPR_Count","SGPR_Count","Counter_Name",'
          '"Counter_Value","Start_Timestamp","End_Timestamp"')

CFG = "idfx::Cfg<3, 3, true, 4, 2, false>"
Y64 = "void idfx::k_sweep_march<%s, double, 1, true, 2, false, false>(idfx::GridDesc)" % CFG
Z64 = "void idfx::k_sweep_march<%s, double, 2, true, 2, false, true>(idfx::GridDesc)" % CFG
Z32 = "void idfx::k_sweep_march<%s, float, 2, true, 2, false, true>(idfx::GridDesc)" % CFG
CT = "void idfx::k_ct3d<%s, double>(idfx::GridDesc)" % CFG


def _row(disp, name, counter, value):
    return ('%d,%d,"Agent 1",1,1,1,512,10,"%s",256,0,0,128,0,48,"%s",%.6f,1,2'
            % (disp, disp, name, counter, value))


def _table(tmp_path, rows):
    path = tmp_path / "run_counter_collection.csv"
    path.write_text("\n".join([HEADER] + rows) + "\n")
    return str(path)


def test_counter_average_per_dispatch(tmp_path):
    rows = [
        _row(1, Z64, "FETCH_SIZE", 100.0), _row(1, Z64, "FETCH_SIZE", 50.0),  # one dispatch
        _row(2, Z64, "FETCH_SIZE", 170.0),
        _row(3, Y64, "FETCH_SIZE", 999.0),   # the y sweep: same template, direction 1
        _row(4, Z32, "FETCH_SIZE", 777.0),   # the FP32 instantiation
        _row(5, Z64, "WRITE_SIZE", 555.0),   # another counter
        _row(6, CT, "FETCH_SIZE", 40.0),
    ]
    path = _table(tmp_path, rows)
    # z: dispatches 1 (100 + 50) and 2 (170)
    assert bench.kernel_counter_average(path, "FETCH_SIZE", "sweep_z") == pytest.approx(160.0)
    assert bench.kernel_counter_average(path, "FETCH_SIZE", "sweep_y") == pytest.approx(999.0)
    assert bench.kernel_counter_average(path, "WRITE_SIZE", "sweep_z") == pytest.approx(555.0)
    assert bench.kernel_counter_average(path, "FETCH_SIZE", "ct_fused") == pytest.approx(40.0)
    assert bench.kernel_counter_average(path, "SQ_INSTS_VALU", "sweep_z") is None
    assert bench.kernel_counter_average(path, "FETCH_SIZE", "sweep_x") is None


def test_region_table_names_each_kernel_once():
    """Each single-kernel region's filter selects exactly one FP64 kernel name."""
    names = [Y64, Z64, CT, "void idfx::k_sweep_x<%s, double, true, 2, false>(x)" % CFG]
    for region in ("sweep_x", "sweep_y", "sweep_z", "ct_fused"):
        prefix, sub = bench.SINGLE_KERNEL_REGIONS[region]
        hits = [n for n in names if prefix in n and sub in n]
        assert len(hits) == 1, (region, hits)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
