"""bench.py's measurement arithmetic on the CPU: algorithmic bytes and flops
per hop launch, the roofline of the class with the largest estimated total
time (active-RHS weighting of the fused CG kernels), the committed HBM
traffic summary, and the refusal of more than one GPU."""

import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

NX, NT = 288, 64


def test_kernel_bytes_flops():
    by, fl, prec = bench.kernel_bytes_flops("hop_M_f64_cgb", 128, NX, NT)
    assert prec == "f64"
    assert by == 6 * NT * NX * 16 * 128            # x, r, p, u read/written, c128
    assert fl == 4.0 * NX * NX * NT * 128           # real E x complex block
    by, fl, prec = bench.kernel_bytes_flops("hop_Mdag_bf16x3_cga", 128, NX, NT)
    assert prec == "bf16x3"
    assert by == 4 * NT * NX * 8 * 128              # c64 vectors
    assert fl == 3 * 4.0 * NX * NX * NT * 128       # three bf16 products


def test_roofline_picks_largest_total_and_weights_active_rhs():
    # name launches total_ms timed: the f64 CGB class has the largest estimate
    text = "\n".join([
        "hop_M_f64_cgb 1000 3.2 50",        # 64 us average, 1000 launches
        "hop_Mdag_f64_cga 1000 3.0 50",     # 60 us
        "hop_M_bf16x3_cgb 100 0.34 10",     # 34 us, 100 launches
    ])
    rep = {"rhs_iters_f64": 64000, "rhs_iters_low": 12800}
    rf, rows = bench.roofline_from_prof(text, 128, NX, NT, rep, traffic=123)
    assert len(rows) == 3
    assert rf["kernel"] == "hop_M_f64_cgb" and rf["bound"] == "mfma"
    assert rf["avg_us"] == pytest.approx(64.0)
    active = 64000 / 1000                         # active RHS per launch
    flops = 4.0 * NX * NX * NT * active
    assert rf["algorithmic_flops"] == pytest.approx(flops)
    assert rf["achieved"] == pytest.approx(flops / 64e-6 / 1e12, rel=1e-3)
    assert rf["frac"] == pytest.approx(rf["achieved"] / 78.6, rel=1e-3)
    assert rf["traffic"] == 123 and rf["timed_launches"] == 50


def test_roofline_of_reported_work_classes():
    """Classes other than the hop report their own flops and bytes (summed
    over the timed launches): the batched GEMM is priced on the matrix cores
    of its precision, the double-double hop on the FP64 vector units, the
    Gauss-Jordan helpers on HBM."""
    text = "\n".join([
        "bgemm_f32_cplx_ph0_64 100 10.0 50 5.0e11 1.0e10",   # 200 us, 1e10 flops per launch
        "dd_hop_m_resid 4 3.2 4 8.0e10 4.0e8",               # 800 us
        "gj_panel_f64 10 1.0 5 1e6 5.0e7",                   # 200 us
        "chunkprod_f64 2 1.0 1 3.0e10 1.0e7",                # 1 ms, 3e10 flops
        "chain_f32_ph1 6 0.6 3 3.0e9 1.0e8",                 # 200 us, 1e9 flops per launch
    ])
    rf, rows = bench.roofline_from_prof(text, 128, NX, NT, {})
    assert rf["kernel"] == "bgemm_f32_cplx_ph0_64"
    assert rf["bound"] == "mfma" and rf["peak"] == 157.3
    assert rf["achieved"] == pytest.approx(1e10 / 200e-6 / 1e12, rel=1e-3)
    table = {d["kernel"]: d for d in bench.kernel_table(rows)}
    assert table["dd_hop_m_resid"]["bound"] == "valu"
    assert table["dd_hop_m_resid"]["achieved"] == pytest.approx(2e10 / 800e-6 / 1e12, rel=1e-3)
    assert table["chunkprod_f64"]["bound"] == "mfma"
    assert table["chunkprod_f64"]["achieved"] == pytest.approx(3e10 / 1e-3 / 1e12, rel=1e-3)
    assert table["chunkprod_f64"]["frac"] == pytest.approx(30.0 / 78.6, rel=1e-3)
    assert table["chain_f32_ph1"]["bound"] == "mfma"
    assert table["chain_f32_ph1"]["frac"] == pytest.approx(5.0 / 157.3, rel=1e-3)
    assert table["gj_panel_f64"]["bound"] == "hbm"
    assert table["gj_panel_f64"]["achieved"] == pytest.approx(1e7 / 200e-6 / 1e9, rel=1e-3)


def test_bf16_roofline_is_hbm_bound():
    rf, _ = bench.roofline_from_prof("hop_M_bf16x3_plain 10 0.25 10", 128, NX, NT, {})
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s"
    assert rf["achieved"] == pytest.approx(2 * NT * NX * 8 * 128 / 25e-6 / 1e9, rel=1e-3)


def test_committed_traffic_covers_the_cg_kernels():
    for k in ("hop_M_f64_cgb", "hop_Mdag_f64_cga"):
        t = bench.read_traffic(k)
        by, _, _ = bench.kernel_bytes_flops(k, 117.7, NX, NT)
        # measured HBM bytes at least the algorithmic ones, and not wildly above
        assert t is not None and by <= t <= 2 * by


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode != 0
    assert "one GPU" in p.stdout + p.stderr
