"""The profile summariser on synthetic rocprofv3 CSV files whose answers
are known by construction: the closure-interval window and its kernels,
the SQ counter derivations, and the interval timeline."""
import csv

from starpolymersmd_amd import profsum


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def _trace(path, kernels):
    """kernels: (name, start_ns, end_ns, workgroups, workgroup_size, queue)"""
    _write(path, ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X", "Workgroup_Size_X", "Queue_Id"],
           [(n, a, b, g * wg, wg, q) for n, a, b, g, wg, q in kernels])


def test_closure_summary_window_and_grids(tmp_path):
    p = tmp_path / "t.csv"
    _trace(p, [
        ("mlpk::k_train_fused(args)", 0, 60_000, 256, 512, 1),          # before the window
        ("void spmd::k_md<4>(x)", 100_000, 214_000, 256, 512, 1),       # window opens
        ("spmd::k_closure_features(x)", 214_000, 270_000, 256, 1024, 1),
        ("void mlpk::k_forward_fused<1>(x)", 270_000, 325_000, 512, 512, 1),
        ("spmd::k_closure_kick(x)", 325_000, 336_000, 1201, 256, 1),
        ("void spmd::k_rows_redo<2>(x)", 336_000, 350_000, 256, 256, 1),  # an explicit run's kernel: left out
        ("void spmd::k_md<4>(x)", 350_000, 466_000, 256, 512, 1),       # window closes
        ("void spmd::k_md<4>(x)", 500_000, 600_000, 16, 512, 1),        # after the window
    ])
    out = profsum.closure_summary(str(p), stars=256)
    assert "| spmd::k_md<4> | 256 | 2 | 115.00 |" in out        # (114 + 116) / 2
    assert "| spmd::k_closure_features | 256 | 1 | 56.00 |" in out
    assert "| mlpk::k_forward_fused<1> | 512 | 1 | 55.00 |" in out
    assert "| spmd::k_closure_kick | 1201 | 1 | 11.00 |" in out
    window = out.split("Launches of the")[1]
    assert "k_rows_redo" not in window and "k_train_fused" not in window
    assert "| spmd::k_md<4> | 3 | 0.33 | 110.00 |" in out     # whole-run table (total in ms): all three k_md


def test_sq_summary_derivations(tmp_path):
    p1, p2 = tmp_path / "p1.csv", tmp_path / "p2.csv"
    hdr = ["Kernel_Name", "Counter_Name", "Counter_Value"]
    k = "void spmd::k_collide_rows<256, 1024, 64, false>(spmd::DevState, double)"
    _write(p1, hdr, [(k, "SQ_WAVES", 1000), (k, "SQ_WAVE_CYCLES", 800_000), (k, "SQ_ACTIVE_INST_VALU", 80_000),
                     (k, "SQ_LDS_BANK_CONFLICT", 3_000), (k, "SQ_WAVES", 1000)])
    _write(p2, hdr, [(k, "SQ_BUSY_CU_CYCLES", 160_000), (k, "SQ_WAIT_ANY", 440_000),
                     (k, "SQ_ACTIVE_INST_ANY", 160_000), (k, "SQ_INSTS_VALU", 859_000),
                     (k, "SQ_LDS_IDX_ACTIVE", 6_000), (k, "GRBM_GUI_ACTIVE", 20_000)])
    out = profsum.sq_summary([str(p1), str(p2)], ["k_collide_rows"])
    row = [l for l in out.splitlines() if "k_collide_rows" in l][0]
    cells = [c.strip() for c in row.strip("|").split("|")]
    # waves/SIMD 800k/160k; waiting 440k/800k; issuing 160k/800k; VALU 80k/800k;
    # 859k/1000 VALU per wave; conflict 3k/(6k-3k); share 3k/(20k/8)/256
    assert cells[1:] == ["5.0", "55%", "20%", "10%", "859", "1.00", "0.5%"]


def test_timeline_gaps_and_queues(tmp_path):
    p = tmp_path / "t.csv"
    _trace(p, [
        ("void spmd::k_collide_rows<256, 1024, 64, false>(x)", 0, 50_000, 4096, 256, 1),
        ("void spmd::k_md<2>(x)", 55_000, 95_000, 1, 512, 1),
        ("spmd::k_prologue(x)", 60_500, 64_000, 1, 64, 4),
        ("void spmd::k_advect<2>(x)", 64_000, 94_000, 5120, 256, 4),
        ("void spmd::k_collide_rows<256, 1024, 64, false>(x)", 110_000, 160_000, 4096, 256, 1),
    ])
    out = profsum.timeline(str(p), intervals=[1])
    assert "| k_md<2> | 1 | 5.00 | 45.00 | 40.00 |" in out
    assert "| k_prologue | 4 | 10.50 | 14.00 | 3.50 |" in out
    assert "| k_collide_rows<256, 1024, 64, false> | 1 | 60.00 | 110.00 | 50.00 |" in out
