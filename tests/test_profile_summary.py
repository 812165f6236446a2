"""The profile summarizer on synthetic rocprofv3 CSVs: only dispatches inside
the window count, repeated dispatches of one kernel are labelled by their
order, counters are averaged per dispatch and HBM bytes are scaled by the copy
calibration."""
import csv

from sphinxsys_amd import profile_summary as ps

P2 = "void sphx::(anonymous namespace)::k_pass2<3, 1, true, true, false>(sphx::(anonymous namespace)::FluidArgs<float>, double, unsigned long long*)"
TILE = "void sphx::(anonymous namespace)::k_relation_tile<3>(sphx::(anonymous namespace)::BuildArgs)"
COPY = "__amd_rocclr_copyBuffer"


def write_csv(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_window_selection_labels_and_traffic(tmp_path):
    (tmp_path / "trace").mkdir()
    (tmp_path / "pmc").mkdir()
    # dispatch id, kernel, start, end (ns): two outside the window (before, after)
    disp = [(1, P2, 500, 1500), (2, P2, 2000, 3000), (3, TILE, 3000, 6000), (4, TILE, 6000, 7000),
            (5, P2, 7000, 8000), (6, TILE, 8000, 11000), (7, TILE, 11000, 12000), (8, P2, 30000, 31000),
            (9, COPY, 40000, 41000)]
    write_csv(tmp_path / "trace" / "run_kernel_trace.csv",
              ["Dispatch_Id", "Kernel_Name", "Start_Timestamp", "End_Timestamp"],
              [(d, k, s, e) for d, k, s, e in disp])
    rows = []
    for d, k, s, e in disp:
        if k == COPY:
            rows += [(d, k, "FETCH_SIZE", 524288.0, s, e), (d, k, "WRITE_SIZE", 1048576.0, s, e)]
        else:
            rows += [(d, k, "SQ_WAVES", 100.0, s, e), (d, k, "SQ_INSTS_VALU", 100.0 * d, s, e),
                     (d, k, "FETCH_SIZE", 1000.0, s, e), (d, k, "WRITE_SIZE", 500.0, s, e)]
    write_csv(tmp_path / "pmc" / "run_counter_collection.csv",
              ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"],
              rows)
    (tmp_path / "bench.err").write_text("x\nprofile-window headline 1800 20000\nprofile-window copy_calibration 39000 42000\n")
    out = tmp_path / "s.md"
    log = str(tmp_path / "bench.err")
    ps.main(["--trace", f"{tmp_path / 'trace'}={log}", "--pmc", f"{tmp_path / 'pmc'}={log}",
             "--alternate", "k_relation_tile=inner,contact", "--out", str(out)])
    text = out.read_text()
    # 6 dispatches in the window: pass 2 twice (1 us each), the tile build as inner (3 us) and contact (1 us)
    # busy: [2000, 12000] covered without gaps = 10 us of an 18.2 us window
    assert "(6 dispatches); the GPU runs at least one kernel 0.0 ms (54.9% of the window)" in text
    assert "| `k_pass2<3, 1, true, true, false>` | 2 | 1.0 |" in text
    assert "| `k_relation_tile<3> [inner]` | 2 | 3.0 |" in text
    assert "| `k_relation_tile<3> [contact]` | 2 | 1.0 |" in text
    # pass 2: VALU per wave = mean(200, 500) / 100; bytes = 1000 KiB x 2048 + 500 KiB x 1024
    assert "| `k_pass2<3, 1, true, true, false>` | 1.0 | 4 |" in text
    assert "traffic acoustic_pass2 2560000" in text
    assert "traffic acoustic_pass1" not in text


def test_short_names():
    assert ps.short_name(P2) == "k_pass2<3, 1, true, true, false>"
    assert ps.short_name("sphx::(anonymous namespace)::k_cll_clear(sphx::(anonymous namespace)::CllScene)") == \
        "k_cll_clear"
    assert ps.short_name(COPY) == COPY
