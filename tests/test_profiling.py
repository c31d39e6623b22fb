"""The rocprofv3 summary tool on small databases of known counts (no GPU)."""
import os
import sqlite3

from pyop2_amd import profiling


def _counter_db(path, counter, rows):
    """rows: (dispatch id, kernel name, value in KiB), one row per dispatch."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (dispatch_id, kernel_name, counter_name, value)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?)",
                    [(d, k, counter, v) for d, k, v in rows])
    con.commit()
    con.close()


def _runs(tmp_path):
    # two loop launches: each is one assembly dispatch and three residual
    # dispatches (one per colour) of 10, 20 and 30 KiB fetched, 1, 2, 3
    # written; the copy kernel moves 4 KiB each way, which its FETCH_SIZE
    # reads as 2 KiB and its WRITE_SIZE as 4 KiB
    fetch, write = [], []
    d = 0
    for _ in range(2):
        d += 1
        fetch.append((d, "assembly_kernel", 100.0))
        write.append((d, "assembly_kernel", 50.0))
        for kib in (10.0, 20.0, 30.0):
            d += 1
            fetch.append((d, "residual_kernel", kib))
            write.append((d, "residual_kernel", kib / 10))
    d += 1
    fetch.append((d, "copy_kernel", 2.0))
    write.append((d, "copy_kernel", 4.0))
    _counter_db(str(tmp_path / "f" / "pmc.db"), "FETCH_SIZE", fetch)
    _counter_db(str(tmp_path / "w" / "pmc.db"), "WRITE_SIZE", write)
    return [str(tmp_path / "f")], [str(tmp_path / "w")]


def test_traffic_per_dispatch_is_calibrated_on_the_copy_kernel(tmp_path):
    f, w = _runs(tmp_path)
    _, res = profiling.traffic(f, w, "assembly_kernel", "copy_kernel", 4096.0)
    # fetch factor 4096 / 2048 = 2, write factor 1
    assert res == {"assembly_kernel": 100 * 1024 * 2 + 50 * 1024}


def test_traffic_per_launch_sums_a_kernel_dispatched_once_per_colour(tmp_path):
    f, w = _runs(tmp_path)
    _, per_dispatch = profiling.traffic(f, w, "residual_kernel", "copy_kernel", 4096.0)
    lines, per_launch = profiling.traffic(f, w, "residual_kernel", "copy_kernel", 4096.0,
                                          per_sub="assembly_kernel")
    one_launch = sum(k * 1024 * 2 + k / 10 * 1024 for k in (10.0, 20.0, 30.0))
    assert abs(per_launch["residual_kernel"] - one_launch) < 1e-6
    assert abs(per_dispatch["residual_kernel"] - one_launch / 3) < 1e-6
    assert any("per launch: 2 dispatches of assembly_kernel" in ln for ln in lines)
