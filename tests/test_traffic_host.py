"""The HBM-traffic analysis of two rocprofv3 counter passes (host logic only):
calibration scaling and per-class traffic / algorithmic ratios."""
import json

from doublyexcited_amd.traffic import analyse

HDR = "Dispatch_Id,Kernel_Name,Counter_Name,Counter_Value\n"


def _write(d, counter, rows):
    d.mkdir(parents=True, exist_ok=True)
    (d / "run_counter_collection.csv").write_text(
        HDR + "".join(f'{i},"{n}",{counter},{v}\n' for i, n, v in rows))


def test_ratios_from_calibrated_counters(tmp_path):
    gib_kib = (1 << 30) / 1024.0
    # the calibration multiply: FETCH reports half of the bytes read, WRITE all of them
    fetch = [(1, "elementwise_kernel<mul>", gib_kib / 2)]
    write = [(1, "elementwise_kernel<mul>", gib_kib)]
    # two combine launches, each reading 2 GB and writing 1 GB (FETCH at half scale)
    for i in (2, 3):
        fetch.append((i, "void k_combine_w<12, 2, 8>(...)", 2e9 / 1024 / 2))
        write.append((i, "void k_combine_w<12, 2, 8>(...)", 1e9 / 1024))
    # one Gram launch reading 1 GB, and its partial-sum reduction reading 1 MB
    fetch += [(4, "void k_gram_mfma<2, 1, 32>(...)", 1e9 / 1024 / 2), (5, "k_reduce_partials", 1e6 / 1024 / 2)]
    write += [(4, "void k_gram_mfma<2, 1, 32>(...)", 0.0), (5, "k_reduce_partials", 0.0)]
    _write(tmp_path / "f", "FETCH_SIZE", fetch)
    _write(tmp_path / "w", "WRITE_SIZE", write)
    stats = tmp_path / "stats.json"
    # algorithmic: combine 3 GB per launch, Gram 0.9 GB per launch
    stats.write_text(json.dumps({"totals": {"combine": [2, 6e9, 0], "gram": [1, 0.9e9, 0]}}))
    cf, cw, fs, ws, table, ratios = analyse(tmp_path / "f", tmp_path / "w", stats)
    assert abs(fs - 2.0) < 1e-12 and abs(ws - 1.0) < 1e-12
    assert abs(ratios["combine"] - 1.0) < 1e-9
    assert abs(ratios["gram"] - (1e9 + 1e6) / 0.9e9) < 1e-9
    assert {row[0]: row[1] for row in table} == {"combine": 2, "gram": 1}
