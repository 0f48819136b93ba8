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
    ns = "void dex::(anonymous namespace)::"
    # torch's fill, then the calibration multiply: FETCH reports half of the bytes read, WRITE all
    # of them (the fill writes as much and reads nothing)
    fill = "void at::native::vectorized_elementwise_kernel<4, at::native::FillFunctor<c10::complex<double> > >(int)"
    mul = "void at::native::vectorized_elementwise_kernel<4, at::native::AUnaryFunctor<c10::complex<double> > >(int)"
    fetch = [(1, fill, 20.0), (2, mul, gib_kib / 2)]
    write = [(1, fill, gib_kib), (2, mul, gib_kib)]
    # a library kernel outside the classes with a larger count must not be taken for the calibration
    fetch.append((3, "dex::(anonymous namespace)::k_singlet(HIP_vector_type<double, 2u>*, long, double)", 4 * gib_kib))
    write.append((3, "dex::(anonymous namespace)::k_singlet(HIP_vector_type<double, 2u>*, long, double)", 2 * gib_kib))
    # two combine launches, each reading 2 GB and writing 1 GB (FETCH at half scale)
    for i in (4, 5):
        fetch.append((i, ns + "k_combine_w<12, 2, 8>(dex::ColDesc const*, int, int)", 2e9 / 1024 / 2))
        write.append((i, ns + "k_combine_w<12, 2, 8>(dex::ColDesc const*, int, int)", 1e9 / 1024))
    # one Gram launch reading 1 GB, and its partial-sum reduction reading 1 MB
    fetch += [(6, ns + "k_gram_mfma<2, 1, 32>(dex::ColDesc const*, int)", 1e9 / 1024 / 2),
              (7, "dex::(anonymous namespace)::k_reduce_partials(HIP_vector_type<double, 2u> const*, int)",
               1e6 / 1024 / 2)]
    write += [(6, ns + "k_gram_mfma<2, 1, 32>(dex::ColDesc const*, int)", 0.0),
              (7, "dex::(anonymous namespace)::k_reduce_partials(HIP_vector_type<double, 2u> const*, int)", 0.0)]
    # a column-dot launch and a residual step (both coldots launches), each reading 1 GB and writing
    # 0.5 GB with its reduction
    for i, k in ((8, "k_coldots(dex::ColDesc const*, int, long, double*)"),
                 (9, "k_resid(dex::ColDesc const*, int, double const*)")):
        fetch.append((i, "dex::(anonymous namespace)::" + k, 1e9 / 1024 / 2))
        write.append((i, "dex::(anonymous namespace)::" + k, 0.5e9 / 1024))
    _write(tmp_path / "f", "FETCH_SIZE", fetch)
    _write(tmp_path / "w", "WRITE_SIZE", write)
    stats = tmp_path / "stats.json"
    # algorithmic: combine 3 GB per launch, Gram 0.9 GB per launch, coldots 1.5 GB per launch
    stats.write_text(json.dumps({"totals": {"combine": [2, 6e9, 0], "gram": [1, 0.9e9, 0],
                                            "coldots": [2, 3e9, 0]}}))
    cf, cw, fs, ws, table, ratios = analyse(tmp_path / "f", tmp_path / "w", stats)
    assert abs(fs - 2.0) < 1e-12 and abs(ws - 1.0) < 1e-12
    assert abs(ratios["combine"] - 1.0) < 1e-9
    assert abs(ratios["gram"] - (1e9 + 1e6) / 0.9e9) < 1e-9
    assert abs(ratios["coldots"] - 1.0) < 1e-9
    assert {row[0]: row[1] for row in table} == {"combine": 2, "gram": 1, "coldots": 2}
