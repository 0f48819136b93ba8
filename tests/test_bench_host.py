"""Host-side pieces of the bench's served-stream summary (no GPU): the
mixed-vs-FP64 eigenvalue matching and the latency statistics."""
import numpy as np

import bench
from doublyexcited_amd.serve import Result


def _res(rid, e, res, ok=True):
    return Result(rid, ok, np.asarray(e, complex), np.asarray(res, float), int(np.sum(np.asarray(res) <= 1e-7)),
                  10, 1.0, 1.0)


def test_match_forms_compares_converged_pairs_of_the_same_request():
    res64 = [_res(0, [-0.78 - 0.1j, -0.70 - 0.2j], [1e-8, 1e-3]),   # second FP64 pair not converged
             _res(1, [-0.60 - 0.3j], [1e-9]),
             _res(2, [-0.5], [1e-9], ok=False)]
    resmx = [_res(0, [-0.70 - 0.2j + 2e-9, -0.78 - 0.1j + 3e-9], [1e-8, 1e-8]),
             _res(1, [-0.60 - 0.3j - 1e-9], [1e-3]),                  # mixed pair not converged: not compared
             _res(2, [-0.5], [1e-9])]                                 # failed FP64 request: skipped
    de, unmatched = bench.match_forms(res64, resmx, 1e-7)
    # -0.78-0.1i matches within 3e-9; -0.70-0.2i has no converged FP64 counterpart
    assert len(de) == 1 and abs(de[0] - 3e-9) < 1e-15
    assert unmatched == 1


def test_latency_stats():
    assert bench.latency_stats([]) is None
    st = bench.latency_stats([1.0, 2.0, 3.0, 4.0])
    assert st["n"] == 4 and st["max"] == 4.0 and st["p50"] == 2.5 and 3.9 < st["p99"] <= 4.0
