"""Host-side logic of the learned tuning: training labels from timing
ladders (no GPU)."""
import numpy as np

import tuning


def test_robust_best_takes_centre_of_flat_optimum():
    v = [1, 2, 4, 8, 16, 32, 64]
    # flat within 5% from 8 to 32 -> centre 16
    t = [9.0, 5.0, 2.0, 1.02, 1.0, 1.04, 1.5]
    assert tuning.robust_best(v, t) == 4.0
    # a single clear minimum
    assert tuning.robust_best(v, [5, 4, 3, 2, 1.0, 2, 3]) == 4.0
    assert tuning.robust_best([4], [1.0]) == 2.0


def test_dataset_labels_per_graph_and_k():
    feat = np.tile(np.arange(16, dtype=np.float32), (10, 1))
    kind = np.array([0, 0, 0, 0, 0, 0, 1, 1, 1, 1], np.int8)
    value = np.array([16, 64, 256, 16, 64, 256, 0, 1, 3, 6], np.float64)
    ms = np.array([3.0, 1.0, 2.0, 1.0, 2.0, 3.0, 5.0, 4.0, 4.1, 6.0], np.float64)
    gid = np.array([11, 11, 11, 18, 18, 18, 10, 10, 10, 10], np.int64)
    X, Y = tuning.dataset({"feat": feat, "kind": kind, "value": value, "ms": ms, "gid": gid})
    assert X.shape == (2, 16)
    assert Y[:, 0].tolist() == [6.0, 4.0]  # log2 64 for k=1, log2 16 for k=8
    # SCC schedules 1 and 3 within 5% -> label 1 (lower median of {1, 3}, no log)
    assert Y[:, 1].tolist() == [1.0, 1.0]
    ms[7] = 4.5  # schedule 3 alone fastest
    X, Y = tuning.dataset({"feat": feat, "kind": kind, "value": value, "ms": ms, "gid": gid})
    assert Y[:, 1].tolist() == [3.0, 3.0]


def test_schedule_table_holds_the_default():
    assert tuning.SCHEDULES[tuning.DEFAULT_SCHEDULE] == (1, 3)  # pivot rule 1, 3 trim rounds
    assert len(set(tuning.SCHEDULES)) == len(tuning.SCHEDULES) == 9
