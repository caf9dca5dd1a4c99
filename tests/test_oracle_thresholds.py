"""The CPU restatement of the per-voxel threshold optimisation, pinned by a
separable case with a known plateau and by numpy confusion counts with the
textbook score definitions."""
import numpy as np


def test_separable_voxel_picks_plateau_centre(oracle):
    rng = np.random.default_rng(0)
    pos = rng.uniform(0.601, 1.0, 40)
    neg = rng.uniform(0.0, 0.399, 60)
    probas = np.concatenate([pos, neg]).astype(np.float32)
    labels = np.concatenate([np.ones(40), np.zeros(60)]).astype(np.float32)
    thr, sc = oracle.optimal_thresholds(labels, probas, 1, 101, metric=2)
    assert np.all(sc[0, 40:61] == 1.0) and sc[0, 39] < 1.0 and sc[0, 61] < 1.0
    assert thr[0] == np.float32(50 / 100)


def test_scores_match_numpy_confusion_counts(oracle):
    rng = np.random.default_rng(1)
    V, S, T = 20, 200, 101
    probas = rng.random((V, S)).astype(np.float32)
    labels = (rng.random((V, S)) < probas * 0.8 + 0.1).astype(np.float32)
    t = (np.arange(T, dtype=np.float32) / np.float32(T - 1))
    pred = probas[:, None, :] > t[None, :, None]          # [V][T][S]
    truth = labels[:, None, :] == 1.0
    tp = (pred & truth).sum(-1).astype(np.float64)
    fp = (pred & ~truth).sum(-1).astype(np.float64)
    fn = (~pred & truth).sum(-1).astype(np.float64)
    tn = (~pred & ~truth).sum(-1).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        prec = np.where(tp > 0, tp / (tp + fp), 0.0)
        rec = np.where(tp > 0, tp / (tp + fn), 0.0)
        expect = {
            0: prec,
            1: rec,
            2: np.where(prec + rec > 0, 2 * prec * rec / (prec + rec), 0.0),
            6: np.where((tp + fp) * (tp + fn) * (tn + fp) * (tn + fn) > 0,
                        (tp * tn - fp * fn) / np.sqrt((tp + fp) * (tp + fn) * (tn + fp) * (tn + fn)), 0.0),
            8: (tp + tn) / (tp + tn + fp + fn),
            9: 0.5 * (np.where(tp > 0, tp / (tp + fn), 0) + np.where(tn > 0, tn / (tn + fp), 0)),
            11: np.where(tp > 0, tp / (tp + fn), 0) + np.where(tn > 0, tn / (tn + fp), 0) - 1,
        }
    for metric, e in expect.items():
        thr, sc = oracle.optimal_thresholds(labels, probas, V, T, metric=metric)
        np.testing.assert_allclose(sc, e, rtol=2e-6, atol=2e-7, err_msg=f"metric {metric}")
        best = np.argmax(sc, axis=1)
        assert np.all(np.abs(thr - best / (T - 1)) <= 0.5 + 1e-6)  # within the plateau after the maximum
        assert np.all(sc[np.arange(V), np.round(thr * (T - 1)).astype(int)] >= sc.max(axis=1) - 1e-4)
