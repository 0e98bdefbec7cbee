"""The statistics of the field-parity test on synthetic fields: pure sampling noise passes,
a regional bias of a few percent or a swapped velocity component fails."""
import numpy as np

from field_stats import compare


def accept(r):
    return r["bias_l2"] <= 0.02 and r["l2"] <= 0.05 and r["beyond3"] <= 0.02 \
        and r["beyond5"] <= 0.002


def synthetic(seed, m=2500):
    rng = np.random.default_rng(seed)
    truth = np.stack([1e17 * (1 + rng.random(m)), 1000 + 50 * rng.random(m),
                      -100 + 20 * rng.random(m), 200 + 800 * rng.random(m)], axis=1)
    err = 0.01 * np.abs(truth) + np.array([0.0, 5.0, 5.0, 0.0])
    return truth, err, rng


def test_noise_only_passes():
    truth, err, rng = synthetic(1)
    a = truth + err * rng.standard_normal(truth.shape)
    b = truth + err * rng.standard_normal(truth.shape)
    res = compare(a, err, b, err, np.ones(len(truth), bool))
    assert all(accept(r) for r in res.values()), res
    # the noise explains the difference: little systematic part is left
    assert all(r["bias_l2"] < 0.005 for r in res.values()), res


def test_regional_bias_fails():
    truth, err, rng = synthetic(2)
    a = truth + err * rng.standard_normal(truth.shape)
    b = truth + err * rng.standard_normal(truth.shape)
    a[:400, 3] *= 1.05  # 5% too hot near a wall
    res = compare(a, err, b, err, np.ones(len(truth), bool))
    assert not accept(res["T"]) and accept(res["n"])


def test_swapped_components_fail():
    truth, err, rng = synthetic(3)
    a = truth + err * rng.standard_normal(truth.shape)
    b = truth + err * rng.standard_normal(truth.shape)
    a[:, [1, 2]] = a[:, [2, 1]]
    res = compare(a, err, b, err, np.ones(len(truth), bool))
    assert not accept(res["ux"]) and not accept(res["uy"])
