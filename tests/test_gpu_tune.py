"""GPU checks of the dslash launch-shape autotuner (qcuDslashAutotune): the
built-in default is timed first and stays unless a candidate is faster, the
candidates include one site per wavefront, the tuned shape reproduces the
oracle and is no slower than the default when timed again, and the tuning
table survives a save / load round trip."""
import numpy as np
import pytest

from oracle import oracle
from qcu_mrhs_amd import api, fields

pytestmark = pytest.mark.gpu

D, S, H, B = 2, 1, 0, 4
TOL = {D: 1e-13, S: 2e-7, H: 2e-3, B: 1.5e-2}


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(autouse=True)
def _reset():
    yield
    api.dslash_tune("auto")


def _time(fn, reps=20):
    L = api.lib()
    fn()
    api.synchronize()
    best = 1e30
    for _ in range(3):
        api.check(L.qcuTimerStart())
        for _ in range(reps):
            fn()
        best = min(best, L.qcuTimerStopMs() / reps)
    return best * 1e3


@pytest.mark.parametrize("prec", [H, D])
def test_autotuned_shape_matches_oracle_and_is_not_slower(prec, tmp_path):
    L, Nc, R = (16, 16, 16, 32), 3, 16
    print("case autotune", prec, flush=True)
    best, best_us, recs = api.dslash_autotune(L, Nc, R, prec, repeats=5)
    default = recs[0]
    print(f"precision {prec}: default {default}, tuned {best} at {best_us:.1f} us over {len(recs)} candidates")
    # a shape other than the default was faster in the sweep and when re-timed against it
    is_default = all(best[k] == default[k] for k in ("variant", "by", "bz", "spw"))
    assert is_default or best_us < default["us"] * 1.05
    if prec == H:
        assert any(r["variant"] == 3 and r["spw"] == 1 for r in recs)
    # the tuned shape is what automatic launches use now; it reproduces the oracle
    gauge = fields.gauge_field(L, Nc, 31, kind="smooth", eps=0.3)
    src = fields.gaussian_sources(L, Nc, R, 32).reshape(R, 2, -1)
    inp = fields.to_mrhs(src[:, 0], Nc)
    d_g = api.DeviceArray.from_complex(gauge.reshape(-1), prec)
    d_in = api.DeviceArray.from_complex(inp, prec)
    d_out = api.DeviceArray(api.storage_bytes(inp.size, prec))
    api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, 1, 0, precision=prec)
    err = _rel(d_out.complex(inp.size, prec), oracle.dslash(inp, gauge, L, Nc, R, 1))
    assert err < TOL[prec]

    def run():
        api.dslash_mrhs(d_out, d_in, d_g, L, Nc, R, 1, 0, precision=prec)

    t_tuned = _time(run)
    variant = default["variant"]
    api.dslash_tune(variant, default["by"], default["bz"], default["spw"])
    t_default = _time(run)
    api.dslash_tune("auto")
    print(f"  re-timed: tuned {t_tuned:.1f} us, default {t_default:.1f} us")
    assert t_tuned <= 1.05 * t_default

    path = tmp_path / "tune.txt"
    api.dslash_tune_save(path)
    rows = [ln.split() for ln in open(path) if not ln.startswith("#")]
    key = [str(v) for v in (*L, Nc, R, prec)]
    mine = [r for r in rows if r[:7] == key]
    assert mine and [int(v) for v in mine[0][7:]] == [best["variant"], best["by"], best["bz"], best["spw"]]
    assert api.dslash_tune_load(path) == len(rows)
    bad = tmp_path / "bad.txt"
    bad.write_text("1 2 3\n")
    with pytest.raises(api.QcuError):
        api.dslash_tune_load(bad)
