"""The HIP backend against the reference's own outputs (GPU).

tests/golden holds runs of the reference's fca_step loop (built from its
sources, contraction off) with their inputs. The fused and the split kernels
must reproduce them: FP32 bit for bit (h, hmax, tmax, t and dt_base per step),
FP64 within 1e-12 m; balance terms to 1e-9 relative; cap counts exactly.
"""
import numpy as np
import pytest

from conftest import GOLDEN_CASES, load_golden
from floodca_amd import abi as A
from floodca_amd.hip import Sim
from oracle import Oracle


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [A.FCA_HIP_FUSED, A.FCA_HIP_SPLIT], ids=["fused", "split"])
@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_hip_reproduces_reference(name, prec, kernel):
    a, g, cfg, exp = load_golden(name, prec)
    print(f"golden {name} {prec} kernel={kernel}", flush=True)
    acc = Oracle(prec).acc_init(g, cfg)
    with Sim(g, cfg, acc, prec=prec, kernel=kernel, strip_rows=16, record_steps=100000) as s:
        s.run()
        out = s.read()
        log = s.steplog(100000)
    log_ref = exp["log"]
    assert out.step == exp["acc"]["step"] == len(log) == len(log_ref)
    if prec == "f32":
        for k in ("h", "hmax", "tmax"):
            assert np.array_equal(getattr(g, k), exp[k]), k
        assert [r.t for r in log] == list(log_ref[:, 0])
        assert np.array_equal(np.array([r.dt_base for r in log], np.float32),
                              log_ref[:, 2].astype(np.float32))
        assert out.cap_active == exp["acc"]["cap_active"]
    else:
        for k in ("h", "hmax"):
            assert np.abs(getattr(g, k) - exp[k]).max() <= 1e-12, k
    e = exp["acc"]
    assert out.cap_total == e["cap_total"] and out.dt_min_hits == e["dt_min_hits"]
    assert out.t == pytest.approx(e["t"], rel=1e-15)
    for k in ("V", "V_rain", "V_sea", "V_bnd", "clamp_err"):
        assert getattr(out, k) == pytest.approx(e[k], rel=1e-9, abs=1e-9 * max(1.0, abs(e[k])))
