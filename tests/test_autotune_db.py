"""Host logic of the autotuner: tuning database round trip and selection."""
import json

from gpufft_amd import autotune as A


def test_db_round_trip(tmp_path):
    p = str(tmp_path / "sub" / "tuning.json")
    assert A.load_db(p) == {}
    A.save_db({"k": {"choice": [4], "ms": 1.5}}, p)
    assert A.load_db(p)["k"]["choice"] == [4]
    with open(p, "w") as f:
        f.write("not json")
    assert A.load_db(p) == {}


def test_best_of_and_signature():
    assert A.best_of([((3,), 2.0), ((4,), 1.0), ((5,), 1.0)]) == ((4,), 1.0)
    s1 = A.signature("batched", [(0, 16), (1, 17)])
    assert s1 == A.signature("batched", [(0, 16), (1, 17)])
    assert s1 != A.signature("batched", [(0, 16), (1, 16)])
    json.dumps(s1)
