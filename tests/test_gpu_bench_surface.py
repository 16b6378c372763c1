"""bench.py --surface runs the reference's benchmark axes through the drop-in
entry points and emits one JSON line per point (smallest sizes here)."""
import argparse
import io
import json
import contextlib

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_surface_lines(dtype):
    import bench
    args = argparse.Namespace(surface_logn="12-13", surface_dtype=dtype, steps=3, warmup=1)
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        bench.surface(args)
    lines = [json.loads(x) for x in buf.getvalue().splitlines() if x.strip()]
    assert len(lines) == 2 * len(bench.SURFACE)
    for ln in lines:
        assert ln["bench"] in bench.SURFACE and ln["ms"] > 0 and ln["dtype"] == dtype
