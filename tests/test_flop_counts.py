"""The bench's algorithmic work figure for the TI kernels is the reference's
own tregn96 operation count per layer (oracle/count_tregn96_flops.py, a
static count over its Fortran source): the constant in bench.py must be what
the count gives for the reference at hand."""
import os
import re

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_ti_ops_per_layer_is_the_tregn96_count():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "count_tregn96_flops", os.path.join(ROOT, "oracle", "count_tregn96_flops.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    if not os.path.exists(m.SRC):
        pytest.skip("reference source not present")
    tot, memo = m.per_layer()
    with open(os.path.join(ROOT, "bench.py")) as fh:
        const = int(re.search(r"^TI_OPS_PER_LAYER = (\d+)", fh.read(), re.M).group(1))
    assert round(tot.total()) == const
    # the compound-matrix step dominates, as the reference's term-by-term
    # complex expansion of the 6x6 compound suggests
    assert memo["dnka_tregn"].total() > 0.5 * tot.total()
