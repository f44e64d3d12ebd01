"""Golden fixtures (tests/golden, written by make_golden.py): committed
inputs and the SHA-256 of the level-0 offsets, canvas and residual.  The CPU
test pins the oracle against later changes; the GPU test checks the HIP path
reproduces the same bytes."""
import json
import os

import pytest

from tests.golden.make_golden import CASES, HERE, digests, load_inputs, run_oracle

GOLDEN = json.load(open(os.path.join(HERE, "golden.json")))


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(name):
    style, frames = load_inputs(name)
    assert digests(run_oracle(CASES[name], style, frames)) == GOLDEN[name]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_reproduces_golden(name):
    from painttween_amd import Context, _ffi
    case = CASES[name]
    style, frames = load_inputs(name)
    with Context.gpu(0) as g:
        p = g.default_params(**case["params"])
        outs = g.synthesize(p, frames, style, scheme=case.get("scheme", _ffi.PT_S_FMBM_RM),
                            iterations=case["iterations"], fixup_passes=1)
    assert digests(outs) == GOLDEN[name]
