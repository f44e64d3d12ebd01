"""Writes the golden fixtures of the oracle (run from the repository root):

    python tests/golden/make_golden.py

Each case stores its inputs and the SHA-256 of the oracle's level-0 offsets,
canvas and residual.  The reference ships no golden data; these fixtures pin
the oracle (and through the parity tests the HIP path) against later changes
of this repository, not against the reference.
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from painttween_amd import _ffi  # noqa: E402
from painttween_amd import workloads as W  # noqa: E402
from tests.oracle_ctx import OracleContext  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "still_r2_noguides_32": dict(kind="still", size=(32, 32), style=(32, 32), guides=False, seed=101,
                               params=dict(residual_window_size=2, levels=2, orientation_enabled=0), iterations=3),
    "still_r3_guides_48x40": dict(kind="still", size=(48, 40), style=(40, 48), guides=True, seed=102,
                                  params=dict(residual_window_size=3, levels=3, dist_trans_weight=0.5), iterations=2),
    "still_counter_rng_40": dict(kind="still", size=(40, 40), style=(64, 64), guides=False, seed=103,
                                 params=dict(residual_window_size=2, levels=2, rng_mode=_ffi.PT_RNG_COUNTER, seed=7),
                                 iterations=3),
    "anim_fmbm_rm_3f_40x32": dict(kind="anim", size=(40, 32), style=(48, 48), frames=3, seed=104,
                                  params=dict(residual_window_size=2, levels=2, temporal_coherence_weight=0.5),
                                  iterations=2, scheme=_ffi.PT_S_FMBM_RM),
    "anim_fmbm_3f_40x32": dict(kind="anim", size=(40, 32), style=(48, 48), frames=3, seed=105,
                               params=dict(residual_window_size=2, levels=2, hysteresis_weight=0.2),
                               iterations=2, scheme=_ffi.PT_S_FMBM),
}


def inputs(case):
    w, h = case["size"]
    sw, sh = case["style"]
    if case["kind"] == "still":
        style = W.style_pair(sw, sh, seed=case["seed"], guides=case["guides"])
        frames = [W.still_frame(w, h, seed=case["seed"], guides=case["guides"])]
    else:
        style = W.style_pair(sw, sh, seed=case["seed"], guides=True)
        frames = W.animation(w, h, case["frames"], seed=case["seed"])
    return style, frames


def save_inputs(name, style, frames):
    arrays = {"style_" + k: v for k, v in style.items()}
    for i, f in enumerate(frames):
        for k, v in f.items():
            arrays["frame%d_%s" % (i, k)] = v
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)


def load_inputs(name):
    z = np.load(os.path.join(HERE, name + ".npz"))
    style = {k[6:]: z[k] for k in z.files if k.startswith("style_")}
    n = 1 + max(int(k.split("_")[0][5:]) for k in z.files if k.startswith("frame"))
    frames = []
    for i in range(n):
        p = "frame%d_" % i
        frames.append({k[len(p):]: z[k] for k in z.files if k.startswith(p)})
    return style, frames


def run_oracle(case, style, frames):
    with OracleContext() as o:
        p = o.default_params(**case["params"])
        return o.synthesize(p, frames, style, scheme=case.get("scheme", _ffi.PT_S_FMBM_RM),
                            iterations=case["iterations"], fixup_passes=1)


def digests(outs):
    return {k: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
            for k, a in zip(("offsets", "canvas", "residual"), outs)}


def main():
    table = {}
    for name, case in CASES.items():
        style, frames = inputs(case)
        save_inputs(name, style, frames)
        style, frames = load_inputs(name)
        table[name] = digests(run_oracle(case, style, frames))
        print(name, table[name]["offsets"][:16])
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
