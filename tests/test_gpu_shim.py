"""The reference-signature C++ shim (delta_amd/csrc/shim/delta_shim.cpp),
compiled against the reference's own headers into oracle/_ref/shim_test, run
as Delta's engine would call it; its contacts must equal the oracle's bit for
bit.  Skipped where the driver was not built (no reference tree)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
import scenes

pytestmark = pytest.mark.gpu
BIN = os.path.join(ol.ROOT, "oracle", "_ref", "shim_test")


@pytest.mark.parametrize("method", [0, 1, 2, 7])
def test_shim_matches_oracle(method):
    if not os.path.exists(BIN):
        pytest.skip("oracle/_ref/shim_test not built (needs the reference headers)")
    name, A, B = scenes.pair_cases()[0]
    def blocks(m):
        return np.concatenate([m[:, :, ax].reshape(-1) for ax in range(3)])
    txt = f"{len(A)} {len(B)}\n" + " ".join(repr(float(v)) for v in blocks(A)) + "\n" + \
        " ".join(repr(float(v)) for v in blocks(B)) + "\n"
    out = subprocess.run([BIN, str(method)], input=txt, capture_output=True, text=True,
                         timeout=120, check=True).stdout.split("\n")
    n = int(out[0])
    got = np.array([[float.fromhex(t) for t in line.split()] for line in out[1:1 + n]])
    rc, want, _ = ol.or_detect(method, A, B, 1e-4, 1e-4, 1, 1, 7, 9)
    assert n == len(want)
    ref = np.concatenate([want["x"], want["normal"], want["depth"][:, None], want["P"],
                          want["Q"], want["eps_total"][:, None]], axis=1)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
