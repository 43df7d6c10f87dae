"""CPU checks of the pilot's training helpers and of the shipped weights files
(the network itself runs on the GPU: tests/test_gpu_pilot.py)."""
import json

import numpy as np
import pytest


def test_dihedral_is_the_eight_symmetries_of_the_square():
    import torch
    from isp_mw_amd import pilot as P
    a = torch.arange(2 * 3 * 4 * 4, dtype=torch.float32).reshape(2, 3, 4, 4)
    seen = set()
    for k in range(8):
        got = P.dihedral(a, k).numpy()
        want = a.numpy()
        if k & 4:
            want = np.swapaxes(want, -2, -1)
        want = np.rot90(want, k & 3, axes=(-2, -1))
        np.testing.assert_array_equal(got, want)
        seen.add(got.tobytes())
        # batch and plane axes are untouched: each plane maps by itself
        np.testing.assert_array_equal(P.dihedral(a[1:2, 2:3], k).numpy(), got[1:2, 2:3])
    assert len(seen) == 8
    np.testing.assert_array_equal(P.dihedral(a, 0).numpy(), a.numpy())


@pytest.mark.parametrize("kind", ["grey", "rgb"])
def test_shipped_weights_file(kind):
    """The weights file holds exactly the network's parameters (finite, in the
    layout PilotCNN uses), no pickled objects, and the recipe that made it."""
    from isp_mw_amd import pilot as P, train_pilot
    path = P.weights_path(kind)
    with np.load(path, allow_pickle=False) as z:
        files = set(z.files)
        params, channels = z["params"], int(z["channels"])
        meta = json.loads(str(z["meta"]))
    assert files == {"params", "step", "channels", "meta"}  # weights only, no optimizer state
    assert channels == (1 if kind == "grey" else 3)
    n = sum(cin * cout * 9 + cout for cin, cout, _, _, _ in P.layers(channels))
    assert params.dtype == np.float32 and params.shape == (n,)
    assert np.all(np.isfinite(params))
    want = train_pilot.recipe(kind, meta["steps"])
    assert {k: meta[k] for k in want} == json.loads(json.dumps(want))
    assert meta["loss_last"] < meta["loss_first"]
    assert P.PilotCNN.meta_of(path) == meta
