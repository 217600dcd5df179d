"""Converts the reference's own regression fixtures (tests/data/*.pt in the reference repo) into
plain .npz data files under tests/golden/.

The .pt files hold only tensors (the input landscape and dicts of output tensors); they are read
with torch.load(weights_only=True), which executes nothing from the file. The .npz files are
written with allow_pickle disabled. Run once in a container where the reference is mounted:

    python tests/golden/make_fixtures.py /root/reference/tests/data
"""

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))


def main(src):
    land = torch.load(os.path.join(src, "landscape_32x32.pt"), weights_only=True)
    np.savez(os.path.join(HERE, "landscape_32x32.npz"), field=land.numpy().astype(np.float32))
    for mode in ("st", "mt", "gpu"):
        for dual in ("primal", "dual"):
            for flt in ("raw", "flt"):
                name = f"ms_{mode}_{dual}_{flt}"
                d = torch.load(os.path.join(src, name + ".pt"), weights_only=True)
                arrays = {k: v.numpy() for k, v in d.items() if v is not None}
                np.savez(os.path.join(HERE, name + ".npz"), **arrays)
                print(name, {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tests/data")
