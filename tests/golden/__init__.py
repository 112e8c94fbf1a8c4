"""Golden fixtures written from the reference build (see make_golden.py)."""
import glob
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


class Fixture:
    def __init__(self, path):
        z = np.load(path, allow_pickle=False)
        self.arrays = {k: z[k] for k in z.files}
        self.meta = json.loads(bytes(self.arrays["meta"]).decode())
        self.name = self.meta["name"]

    def blocks(self):
        from nuc3d_amd.cases import BlockSpec
        out = []
        for b in range(self.meta["nblocks"]):
            a = lambda k: self.arrays["b%d_%s" % (b, k)]
            faces = [(int(t), int(i), [float(x) for x in v]) for t, i, v in zip(a("ftype"), a("fid"), a("fval"))]
            out.append(BlockSpec(tuple(int(x) for x in a("dims")), faces, a("jac").copy(), a("met").copy(),
                                 a("dxyz").copy(), a("q_in").copy()))
        return out

    def out(self, b, key):
        return self.arrays["b%d_out_%s" % (b, key)]

    def params(self):
        return dict(self.meta["params"])


def all_fixtures():
    return [Fixture(p) for p in sorted(glob.glob(os.path.join(HERE, "*.npz")))]
