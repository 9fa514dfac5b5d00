"""Write ``c3_oracle_fit.npz``: the FP64 oracle CP-ALS run on the 128^4
rank-64 precision workload (tensorly_amd.synthetic.cp_128x4_rank64, float32
data promoted to float64), init='svd', tol=0, for ``sweeps`` sweeps: the
per-sweep relative errors, the reference the GPU precisions' fits are
measured against.

Plain ALS keeps no state beyond the factors, so the run is made of chunks of
``chunk`` sweeps, each started from the previous chunk's factors (the same
iterates as one long run); the fixture is rewritten after every chunk, so an
interrupted run keeps what it reached.  CPU only; about half an hour on 8 cores.

Run from the repository root:  python tests/golden/make_c3_oracle_fit.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import tensorly_np as T  # noqa: E402
from tensorly_amd import synthetic  # noqa: E402

OUT = os.path.join(HERE, "c3_oracle_fit.npz")


def main(sweeps=200, chunk=10):
    t0 = time.time()
    x32, _ = synthetic.cp_128x4_rank64()
    x = x32.astype(np.float64)
    del x32
    errs = []
    init = "svd"
    while len(errs) < sweeps:
        (w, f), e = T.parafac(x, 64, n_iter_max=min(chunk, sweeps - len(errs)), init=init, tol=0,
                              return_errors=True)
        errs.extend(e)
        init = (w, f)
        np.savez_compressed(OUT, errors=np.asarray(errs))
        print(f"{len(errs)} sweeps {time.time() - t0:.0f}s error {errs[-1]:.12f}", flush=True)


if __name__ == "__main__":
    main()
