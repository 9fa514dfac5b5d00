"""Write ``cp512_oracle.npz``: the FP64 oracle CP-ALS trajectory for the
512^3 rank-32 workload (tensorly_amd.synthetic.cp_512_rank32, float32 data
promoted to float64), init='random' with random_state=1, 100 sweeps, tol=0.

The per-sweep relative errors define the target fit for the time-to-fit
measurement (fit = 1 - error after the last sweep).  Takes a few minutes of
CPU; the result is a small fixture.

Run from the repository root:  python tests/golden/make_cp512_oracle.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import tensorly_np as T  # noqa: E402
from tensorly_amd import synthetic  # noqa: E402


def main(n_iter=100):
    t0 = time.time()
    x32, _ = synthetic.cp_512_rank32()
    checksum = np.array([float(x32.astype(np.float64).sum()), float((x32.astype(np.float64) ** 2).sum())])
    x = x32.astype(np.float64)
    del x32
    print(f"data {time.time() - t0:.1f}s", flush=True)
    (w, f), errs = T.parafac(x, 32, n_iter_max=n_iter, init="random", random_state=1, tol=0, return_errors=True)
    print(f"oracle {time.time() - t0:.1f}s final error {errs[-1]:.10f}", flush=True)
    np.savez_compressed(os.path.join(HERE, "cp512_oracle.npz"), errors=np.asarray(errs), checksum=checksum)


if __name__ == "__main__":
    main()
