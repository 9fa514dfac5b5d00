"""Write ``cp512_converged.npz``: the FP64 oracle CP-ALS run on the 512^3
rank-32 workload (tensorly_amd.synthetic.cp_512_rank32, float32 data promoted
to float64), init='random' with random_state=1, tol=0, continued until the
relative error changes by less than 1e-9 between sweeps (or ``max_sweeps``).

Plain ALS keeps no state beyond the factors, so the run is made of chunks of
``chunk`` sweeps, each started from the previous chunk's factors (the same
iterates as one long run); the fixture is rewritten after every chunk, so an
interrupted run keeps what it reached.  The converged error is the target of
the time-to-fit measurement.  CPU only; takes hours on 8 cores.

Run from the repository root:  python tests/golden/make_cp512_converged.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import tensorly_np as T  # noqa: E402
from tensorly_amd import synthetic  # noqa: E402

OUT = os.path.join(HERE, "cp512_converged.npz")


def main(max_sweeps=6000, chunk=25, tol=1e-9):
    t0 = time.time()
    x32, _ = synthetic.cp_512_rank32()
    x = x32.astype(np.float64)
    del x32
    errs = []
    init = "random"
    while len(errs) < max_sweeps:
        (w, f), e = T.parafac(x, 32, n_iter_max=chunk, init=init, random_state=1, tol=0, return_errors=True)
        errs.extend(e)
        init = (w, f)
        converged = len(errs) >= 2 and abs(errs[-2] - errs[-1]) < tol
        np.savez_compressed(OUT, errors=np.asarray(errs), converged=np.asarray(converged))
        print(f"{len(errs)} sweeps {time.time() - t0:.0f}s error {errs[-1]:.12f} change {errs[-2] - errs[-1]:.3e}",
              flush=True)
        if converged:
            break


if __name__ == "__main__":
    main()
