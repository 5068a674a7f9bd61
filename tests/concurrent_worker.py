"""Child process of tests/test_gpu_concurrency.py: runs one exponent for a
number of iterations in short calls and prints Res64, the SHA-256 of the
residue words and the maximum roundoff.

    python tests/concurrent_worker.py P N ITERS CHUNK
"""
import hashlib
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np  # noqa: E402

import cudalucas_amd as L  # noqa: E402


def main():
    p, n, iters, chunk = (int(a) for a in sys.argv[1:5])
    with L.LLContext(p, n, flags=L.LL_OPT_FORCE_SWEEPS) as c:
        c.load_initial(0)
        worst = 0.0
        for _ in range(iters // chunk):
            worst = max(worst, c.iterate(chunk))
        w, r = c.residue()
    digest = hashlib.sha256(np.ascontiguousarray(w, dtype="<u4").tobytes()).hexdigest()
    print("0x%016x %s %.6f" % (r, digest, worst))


if __name__ == "__main__":
    main()
