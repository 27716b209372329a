"""Writes tests/golden/qaoa_n26_p4.json: the oracle's evaluate_qaoa_cost
(qaoa.rs:179-201) on the 26-node random 3-regular graph (seed 3), p = 4,
angles gamma, beta ~ U(0, pi), plus the amplitudes at 64 fixed indices and the
per-edge <Z_i Z_j>.  The oracle is the C++ restatement in oracle/; the file is
a fixture so GPU tests at full size need not rerun the slow CPU oracle."""
import json
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from ruqu_amd import workloads  # noqa: E402


def main():
    n, p = 26, 4
    edges = workloads.random_regular_graph(n, 3, 3)
    u = workloads.uniforms(workloads.stream_seed(3, 1), 2 * p)
    gammas = list(u[:p] * math.pi)
    betas = list(u[p:] * math.pi)
    w = [1.0] * len(edges)
    cost, amps = orc.qaoa_cost(n, edges, w, gammas, betas, want_state=True)
    idx = (workloads.uniforms(workloads.stream_seed(3, 2), 64) * (1 << n)).astype(np.int64)
    zz = [orc.expectation(amps, n, [(i, 3), (j, 3)]) for i, j in edges[:8]]
    out = {
        "n": n, "p": p, "edges": edges, "gammas": gammas, "betas": betas, "cost": cost,
        "indices": idx.tolist(),
        "amps_re": [float(amps[i].real) for i in idx], "amps_im": [float(amps[i].imag) for i in idx],
        "amps_re_hex": [float(amps[i].real).hex() for i in idx], "amps_im_hex": [float(amps[i].imag).hex() for i in idx],
        "zz_first8": zz,
        "norm": float(np.sum(np.abs(amps) ** 2)),
    }
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "qaoa_n26_p4.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("cost", cost)


if __name__ == "__main__":
    main()
