"""Regenerate the golden fixtures of tests/golden/ (run in the build container).

    python tests/golden/make_fixtures.py [/path/to/reference]

Writes data only (inputs and expected outputs):
  reference_kats.json   known-answer data held by the reference's own tests:
      - transition matrices of apps/libs/cma_utils/tests/test_transition_matrix.cpp:7-26
      - the 16-compartment flow matrix and neighbour table of
        apps/libs/cma_utils/tests/test_transport.cpp:96-373 (numbers read from the file's text)
      - the 0-D case of apps/api/tests/data/0d (vofL.raw: u32 count + f64 volumes;
        flowL.raw: u32 rows, u32 cols + f64 values)
      - feed flows of apps/libs/simulation/tests/test_feed.cpp:44-123 (f0=1,
        df=alpha=0.1, dt=0.1, 12 updates)
  transport_scipy.npz   one explicit transport step on small random networks,
      computed with scipy.sparse (an independent library) from the same COO
      inputs: m' = m + dt*(C@M - C*sink + S), C' = m'/V.
"""
from __future__ import annotations

import json
import math
import os
import re
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_transport16(ref):
    path = os.path.join(ref, "apps/libs/cma_utils/tests/test_transport.cpp")
    text = open(path).read()
    start = text.index("get_raw_data()\n")
    body = text[start:]
    body = body[body.index("return {{"):]
    body = body[: body.index("}};") + 3]
    nums = [float(x) for x in re.findall(r"[-+]?\d+\.\d*(?:[eE][-+]?\d+)?", body)]
    assert len(nums) == 256, len(nums)
    return np.array(nums).reshape(16, 16)


def read_neighbors16(ref):
    """The neighbour table of the same test (test_transport.cpp:96-114): rows
    in rcmtool's order, padded with repeated ids (the compartment's own id
    where it has fewer neighbours)."""
    path = os.path.join(ref, "apps/libs/cma_utils/tests/test_transport.cpp")
    text = open(path).read()
    body = text[text.index("get_neighbors() //"):]
    body = body[body.index("return {{"):]
    body = body[: body.index("}};") + 3]
    rows = [[int(x) for x in re.findall(r"\d+", r)] for r in re.findall(r"\{([\d,\s]+)\}", body)]
    assert len(rows) == 16 and all(len(r) == 3 for r in rows), rows
    return rows


def read_0d(ref):
    d = os.path.join(ref, "apps/api/tests/data/0d")
    b = open(os.path.join(d, "vofL.raw"), "rb").read()
    (n,) = struct.unpack("<I", b[:4])
    vol = np.frombuffer(b[4:4 + 8 * n], dtype="<f8").tolist()
    b = open(os.path.join(d, "flowL.raw"), "rb").read()
    rows, cols = struct.unpack("<II", b[:8])
    flows = np.frombuffer(b[8:8 + 8 * rows * cols], dtype="<f8").reshape(rows, cols).tolist()
    return {"volumes": vol, "flow_matrix": flows}


def feed_expectations():
    # test_feed.cpp: t advances by d_t after each update; 12 updates (i < n_it+1)
    f0, df, alpha, dt, n_it = 1.0, 0.1, 0.1, 0.1, 11
    ts, lin, ex = [], [], []
    t = 0.0
    for _ in range(n_it + 1):
        ts.append(t)
        lin.append(f0 + t * df)
        ex.append(f0 + math.exp(t * alpha))
        t += dt
    return {"f0": f0, "df": df, "alpha": alpha, "dt": dt, "t": ts, "linear": lin, "exponential": ex,
            "tolerance": 1e-12}


def scipy_transport(seed=11, n_cases=4):
    import scipy.sparse as sp

    rng = np.random.default_rng(seed)
    out = {}
    for k in range(n_cases):
        n = int(rng.integers(3, 40))
        ns = int(rng.integers(1, 4))
        nnz = int(rng.integers(n, 4 * n))
        rows = rng.integers(0, n, nnz).astype(np.uint32)
        cols = rng.integers(0, n, nnz).astype(np.uint32)
        flows = np.exp(rng.uniform(-12, -4, nnz))
        V = rng.uniform(1e-4, 1e-2, n)
        C = rng.uniform(0, 10, (ns, n))
        m = C * V
        S = rng.normal(0, 1e-3, (ns, n))
        sink = np.where(rng.random(n) < 0.2, rng.uniform(0, 1e-3, n), 0.0)
        dt = 0.01
        keep = rows != cols
        F = sp.coo_matrix((flows[keep], (rows[keep], cols[keep])), shape=(n, n)).tocsr()  # sums duplicates
        F.data[F.data <= 0] = 0
        F.eliminate_zeros()
        M = F - sp.diags(np.asarray(F.sum(axis=1)).ravel())
        dmdt = (M.T @ C.T).T - C * sink + S
        m2 = m + dt * dmdt
        C2 = m2 / V
        for name, arr in dict(rows=rows, cols=cols, flows=flows, V=V, C=C, m=m, S=S, sink=sink, m2=m2,
                              C2=C2).items():
            out[f"case{k}_{name}"] = arr
        out[f"case{k}_dt"] = np.array(dt)
    out["n_cases"] = np.array(n_cases)
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    kats = {
        "transition_2x2": {"flows": [[1, 2], [3, 4]], "expected": [[-2, 2], [3, -3]],
                           "source": "apps/libs/cma_utils/tests/test_transition_matrix.cpp:13-19"},
        "transition_3x3": {"flows": [[1, 2, 3], [4, 5, 6], [7, 8, 9]],
                           "expected": [[-5, 2, 3], [4, -10, 6], [7, 8, -15]],
                           "source": "apps/libs/cma_utils/tests/test_transition_matrix.cpp:21-26"},
        "transport16_flows": {"flows": read_transport16(ref).tolist(),
                              "source": "apps/libs/cma_utils/tests/test_transport.cpp:96-373"},
        "transport16_neighbors": {"neighbors": read_neighbors16(ref),
                                  "source": "apps/libs/cma_utils/tests/test_transport.cpp:96-114"},
        "case_0d": dict(read_0d(ref), source="apps/api/tests/data/0d"),
        "feed": dict(feed_expectations(), source="apps/libs/simulation/tests/test_feed.cpp:44-123"),
    }
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
    np.savez_compressed(os.path.join(HERE, "transport_scipy.npz"), **scipy_transport())
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
