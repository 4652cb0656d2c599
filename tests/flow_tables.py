"""Flowmap tables in the form the reference's hydrodynamic state hands to the
step (ReactorDomain::update and set_transition, apps/libs/mc/src/domain.cpp:44-74,
apps/libs/simulation/src/implScalar.cpp:316-320), built here from a flow
matrix: neighbour rows in a chosen order, padded to a common width, their
cumulative leaving probabilities, the out-flows, and the transition matrix
as COO triplets."""
import numpy as np


def canonical(n, rows, cols, flows):
    """Dense F(i, j): off-diagonal flows, duplicates summed in input order,
    non-positive totals dropped (what both flowmap builders do)."""
    F = np.zeros((n, n))
    for i, j, q in zip(rows, cols, flows):
        if i != j:
            F[i, j] += q
    F[F <= 0.0] = 0.0
    return F


def tables(n, rows, cols, flows, order="ascending", pad="last", extra=0, split=False, seed=0):
    """(nbr [n][m] u64, cdf [n][m], out_flow [n], t_rows, t_cols, t_vals).

    order: "ascending" neighbour ids, or "shuffled" (a random order per row);
    pad: "last" repeats the row's last neighbour (the flows-based builder's
    padding), "self" the compartment's own id (the reference's neighbour
    fixture, apps/libs/cma_utils/tests/test_transport.cpp:96-114); padded
    columns keep the row's last cumulative value.  extra adds padded columns.
    split: each off-diagonal triplet is given as two halves, in shuffled
    triplet order (the sum of the halves is the flow rounded once more)."""
    rng = np.random.default_rng(seed)
    F = canonical(n, rows, cols, flows)
    deg = (F > 0).sum(axis=1)
    m = max(1, int(deg.max()) if n else 1) + extra
    nbr = np.zeros((n, m), np.uint64)
    cdf = np.zeros((n, m))
    out = np.zeros(n)
    for i in range(n):
        js = np.nonzero(F[i] > 0)[0]
        if order == "shuffled":
            js = rng.permutation(js)
        acc = 0.0
        for q in F[i, js]:
            acc += q
        out[i] = acc
        run = 0.0
        for c, j in enumerate(js):
            run += F[i, j]
            cdf[i, c] = run / acc
            nbr[i, c] = j
        last = js[-1] if len(js) else i
        for c in range(len(js), m):
            nbr[i, c] = i if pad == "self" else last
            cdf[i, c] = 1.0 if len(js) else 0.0
    ti, tj = np.nonzero(F)
    tv = F[ti, tj]
    t_rows, t_cols, t_vals = list(ti), list(tj), list(tv)
    if split:
        h = tv * 0.5
        t_rows, t_cols, t_vals = list(ti) * 2, list(tj) * 2, list(h) + list(tv - h)
    t_rows += list(range(n))
    t_cols += list(range(n))
    t_vals += list(-out)
    t_rows, t_cols, t_vals = np.array(t_rows, np.uint64), np.array(t_cols, np.uint64), np.array(t_vals)
    if order == "shuffled":
        p = rng.permutation(len(t_vals))
        t_rows, t_cols, t_vals = t_rows[p], t_cols[p], t_vals[p]
    return nbr, cdf, out, t_rows, t_cols, t_vals


def apply_tables(case, run, particles=True, **kw):
    """case.apply(run) with every flowmap given as tables."""
    if case.params is not None:
        run.set_params(case.params)
    for k, (V, r, c, q) in enumerate(case.maps):
        run.add_flowmap_tables(V, *tables(case.n_c, r, c, q, seed=k, **kw))
    run.set_steps_per_map(case.steps_per_map)
    for f in case.feeds:
        run.add_feed(*f)
    run.set_concentrations(case.C0)
    if case.gas is not None:
        run.set_gas_transfer(*case.gas)
    if particles:
        run.init_particles(case.n_particles, case.capacity or 2 * case.n_particles, case.x0)
    return run
