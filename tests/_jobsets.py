"""Job sets of the batched MI and TE drivers, for tests that run the same
batch through ent_batch_densities_dev and compare per-point densities.

They list the jobs ent_mi_matrix_dev / ent_te_matrix_dev build (mi.cpp:37-52,
transs.cpp:42-74: every column pair's joint density, the lagged 2-D and 3-D
densities for every lag, the 1-D densities), each with its own output slice,
so the engine plans, sorts and tiles this batch as it does the driver's."""


def mi_jobs(n_rows, n_cols, h):
    """2-D jobs of all column pairs, then the 1-D jobs.  Returns (jobs,
    out_slot) with out_slot[(i, j)] / out_slot[(c,)] = first output slot."""
    jobs, slot, nxt = [], {}, 0
    for i in range(n_cols):
        for j in range(i + 1, n_cols):
            jobs.append({"n_dim": 2, "n": n_rows, "n_rows": n_rows, "off": [i * n_rows, j * n_rows],
                         "h": [h[i], h[j]], "out_off": nxt})
            slot[(i, j)] = nxt
            nxt += n_rows
    for c in range(n_cols):
        jobs.append({"n_dim": 1, "n": n_rows, "n_rows": n_rows, "off": [c * n_rows], "h": [h[c]], "out_off": nxt})
        slot[(c,)] = nxt
        nxt += n_rows
    return jobs, slot, nxt


def te_jobs(n_rows, n_cols, h, taus):
    """Per lag the 3-D densities p(x_t, x_t+tau, y_t) of both directions of
    every pair (the second listed as (y_t, y_t+tau, x_t)), then the pairs'
    2-D densities, the lagged 2-D densities and the 1-D densities.  Returns
    (jobs, slot, total) with slot[("itau", t, i, j)], slot[("jtau", t, i, j)],
    slot[("pij", i, j)], slot[("ptau", t, i)], slot[("p", i)]."""
    jobs, slot = [], {}
    nxt = 0

    def add(key, cols, lags, hh):
        nonlocal nxt
        n = n_rows - max(lags)
        jobs.append({"n_dim": len(cols), "n": n, "n_rows": n_rows,
                     "off": [c * n_rows + t for c, t in zip(cols, lags)], "h": list(hh), "out_off": nxt})
        slot[key] = nxt
        nxt += n

    for t, tau in enumerate(taus):
        for i in range(n_cols):
            for j in range(i + 1, n_cols):
                add(("itau", t, i, j), (i, i, j), (0, tau, 0), (h[i], h[i], h[j]))
                add(("jtau", t, i, j), (j, j, i), (0, tau, 0), (h[j], h[j], h[i]))
    for i in range(n_cols):
        for j in range(i + 1, n_cols):
            add(("pij", i, j), (i, j), (0, 0), (h[i], h[j]))
    for t, tau in enumerate(taus):
        for i in range(n_cols):
            add(("ptau", t, i), (i, i), (0, tau), (h[i], h[i]))
    for i in range(n_cols):
        add(("p", i), (i,), (0,), (h[i],))
    return jobs, slot, nxt
