"""Reader for signalAlign's per-read input file (.npRead).

The layout follows nanopore_loadNanoporeReadFromFile (impl/nanopore.c:98-424):
whitespace-separated fields on 14 lines --

  1  read length, template events, complement events, template read length,
     complement read length, template scale, shift, var, scale_sd, var_sd, drift,
     complement scale, shift, var, scale_sd, var_sd, drift, has-2D flag (18 fields)
  2  2D read (only when the flag is set)
  3  template read
  4  template strand event map (one event index per template base)
  5  complement read (only with 2D)
  6  complement strand event map
  7  template 2D event map (one per 2D base)
  8  template events: mean, stdv, length, start per event
  9  complement 2D event map
  10 complement events
  11 template model states (k-mer strings, one per event)
  12 template P(model state)
  13 complement model states
  14 complement P(model state)

Field counts are checked as the reference checks them; a mismatch raises
ValueError naming the line.
"""
import numpy as np

PARAM_NAMES = ("scale", "shift", "var", "scale_sd", "var_sd", "drift")


def _ints(tokens, what):
    try:
        return np.array([int(t) for t in tokens], dtype=np.int64)
    except ValueError as e:
        raise ValueError(f"error loading in {what}") from e


def _floats(tokens, what):
    try:
        return np.array([float(t) for t in tokens], dtype=np.float64)
    except ValueError as e:
        raise ValueError(f"error loading in {what}") from e


def load_npread(path):
    """Returns a dict: read_length, template/complement {events (n, 4) f64, params
    (dict of PARAM_NAMES), read (str), strand_event_map, event_map (2D), model_state
    (list of str), p_model}, two_d (bool)."""
    with open(path) as f:
        lines = f.read().split("\n")
    lines += [""] * (14 - len(lines))
    h = lines[0].split()
    if len(h) != 18:
        raise ValueError(f"{path}: incorrect line 1, got {len(h)} tokens, should get 18")
    read_len, n_t, n_c, t_len, c_len = (int(x) for x in h[:5])
    vals = [float(x) for x in h[5:17]]
    two_d = int(h[17]) != 0
    out = {"read_length": read_len, "two_d": two_d}
    strands = {"template": {"params": dict(zip(PARAM_NAMES, vals[0:6])), "n": n_t, "len": t_len},
               "complement": {"params": dict(zip(PARAM_NAMES, vals[6:12])), "n": n_c, "len": c_len}}
    out["two_d_read"] = lines[1].split()[0] if two_d and lines[1].split() else ""
    t = strands["template"]
    t["read"] = lines[2].split()[0][:t_len] if lines[2].split() else ""
    t["strand_event_map"] = _ints(lines[3].split(), "template strand event map")
    if len(t["strand_event_map"]) != t_len:
        raise ValueError(f"template event map is not the correct length, should be {t_len}, "
                         f"got {len(t['strand_event_map'])}")
    c = strands["complement"]
    c["read"] = lines[4].split()[0][:c_len] if two_d and lines[4].split() else ""
    c["strand_event_map"] = _ints(lines[5].split(), "complement strand event map")
    if len(c["strand_event_map"]) != c_len:
        raise ValueError(f"complement event map is not the correct length, should be {c_len}, "
                         f"got {len(c['strand_event_map'])}")
    for s, ln_map, ln_ev, ln_ms, ln_pm in (("template", 6, 7, 10, 11), ("complement", 8, 9, 12, 13)):
        d = strands[s]
        d["event_map"] = _ints(lines[ln_map].split(), f"{s} 2D event map")
        if len(d["event_map"]) != read_len:
            raise ValueError(f"{s} 2D event map is not the correct length, should be {read_len}, "
                             f"got {len(d['event_map'])}")
        ev = _floats(lines[ln_ev].split(), f"{s} events")
        if len(ev) != 4 * d["n"]:
            raise ValueError(f"incorrect number of {s} events, should be {d['n']}, got {len(ev) // 4}")
        d["events"] = ev.reshape(-1, 4)
        ms = lines[ln_ms].split()
        if len(ms) != d["n"]:
            raise ValueError(f"incorrect number of {s} model states, got {len(ms)}")
        d["model_state"] = ms
        pm = _floats(lines[ln_pm].split(), f"{s} P(model_state)")
        if len(pm) != d["n"]:
            raise ValueError(f"incorrect number of {s} model probs, got {len(pm)}")
        d["p_model"] = pm
    for s in strands:
        del strands[s]["n"], strands[s]["len"]
    out.update(strands)
    return out


def load_nhdp(path):
    """Reader for a serialised nanopore HDP (.nhdp): deserialize_nhdp
    (impl/nanopore_hdp.c:1071-1098) around deserialize_hdp (impl/hdp.c:3047-3317).

    Keeps what the alignment's emissions read: the sampling grid, each Dirichlet
    process's posterior predictive density and spline slopes on it, and for each
    k-mer (DP id = kmer_id over the sorted alphabet) its first observed ancestor
    (dir_proc_density, impl/hdp.c:2583-2607; a DP is observed when a data point is
    assigned to it or to a descendant, mark_observed_dps, impl/hdp.c:1127-1155).
    The Gibbs-sampling state (factor tree, gamma samples) is skipped.

    Returns a dict: alphabet (sorted), kmer_length, grid, density (rows), slopes (rows),
    kmer_row (row of each k-mer), base (mu, nu, alpha, beta), n_dps."""
    from . import models
    with open(path) as f:
        lines = f.read().split("\n")
    pos = 0

    def nxt():
        nonlocal pos
        if pos >= len(lines):
            raise ValueError(f"{path}: unexpected end of file")
        line = lines[pos]
        pos += 1
        return line

    alphabet_size = int(nxt().split()[0])
    alphabet = nxt().split()[0][:alphabet_size]
    kmer_length = int(nxt().split()[0])
    splines_finalized = int(nxt().split()[0]) != 0
    has_data = int(nxt().split()[0]) != 0
    sample_gamma = int(nxt().split()[0]) != 0
    num_dps = int(nxt().split()[0])
    dp_ids = None
    if has_data:
        nxt()  # the data points themselves
        dp_ids = np.array([int(t) for t in nxt().split()], dtype=np.int64)
    base = tuple(float(t) for t in nxt().split()[:4])
    g = nxt().split()
    grid_start, grid_stop, grid_length = float(g[0]), float(g[1]), int(g[2])
    nxt()  # gamma
    if sample_gamma:
        for _ in range(4):  # gamma alpha, gamma beta, w, s
            nxt()
    parent = np.full(num_dps, -1, dtype=np.int64)
    for i in range(num_dps):
        tok = nxt().split()
        if tok[0] != "-":
            parent[i] = int(tok[0])
    observed = np.zeros(num_dps, dtype=bool)
    if has_data:
        for i in np.unique(dp_ids):
            j = int(i)
            while j >= 0 and not observed[j]:
                observed[j] = True
                j = int(parent[j])
    density = np.zeros((num_dps, grid_length))
    if has_data:
        for i in range(num_dps):
            tok = nxt().split()
            if tok:
                if len(tok) != grid_length:
                    raise ValueError(f"{path}: DP {i} posterior predictive has {len(tok)} values")
                density[i] = [float(t) for t in tok]
    if not splines_finalized:
        raise ValueError(f"{path}: spline slopes are not finalised (dir_proc_density refuses such an HDP)")
    slopes = np.zeros((num_dps, grid_length))
    for i in range(num_dps):
        tok = nxt().split()
        if tok:
            slopes[i] = [float(t) for t in tok]
    grid = np.array(models.linspace(grid_start, grid_stop, grid_length))
    # rows: the observed DPs; each k-mer reads its first observed ancestor's row
    alpha_sorted = "".join(sorted(alphabet))
    n_kmers = len(alpha_sorted) ** kmer_length
    rows = np.flatnonzero(observed)
    row_of = np.full(num_dps, -1, dtype=np.int64)
    row_of[rows] = np.arange(len(rows))
    kmer_row = np.empty(n_kmers, dtype=np.int64)
    for k in range(n_kmers):
        j = k
        while j >= 0 and not observed[j]:
            j = int(parent[j])
        if j < 0:
            raise ValueError(f"{path}: k-mer {k} has no observed ancestor")
        kmer_row[k] = row_of[j]
    return {"alphabet": alpha_sorted, "kmer_length": kmer_length, "grid": grid, "density": density[rows],
            "slopes": slopes[rows], "kmer_row": kmer_row, "base": base, "n_dps": num_dps}
