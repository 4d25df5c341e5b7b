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
