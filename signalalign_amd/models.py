"""Model tables for the three-state signal-to-reference pair-HMM.

Host-side mirror of the reference's model handling:
  * `load_model_file` restates stateMachine3_loadFromFile
    (impl/stateMachine.c:1328-1426) with stateMachine3_loadTransitionsFromFile
    (:1090-1146): log transitions, the match table copied to the extra-event
    (gapY) table with level_sd x 1.75, gapX emission log(0.1)
    (stateMachine3_construct, :1474-1476), GAP_SWITCH_TO_X = log(0) from
    stateMachine3_setTransitionsToNanoporeDefaults (:1077-1088).
  * `scale_noise` restates emissions_signal_scaleNoise (:609-629).
  * synthetic ONT-style Gaussian tables and HDP density grids for benchmarks.

Tables are float64 numpy arrays in the reference's units.
"""
import math

import numpy as np

# Transition slots in the field order of struct _StateMachine3 (inc/stateMachine.h:141-160).
T_MATCH_CONTINUE, T_MATCH_FROM_GAP_X, T_MATCH_FROM_GAP_Y, T_GAP_OPEN_X, T_GAP_OPEN_Y, \
    T_GAP_EXTEND_X, T_GAP_EXTEND_Y, T_GAP_SWITCH_TO_X, T_GAP_SWITCH_TO_Y = range(9)

EXTRA_EVENT_NOISE_MULTIPLIER = 1.75
LOG_GAP_X = -2.3025850929940455  # log(0.1)

# Line 1 of the shipped .model files (match->match ... gapY->gapY, likelihood).
DEFAULT_TRANSITION_LINE = [0.790158882824, 0.196524254983, 0.0133168621928, 0.803475745017,
                           0.196524254983, 0.0, 0.986683137807, 1e-09, 0.0133168621929]


def transitions_from_line(vals):
    """impl/stateMachine.c:1090-1146: nine probabilities (row-major from->to over
    match, gapX, gapY) to the nine log-transition slots."""
    t = np.full(9, -np.inf)
    with np.errstate(divide="ignore"):
        t[T_MATCH_CONTINUE] = math.log(vals[0]) if vals[0] > 0 else -np.inf
        t[T_GAP_OPEN_X] = math.log(vals[1]) if vals[1] > 0 else -np.inf
        t[T_GAP_OPEN_Y] = math.log(vals[2]) if vals[2] > 0 else -np.inf
        t[T_MATCH_FROM_GAP_X] = math.log(vals[3]) if vals[3] > 0 else -np.inf
        t[T_GAP_EXTEND_X] = math.log(vals[4]) if vals[4] > 0 else -np.inf
        t[T_MATCH_FROM_GAP_Y] = math.log(vals[6]) if vals[6] > 0 else -np.inf
        t[T_GAP_SWITCH_TO_Y] = math.log(vals[7]) if vals[7] > 0 else -np.inf
        t[T_GAP_EXTEND_Y] = math.log(vals[8]) if vals[8] > 0 else -np.inf
    t[T_GAP_SWITCH_TO_X] = -np.inf
    return t


class GaussianModel:
    """Tables of a StateMachine3 with Gaussian/inverse-Gaussian emissions."""

    def __init__(self, alphabet, kmer_length, transitions, match, gapy=None, gapx=None):
        self.alphabet = "".join(sorted(alphabet))
        if len(set(self.alphabet)) != len(self.alphabet):
            raise ValueError("alphabet characters must be distinct")
        self.kmer_length = int(kmer_length)
        self.n_kmers = len(self.alphabet) ** self.kmer_length
        self.transitions = np.asarray(transitions, dtype=np.float64).copy()
        self.match = np.asarray(match, dtype=np.float64).reshape(self.n_kmers, 5).copy()
        if gapy is None:
            gapy = self.match.copy()
            gapy[:, 1] *= EXTRA_EVENT_NOISE_MULTIPLIER
        self.gapy = np.asarray(gapy, dtype=np.float64).reshape(self.n_kmers, 5).copy()
        self.gapx = (np.full(self.n_kmers, LOG_GAP_X) if gapx is None
                     else np.asarray(gapx, dtype=np.float64).copy())


class HdpModel:
    """Tables of a StateMachine3_HDP: the k-mer density grid (posterior predictive
    densities and spline slopes on a uniform grid), the row of the first observed
    ancestor DP for each k-mer, and the level means used to descale events."""

    def __init__(self, alphabet, kmer_length, transitions, grid, density, slopes, kmer_row, level_mean):
        self.alphabet = "".join(sorted(alphabet))
        self.kmer_length = int(kmer_length)
        self.n_kmers = len(self.alphabet) ** self.kmer_length
        self.transitions = np.asarray(transitions, dtype=np.float64).copy()
        self.grid = np.asarray(grid, dtype=np.float64).copy()
        self.density = np.asarray(density, dtype=np.float64).reshape(-1, len(self.grid)).copy()
        self.slopes = np.asarray(slopes, dtype=np.float64).reshape(-1, len(self.grid)).copy()
        self.kmer_row = np.asarray(kmer_row, dtype=np.int64).copy()
        self.level_mean = np.asarray(level_mean, dtype=np.float64).copy()


def merge_strand_models(mt, mc):
    """One HDP table holding two models of the same k-mer length, grid and transitions
    (a read's template and complement models): the alphabet doubles, the second
    model's k-mers spelled in lower case (a complement strand's reference lower-cased
    reads them), so a batch can hold reads of both strands under one resident table.
    Each k-mer of one case keeps its model's density row and level mean; mixed-case
    k-mers never occur and point at row 0.  Results equal those of each strand's
    reads run with its own model."""
    if mt.kmer_length != mc.kmer_length or mt.alphabet != mc.alphabet or not np.array_equal(mt.grid, mc.grid) \
            or not np.array_equal(mt.transitions, mc.transitions):
        raise ValueError("the models differ in k-mer length, alphabet, grid or transitions")
    lower = mt.alphabet.lower()
    if set(lower) & set(mt.alphabet):
        raise ValueError("the alphabet has lower-case letters")
    alpha = "".join(sorted(mt.alphabet + lower))
    A, k = len(alpha), mt.kmer_length
    a0 = len(mt.alphabet)
    # digit of each sorted letter in its own model's sorted alphabet, and its case
    pos = {ch: i for i, ch in enumerate(alpha)}
    own = np.array([sorted(mt.alphabet).index(ch) if ch in mt.alphabet else sorted(lower).index(ch) for ch in alpha])
    upper = np.array([ch in mt.alphabet for ch in alpha])
    ids = np.arange(A ** k)
    digits = np.stack([(ids // A ** (k - 1 - j)) % A for j in range(k)], axis=1)
    all_up = upper[digits].all(axis=1)
    all_lo = (~upper[digits]).all(axis=1)
    sub = (own[digits] * (a0 ** np.arange(k - 1, -1, -1))[None, :]).sum(axis=1)
    nrt = mt.density.shape[0]
    kmer_row = np.zeros(A ** k, dtype=np.int64)
    level_mean = np.zeros(A ** k)
    kmer_row[all_up] = mt.kmer_row[sub[all_up]]
    kmer_row[all_lo] = nrt + mc.kmer_row[sub[all_lo]]
    level_mean[all_up] = mt.level_mean[sub[all_up]]
    level_mean[all_lo] = mc.level_mean[sub[all_lo]]
    del pos
    return HdpModel(alpha, k, mt.transitions, mt.grid, np.concatenate([mt.density, mc.density]),
                    np.concatenate([mt.slopes, mc.slopes]), kmer_row, level_mean)


def load_model_file(path):
    """Parse a signalAlign .model file (impl/stateMachine.c:1328-1426)."""
    with open(path) as fh:
        head = fh.readline().split()
        if len(head) != 4:
            raise ValueError("model header must have 4 fields")
        state_number, alphabet_size, alphabet, kmer_length = int(head[0]), int(head[1]), head[2], int(head[3])
        if state_number != 3:
            raise ValueError("only three-state models are supported")
        if len(alphabet) != alphabet_size:
            raise ValueError("alphabet size mismatch")
        trans = fh.readline().split()
        if len(trans) != state_number * state_number + 1:
            raise ValueError("bad transitions line")
        transitions = transitions_from_line([float(v) for v in trans[:9]])
        vals = np.array([float(v) for v in fh.readline().split()], dtype=np.float64)
    n = alphabet_size ** kmer_length
    if len(vals) != 5 * n:
        raise ValueError(f"expected {5 * n} emission values, got {len(vals)}")
    return GaussianModel(alphabet, kmer_length, transitions, vals.reshape(n, 5))


def write_model_file(model, path, transition_line=None):
    """Write a GaussianModel in the .model format load_model_file reads (header,
    nine transition probabilities + likelihood, the match table).  The
    transitions written are `transition_line` (probabilities; the shipped
    default when None), so the model must have been built from them."""
    line = list(DEFAULT_TRANSITION_LINE if transition_line is None else transition_line)
    if not np.array_equal(transitions_from_line(line), model.transitions):
        raise ValueError("model transitions do not come from this transition line")
    with open(path, "w") as fh:
        fh.write(f"3\t{len(model.alphabet)}\t{model.alphabet}\t{model.kmer_length}\n")
        fh.write("\t".join(repr(float(v)) for v in line) + "\t0.0\n")
        fh.write("\t".join(repr(float(v)) for v in model.match.reshape(-1)) + "\n")


def scale_noise(model, scale_sd, var_sd):
    """emissions_signal_scaleNoise (impl/stateMachine.c:609-629), including the
    reference's use of the match row's lambda for the gapY noise sd."""
    m = model.match.copy()
    g = model.gapy.copy()
    m[:, 2] *= scale_sd
    m[:, 4] *= var_sd
    m[:, 3] = np.sqrt(m[:, 2] ** 3 / m[:, 4])
    g[:, 2] *= scale_sd
    g[:, 4] *= var_sd
    g[:, 3] = np.sqrt(g[:, 2] ** 3 / m[:, 4])
    return GaussianModel(model.alphabet, model.kmer_length, model.transitions, m, g, model.gapx)


def scale_model(model, scale, shift, var, scale_sd, var_sd):
    """emissions_signal_scaleModel (impl/stateMachine.c:631-667): the tables moved to a
    read's scale (level mean x scale + shift, level sd x var, noise mean x scale_sd,
    noise lambda x var_sd, noise sd recomputed; the gapY noise sd again from the match
    row's lambda, as the reference computes it)."""
    m = model.match.copy()
    g = model.gapy.copy()
    for t in (m, g):
        t[:, 0] = t[:, 0] * scale + shift
        t[:, 1] = t[:, 1] * var
        t[:, 2] = t[:, 2] * scale_sd
        t[:, 4] = t[:, 4] * var_sd
    m[:, 3] = np.sqrt(m[:, 2] ** 3.0 / m[:, 4])
    g[:, 3] = np.sqrt(g[:, 2] ** 3.0 / m[:, 4])
    return GaussianModel(model.alphabet, model.kmer_length, model.transitions, m, g, model.gapx)


def kmer_ids(seq, alphabet, kmer_length):
    """Vectorised kmer_id (impl/nanopore_hdp.c:360-393) of every k-mer of `seq`."""
    alphabet = "".join(sorted(alphabet))
    lut = np.full(256, -1, dtype=np.int64)
    for i, c in enumerate(alphabet):
        lut[ord(c)] = i
    codes = lut[np.frombuffer(seq.encode() if isinstance(seq, str) else bytes(seq), dtype=np.uint8)]
    n = len(codes) - kmer_length + 1
    if n <= 0:
        return np.zeros(0, dtype=np.int64)
    ids = np.zeros(n, dtype=np.int64)
    bad = np.zeros(n, dtype=bool)
    A = len(alphabet)
    for j in range(kmer_length):
        c = codes[j:j + n]
        bad |= c < 0
        ids = ids * A + c
    ids[bad] = -1
    return ids


def synthetic_gaussian_model(alphabet, kmer_length, seed, noise=True):
    """ONT-style table: level_mean ~ U(60,110) pA, level_sd ~ U(1,3) pA; noise
    mean ~ U(0.8,2.0), lambda ~ U(1.4,4.0), noise_sd = sqrt(nm^3/lambda)
    (the relation of impl/stateMachine.c:617-618)."""
    alphabet = "".join(sorted(alphabet))
    n = len(alphabet) ** kmer_length
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x6d6f64]))
    t = np.zeros((n, 5))
    t[:, 0] = rng.uniform(60.0, 110.0, n)
    t[:, 1] = rng.uniform(1.0, 3.0, n)
    t[:, 2] = rng.uniform(0.8, 2.0, n)
    t[:, 4] = rng.uniform(1.4, 4.0, n)
    t[:, 3] = np.sqrt(t[:, 2] ** 3 / t[:, 4])
    return GaussianModel(alphabet, kmer_length, transitions_from_line(DEFAULT_TRANSITION_LINE), t)


def synthetic_hdp_model(alphabet, kmer_length, seed, grid_length=2000, grid_start=30.0, grid_stop=130.0):
    """HDP-style density grid: each k-mer's density is a mixture of 1-4 normals
    (means U(60,110) pA, sd U(1,3) pA, Dirichlet(1) weights) tabulated on
    linspace(grid_start, grid_stop, grid_length) (impl/hdp_math_utils.c:498-510),
    with knot slopes from the spline system of impl/hdp_math_utils.c:402-442 and
    descaling level means from the grid expected value (impl/hdp.c:2772-2784)."""
    alphabet = "".join(sorted(alphabet))
    n = len(alphabet) ** kmer_length
    rng = np.random.default_rng(np.random.SeedSequence([seed, 0x686470]))
    grid = linspace(grid_start, grid_stop, grid_length)
    dens = np.zeros((n, grid_length))
    ncomp = rng.integers(1, 5, n)
    for i in range(n):
        k = int(ncomp[i])
        mu = rng.uniform(60.0, 110.0, k)
        sd = rng.uniform(1.0, 3.0, k)
        w = rng.dirichlet(np.ones(k))
        z = (grid[None, :] - mu[:, None]) / sd[:, None]
        dens[i] = (w[:, None] * np.exp(-0.5 * z * z) / (sd[:, None] * math.sqrt(2 * math.pi))).sum(0)
    slopes = spline_knot_slopes_rows(grid, dens)
    level_mean = np.array([expected_value(grid, dens[i]) for i in range(n)])
    return HdpModel(alphabet, kmer_length, transitions_from_line(DEFAULT_TRANSITION_LINE), grid, dens, slopes,
                    np.arange(n, dtype=np.int64), level_mean)


def linspace(start, stop, n):
    """impl/hdp_math_utils.c:498-510 (start + i*dx, last point exactly stop)."""
    dx = (stop - start) / float(n - 1)
    g = start + np.arange(n, dtype=np.float64) * dx
    g[-1] = stop
    return g


def expected_value(grid, distr):
    """impl/hdp.c:2772-2784 (sequential sum, as the reference)."""
    e = 0.0
    for i in range(1, len(grid)):
        e += grid[i] * distr[i] * (grid[i] - grid[i - 1])
    return e


def spline_knot_slopes_rows(x, Y):
    """impl/hdp_math_utils.c:402-442 applied to every row of Y (vectorised over
    rows; each row follows the same operation sequence as the reference)."""
    Y = np.atleast_2d(np.asarray(Y, dtype=np.float64))
    x = np.asarray(x, dtype=np.float64)
    L = len(x)
    f = L - 1
    R = Y.shape[0]
    cc = np.zeros((R, L)); rc = np.zeros(L); rh = np.zeros((R, L))
    right0 = 1.0 / (x[1] - x[0])
    center0 = 2.0 * right0
    rhs0 = 3.0 * (Y[:, 1] - Y[:, 0]) * right0 * right0
    c_prev = np.full(R, center0); r_prev = right0; rhs_prev = rhs0
    for idx in range(1, f):
        left = 1.0 / (x[idx] - x[idx - 1])
        right = 1.0 / (x[idx + 1] - x[idx])
        center = 2.0 * (left + right)
        rhs = 3.0 * ((Y[:, idx] - Y[:, idx - 1]) * left * left + (Y[:, idx + 1] - Y[:, idx]) * right * right)
        center = center - left * r_prev / c_prev
        rhs = rhs - left * rhs_prev / c_prev
        cc[:, idx] = center; rc[idx] = right; rh[:, idx] = rhs
        c_prev = center; r_prev = right; rhs_prev = rhs
    K = np.zeros((R, L))
    left = 1.0 / (x[f] - x[f - 1])
    center = 2.0 * left
    rhs = 3.0 * (Y[:, f] - Y[:, f - 1]) * left * left
    K[:, f] = (rhs * c_prev - rhs_prev * left) / (center * c_prev - r_prev * left)
    for idx in range(f - 1, 0, -1):
        K[:, idx] = (rh[:, idx] - rc[idx] * K[:, idx + 1]) / cc[:, idx]
    K[:, 0] = (rhs0 - right0 * K[:, 1]) / center0
    return K
