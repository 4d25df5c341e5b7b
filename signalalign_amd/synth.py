"""Synthetic nanopore reads for the benchmark and parity workloads.

Every read is generated from a counter-based seed (base_seed, read index, stream)
so the CPU oracle and the GPU see identical bytes.  The recipe:
  * reference: uniform random ACGT; for the ACEGT alphabet, C -> E at a given
    fraction of CpG sites (and the complementary C of those sites on the reverse
    strand);
  * events: walk the reference k-mers with stay probability 0.3, skip probability
    0.05 (else step 1); each event mean ~ N(level_mean*scale + shift, level_sd)
    and noise ~ InverseGaussian(noise_mean, lambda) for a Gaussian model, or the
    mean drawn from the k-mer's density grid for an HDP model;
  * anchors: every 100 events the true (k-mer, event) pair with the event index
    jittered uniformly by +-3, kept strictly increasing and inside the matrix;
  * per-read scaling: scale ~ U(0.95, 1.05), shift ~ U(-5, 5), var = 1.
"""
import numpy as np

from . import models as _models

COMPLEMENT = bytes.maketrans(b"ACGT", b"TGCA")


def rng_for(seed, read, stream):
    return np.random.default_rng(np.random.SeedSequence([int(seed), int(read), int(stream)]))


def random_reference(n, rng, methyl_cpg=0.0):
    """Uniform ACGT of length n; with methyl_cpg > 0, that fraction of CpG sites has
    C -> E.  Returns (forward, reverse-complement) as bytes."""
    codes = rng.integers(0, 4, n)
    seq = np.frombuffer(b"ACGT", dtype=np.uint8)[codes].copy()
    rc = np.frombuffer(bytes(seq[::-1]).translate(COMPLEMENT), dtype=np.uint8).copy()
    if methyl_cpg > 0 and n > 1:
        cpg = np.nonzero((seq[:-1] == ord("C")) & (seq[1:] == ord("G")))[0]
        meth = cpg[rng.random(len(cpg)) < methyl_cpg]
        seq[meth] = ord("E")
        rc[n - 2 - meth] = ord("E")
    return bytes(seq), bytes(rc)


def walk(lX, rng, stay=0.3, skip=0.05):
    """True path: k-mer index of each event (stay 0.3, skip 0.05, step otherwise)."""
    if lX <= 0:
        return np.zeros(0, dtype=np.int64)
    n_max = int(lX / (1.0 - stay + skip) * 1.5) + 64
    u = rng.random(n_max)
    moves = np.where(u < stay, 0, np.where(u < stay + skip, 2, 1))
    pos = np.concatenate([[0], np.cumsum(moves)[:-1]])
    while pos[-1] < lX:  # extremely unlikely: extend
        u = rng.random(n_max)
        more = np.where(u < stay, 0, np.where(u < stay + skip, 2, 1))
        pos = np.concatenate([pos, pos[-1] + np.cumsum(more)])
    return pos[pos < lX]


def anchors_from_path(path, lX, lY, rng, every=100, jitter=3):
    out = []
    px, py = -1, -1
    for j in range(every, lY, every):
        x = int(path[j])
        y = j + int(rng.integers(-jitter, jitter + 1))
        if x > px and y > py and x < lX and y < lY and y >= 0:
            out.append((x, y))
            px, py = x, y
    return np.array(out, dtype=np.int64).reshape(-1, 2)


def _quantiles(model, n=4097):
    """Inverse CDF of each density row on n uniform probabilities (piecewise linear
    CDF over the grid), cached on the model: event levels are drawn from it."""
    q = getattr(model, "_quantile_table", None)
    if q is None:
        g = model.grid
        mass = 0.5 * (model.density[:, 1:] + model.density[:, :-1]) * np.diff(g)[None, :]
        cdf = np.concatenate([np.zeros((len(mass), 1)), np.cumsum(mass, axis=1)], axis=1)
        cdf /= cdf[:, -1:]
        p = np.linspace(0.0, 1.0, n)
        q = np.stack([np.interp(p, cdf[r], g) for r in range(len(cdf))])
        model._quantile_table = q
    return q


def simulate_read(model, ref, rng, scale=None, shift=None, var=1.0):
    """Events and anchors for one strand of `ref` under `model` (GaussianModel or
    HdpModel).  Returns dict(ref, events (n,4), anchors (m,2), scale, shift, var, path)."""
    k = model.kmer_length
    lX = len(ref) - k + 1
    kids = _models.kmer_ids(ref, model.alphabet, k)
    if scale is None:
        scale = float(rng.uniform(0.95, 1.05))
    if shift is None:
        shift = float(rng.uniform(-5.0, 5.0))
    path = walk(lX, rng)
    lY = len(path)
    kk = kids[path]
    ev = np.zeros((lY, 4))
    if isinstance(model, _models.HdpModel):
        rows = model.kmer_row[kk]
        q = _quantiles(model)
        u = rng.random(lY) * (q.shape[1] - 1)
        i0 = np.minimum(u.astype(np.int64), q.shape[1] - 2)
        f = u - i0
        level = q[rows, i0] * (1.0 - f) + q[rows, i0 + 1] * f
        mu = model.level_mean[kk]
        ev[:, 0] = scale * mu + shift + var * (level - mu)
        ev[:, 1] = rng.uniform(0.8, 2.0, lY)
    else:
        t = model.match[kk]
        ev[:, 0] = rng.normal(t[:, 0] * scale + shift, t[:, 1])
        ev[:, 1] = rng.wald(t[:, 2], t[:, 4])
    ev[:, 2] = 0.01
    ev[:, 3] = np.arange(lY) * 0.01
    anchors = anchors_from_path(path, lX, lY, rng)
    return {"ref": ref, "events": ev, "anchors": anchors, "scale": scale, "shift": shift, "var": var, "path": path}


def lognormal_lengths(n, median, sigma, lo, hi, rng):
    return np.clip(np.round(np.exp(np.log(median) + sigma * rng.standard_normal(n))), lo, hi).astype(np.int64)


# ---------------------------------------------------------------------------
# Workloads (the five configurations)
# ---------------------------------------------------------------------------

def gaussian_model(alphabet="ACGT", k=6, seed=1):
    return _models.synthetic_gaussian_model(alphabet, k, seed)


def hdp_model(alphabet="ACEGT", k=5, seed=3, grid_length=2000):
    return _models.synthetic_hdp_model(alphabet, k, seed, grid_length=grid_length)


def single_read(model, ref_len, seed, read=0, methyl_cpg=0.0):
    rng = rng_for(seed, read, 0)
    fwd, _ = random_reference(ref_len, rng, methyl_cpg)
    return simulate_read(model, fwd, rng_for(seed, read, 1))


def batch_reads(model, n_reads, median, seed, sigma=0.5, lo=1000, hi=50000, methyl_cpg=0.0, strands=1,
                complement_model=None):
    """n_reads reads with log-normal reference lengths; strands=2 adds the
    complement strand (reverse-complement target) of each read."""
    lens = lognormal_lengths(n_reads, median, sigma, lo, hi, rng_for(seed, 0x7fffffff, 0))
    out = []
    for i in range(n_reads):
        rng = rng_for(seed, i, 0)
        fwd, rc = random_reference(int(lens[i]), rng, methyl_cpg)
        r = simulate_read(model, fwd, rng_for(seed, i, 1))
        r["read"] = i
        r["strand"] = 0
        out.append(r)
        if strands == 2:
            r2 = simulate_read(complement_model or model, rc, rng_for(seed, i, 2))
            r2["read"] = i
            r2["strand"] = 1
            out.append(r2)
    return out
