"""Pure-Python restatement of signalMachine's guide-alignment and event-map steps.

TEST INFRASTRUCTURE ONLY: the checker for the library's signalMachine driver
(signalalign_machine.h).  Only tests/ import this module.

Each function cites the reference code it follows.  Pins: the hand-worked
CIGAR cases and the brute-force dominance check in tests/test_machine.py; the
read-scaling and TSV-writer steps are pinned directly on the reference's own
output file (tests/golden/ZymoC_ch_1_file1.sm.forward.tsv), not restated here.
"""


def parse_cigar(line):
    """sonLib cigarRead (an absent submodule with no pinned commit; its format is
    exonerate's): 'cigar: read start end strand ref start end strand score (op len)*'.
    Sequence 1 is the reference, 2 the read.  M = both advance, D = reference
    only (PAIRWISE_INDEL_X), I = read only (PAIRWISE_INDEL_Y)."""
    t = line.split()
    assert t[0] == "cigar:"
    ops = [(t[i], int(t[i + 1])) for i in range(10, len(t), 2)]
    return {"contig2": t[1], "start2": int(t[2]), "end2": int(t[3]), "strand2": t[4] == "+",
            "contig1": t[5], "start1": int(t[6]), "end1": int(t[7]), "strand1": t[8] == "+",
            "score": float(t[9]), "ops": ops}


def rebase(pa):
    """signalUtils_guideAlignmentToRebasedAnchorPairs's first step
    (impl/signalMachineUtils.c:130-152): reference coordinates shifted to start
    at 0 and turned onto the forward strand."""
    pa = dict(pa)
    shift = pa["start1"] if pa["strand1"] else pa["end1"]
    pa["start1"] -= shift
    pa["end1"] -= shift
    if not pa["strand1"]:
        pa["strand1"] = True
        pa["start1"], pa["end1"] = pa["end1"], pa["start1"]
    return pa


def cigar_to_anchors(pa, trim):
    """convertPairwiseForwardStrandAlignmentToAnchorPairs (impl/pairwiseAligner.c:1454-1478)."""
    x, y = pa["start1"], pa["start2"]
    out = []
    for op, n in pa["ops"]:
        if op == "M":
            out += [(x + l, y + l) for l in range(trim, n - trim)]
        if op != "I":
            x += n
        if op != "D":
            y += n
    assert (x, y) == (pa["end1"], pa["end2"])
    return out


def filter_to_remove_overlap(pairs):
    """filterToRemoveOverlap (impl/pairwiseAligner.c:1575-1616) on pairs sorted by (x, y)."""
    keep_b = set()
    mx = my = float("inf")
    for x, y in reversed(pairs):
        if x < mx and y < my:
            keep_b.add((x, y))
        mx, my = min(mx, x), min(my, y)
    out = []
    px = py = float("-inf")
    for x, y in pairs:
        if x > px and y > py and (x, y) in keep_b:
            out.append((x, y))
        px, py = max(px, x), max(py, y)
    return out


def guide_alignment_to_anchors(pa, trim):
    """signalUtils_guideAlignmentToRebasedAnchorPairs (impl/signalMachineUtils.c:130-163)."""
    pa = rebase(pa)
    return pa, filter_to_remove_overlap(sorted(cigar_to_anchors(pa, trim)))


def remap(anchors, event_map, offset):
    """signalUtils_getRemappedAnchorPairs (impl/signalMachineUtils.c:165-169) over
    nanopore_remapAnchorPairsWithOffset (impl/nanopore.c:438-450)."""
    base = int(event_map[offset])
    return filter_to_remove_overlap([(x, int(event_map[y]) - base) for x, y in anchors])
