"""Device-level mutation catalog, run on the CPU.

The reference's catalog (ops/hashtable_mutants.h, ops/aggregate_mutants.h,
ops/join_mutants.h, tsx/asof_mutants.h) plants defects in its operators and
shows its differential tests flag each one. The device path's own hazards sit
in the kernels' decomposition: partitions and buckets, the row-to-block
assignment of the streaming probe and its wave tail, the block-order merge,
and the learned index's routing and in-piece search. Those defects are planted
in CPU models of the kernels (tests/_device_models.py), never in device code,
and each model runs on exactly the inputs the GPU tests give the kernels
(tests/_device_cases.py):

* the clean model equals the oracle on every case (so it is a faithful model
  of a correct kernel, and the GPU tests check the kernels against the same
  oracle on the same inputs);
* every mutant is flagged, by the comparison the GPU tests make, on at least
  one of those cases -- a kernel with that defect would fail its GPU test.
"""
import numpy as np
import pytest

import _device_cases as DC
import _device_models as DM
import _oracle as O
from _compare import diff

JOIN_CASES = [("direct%d" % c, c) for c in range(8)] + [("wide%d" % c, c) for c in (0, 1, 2, 4)]
ASOF_CASES = ([("random%d" % c, c) for c in range(10)] + [("dup_runs_B%d" % b, b) for b in (1, 2, 4, 16, 64)]
              + [("bucket_edge_B%d" % b, b) for b in (4, 16, 64)] + [("bucket_fill_B%d" % b, b) for b in (4, 16, 64)])


def join_inputs(name):
    """(probe, build, value columns summed) of a GPU join test case; the
    model's program is SUM over all-valid value columns plus COUNT(*), group
    by the build's group column."""
    c = int(name[-1])
    if name.startswith("direct"):
        p, b = DC.direct_stream_case(c)
        return p, b, [3]
    p, b, _ = DC.wide_map_case(c)
    return p, b, [1]


_JOIN = {}


def join_case(name):
    if name not in _JOIN:
        p, b, vc = join_inputs(name)
        aggs = [(O.SUM, c) for c in vc] + [(O.COUNT_STAR, 0)]
        want = O.join_aggregate(p, b, [0], [0], 0, [len(p) + 1], aggs)
        _JOIN[name] = (p, b, vc, want)
    return _JOIN[name]


def model_join(p, b, vc, pairs, mutant):
    n = len(p[0][0])
    return DM.fused_join_agg(p[0][0], p[0][1], [p[c][0] for c in vc], b[0][0], b[1][0],
                             DM.stream_grid_blocks(n), pairs, mutant)


def asof_inputs(name):
    if name.startswith("random"):
        p, b, tol, B = DC.asof_random_case(int(name[6:]))
        return p, b, tol, B
    B = int(name.rsplit("B", 1)[1])
    build = {"dup": DC.asof_dup_runs_case, "bucket_edge": DC.asof_bucket_edge_case,
             "bucket_fill": DC.asof_bucket_fill_case}[name.rsplit("_B", 1)[0].replace("_runs", "")]
    p, b = build(B)
    return p, b, None, B


_ASOF = {}


def asof_case(name):
    if name not in _ASOF:
        p, b, tol, B = asof_inputs(name)
        pi, bi = O.asof_pairs(p, b, [0], [0], 1, 1, 1, tolerance=tol)
        want = np.full(len(p[1][0]), O.NO_ROW, dtype=np.uint32)
        want[pi] = bi
        _ASOF[name] = (p, b, tol, B, want)
    return _ASOF[name]


def model_asof(p, b, tol, B, mutant, data_lines=True):
    """The learned-probe model over the valid rows (NULL keys or times never
    match, as in the kernels)."""
    bv = np.ones(len(b[0][0]), dtype=bool)
    for _, valid, _ in b:
        if valid is not None:
            bv &= valid
    pv = np.ones(len(p[0][0]), dtype=bool)
    for _, valid, _ in p:
        if valid is not None:
            pv &= valid
    out = np.full(len(pv), DM.NO_ROW, dtype=np.uint32)
    keep = np.nonzero(bv)[0]
    if len(keep) == 0:
        return out
    no_sym = int(b[0][0][keep].min()) - 1  # NULL probes: a key with no partition
    m = DM.asof_learned(b[0][0][keep], b[1][0][keep], np.where(pv, p[0][0], no_sym), p[1][0], B, mutant, tol,
                        data_lines)
    hit = m != DM.NO_ROW
    out[hit] = keep[m[hit]]
    return out


@pytest.mark.parametrize("name", [n for n, _ in JOIN_CASES])
def test_join_model_equals_oracle(name):
    p, b, vc, want = join_case(name)
    for pairs in (2, 4):
        assert diff(model_join(p, b, vc, pairs, None), want, 1e-12, 1e-12, False) is None


@pytest.mark.parametrize("mutant", DM.JOIN_MUTANTS)
def test_join_mutant_is_flagged(mutant):
    flagged = [name for name, _ in JOIN_CASES
               if diff(model_join(*join_case(name)[:3], 4, mutant), join_case(name)[3], 1e-12, 1e-12, False)
               is not None]
    assert flagged, f"no GPU join case exposes {mutant}"


@pytest.mark.parametrize("name", [n for n, _ in ASOF_CASES])
def test_asof_model_equals_oracle(name):
    p, b, tol, B, want = asof_case(name)
    for data_lines in (True, False):
        got = model_asof(p, b, tol, B, None, data_lines)
        bad = np.nonzero(got != want)[0]
        assert len(bad) == 0, (data_lines, bad[:5], got[bad[:5]], want[bad[:5]])


@pytest.mark.parametrize("mutant", DM.ASOF_MUTANTS)
def test_asof_mutant_is_flagged(mutant):
    flagged = []
    for name, _ in ASOF_CASES:
        p, b, tol, B, want = asof_case(name)
        if (model_asof(p, b, tol, B, mutant) != want).any():
            flagged.append(name)
    assert flagged, f"no GPU as-of case exposes {mutant}"
    if mutant == "router_f32":  # only probes placed at bucket edges expose it
        assert all(n.startswith(("bucket_edge", "bucket_fill")) for n in flagged), flagged
    if mutant.startswith("dl_"):  # data-line defects need the buckets filled to and past its edges
        assert any(n.startswith("bucket_fill") for n in flagged), flagged


def test_models_follow_kernel_layout():
    # Row pairs go to block (pair // (1024 * pairs)) % blocks, the odd tail row
    # to block 0; 16-bit maps take 2^15-key partitions below 4096 * 2^15 keys.
    blk = DM.stream_blocks(2 * 1024 * 4 * 3 + 1, blocks=2, pairs=4)
    assert blk[0] == 0 and blk[2 * 4096] == 1 and blk[2 * 4096 * 2] == 0 and blk[-1] == 0
    gmap, kmin, dup = DM.direct_map(np.arange(70000) + 5, np.arange(70000) % 37)
    assert kmin == 5 and not dup and (gmap == np.arange(70000) % 37).all()
    gmap, _, _ = DM.direct_map(np.arange(70000) + 5, np.arange(70000) % 37, mutant="bucket_edge")
    assert (gmap == 0xFFFF).sum() == 2  # the two full 2^15-key partitions lose one key each
