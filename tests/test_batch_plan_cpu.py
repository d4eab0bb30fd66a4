"""The batched engine's host-side layout (pf_batch_plan, no GPU): the offsets
and tile tables the kernels index with, checked by the library against the
invariants its indexing relies on, for the bench's batched workload under
every schedule the autotuner tries and for the scaled-state shape; the
summary counts are checked here against their closed forms."""
import numpy as np
import pytest

from mkfbodytracker_pdaf_amd import capi
from mkfbodytracker_pdaf_amd.batch import plan

SCHEDULES = [(blk, ppl) for blk in (64, 128, 256) for ppl in (1, 2)]


def _bench_shapes(S=256, seed=7):
    rng = np.random.default_rng(seed)
    Ts = rng.integers(100, 2001, S).astype(np.int32)
    Ns = np.exp(rng.uniform(np.log(1000), np.log(64000), S)).astype(np.int32)
    return Ns, Ts


def _expect(Ns, Ts, tile):
    return dict(ld=int(np.sum((Ns + 511) // 512 * 512)), blocks=int(np.sum((Ns + 255) // 256)),
                tiles=int(np.sum((Ns + tile - 1) // tile)), weight_blocks=int(np.sum((Ns + 255) // 256)),
                nb_max=int(np.max((Ns + 255) // 256)), Tmax=int(np.max(Ts)))


@pytest.mark.parametrize("block,ppl", SCHEDULES)
def test_batched_workload_layout(block, ppl):
    Ns, Ts = _bench_shapes()
    assert plan(Ns, Ts, block, ppl) == _expect(Ns, Ts, block * ppl)


def test_scaled_and_edge_layouts():
    N = np.full(32, 31_250, np.int32); T = np.full(32, 1000, np.int32)
    assert plan(N, T) == _expect(N, T, 256)
    one = np.array([1], np.int32)
    assert plan(one, one, 64, 2) == _expect(one, one, 128)
    big = np.array([4096 * 256, 3, 700], np.int32); Tb = np.array([5, 9, 9], np.int32)
    assert plan(big, Tb, 128, 1) == _expect(big, Tb, 128)
    Nr = np.random.default_rng(3).integers(1, 5000, 4000).astype(np.int32)
    Tr = np.random.default_rng(4).integers(1, 50, 4000).astype(np.int32)
    assert plan(Nr, Tr, 256, 2) == _expect(Nr, Tr, 512)


@pytest.mark.parametrize("N,T,block,ppl", [([0], [5], 256, 1), ([5], [0], 256, 1), ([4096 * 256 + 1], [1], 256, 1),
                                           ([100], [5], 96, 1), ([100], [5], 256, 3)])
def test_bad_layouts_are_refused(N, T, block, ppl):
    with pytest.raises(capi.PfError) as e:
        plan(np.array(N, np.int32), np.array(T, np.int32), block, ppl)
    assert e.value.code == capi.PF_ERR_ARG
