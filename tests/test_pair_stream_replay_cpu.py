"""CPU replay of the pair-stream schedule of the uneven-batch kernel (fwd_ragged_stream_kernel and ps_step in
p2r-tests_amd/csrc/p2r_capi.hip) with every address it forms checked against its buffer.

The replay restates, lane for lane, what one wavefront does with its tile of 128 slots: the ballots and
mbcnt ranks of each step, the table writes tbl[rank] and reads tbl[lane] / tbl[64], the pair codes
(slot | layer << 7 | last << 12 | lane-in-block << 13 | relative block << 23), the next window start,
and the addresses each dealt pair turns into: the LDS state column, the hit row offset inside the tile's
buffer resource, the per-layer chi2 word and the output track block.  All tiles of a batch advance in
lock step (numpy over tiles x slots).  It asserts, at every step of every tile:
  * table indices are in [0, 64], no two lanes write the same entry, and every entry read this step was
    written this step (no stale code from an earlier step is ever dealt);
  * a dealt code decodes to a slot of the batch (< nslot) with the layer's bit set in its mask, not yet
    processed and above every layer already processed for that slot (layers in increasing order), and
    no slot occurs twice in one step;
  * the decoded lane-in-block and relative block are those of the slot, the last-layer flag is right,
    and the LDS word, hit offsets (all 9 fields), chi2 word and output words lie inside their buffers;
  * the window start strictly increases, and at the end every present (layer, slot) pair was processed.
It runs over every tile of the 500-event uneven workload at bsize 16, 32 and 64 and over masks from empty
to full, as the GPU tests use them.

It also pins the 64-bit address arithmetic the kernel uses for its buffer resources: the two 32-bit
halves of a pointer pass through readfirstlane, whose result is a signed int, and are joined again.  A
join that converts the low half straight to 64 bits sign-extends it, which turns every address whose
bit 31 is set into 0xffffffff........, an address no process maps (a GPU memory fault on the first
load).  That is the defect an intermediate build of the pair stream had; the committed join goes through
uint32_t first.
"""
import os
import re

import numpy as np
import pytest

_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "p2r-tests_amd", "csrc",
                    "p2r_capi.hip")
TILE = int(re.search(r"constexpr int kPsTile = (\d+);", open(_SRC).read()).group(1))  # the kernel's tile
POS = 128   # position stride: position = layer * POS + slot
NONE = 0xFFFFFFFF


def _layout(counts, bsize):
    nb = (counts.astype(np.int64) + bsize - 1) // bsize
    offs = np.zeros(len(counts) + 1, np.int64)
    offs[1:] = np.cumsum(nb)
    return offs


def relayout_masks(mask_src, counts, bs_src, bs_dst):
    """The same tracks' masks in the bs_dst layout (event e's track t at slot offs[e]*bs + t)."""
    os_, od = _layout(counts, bs_src), _layout(counts, bs_dst)
    out = np.zeros(int(od[-1]) * bs_dst, np.uint32)
    ev = np.repeat(np.arange(len(counts)), counts)
    tk = np.arange(int(counts.sum())) - np.repeat(np.cumsum(counts) - counts, counts)
    out[od[ev] * bs_dst + tk] = mask_src[os_[ev] * bs_src + tk]
    return out, od


def replay(mask, nslot, bs, nlayer, chi2_per_layer=True, guard_presence=True, tile=TILE):
    """Runs the schedule over all tiles of `tile` slots (at most 128: lane l holds slots l and l + 64;
    positions are layer * 128 + slot whatever the tile); returns (steps taken by the longest tile, pairs
    dealt).  guard_presence=False replays a variant whose lanes write the table whether or not they have
    a pair (a planted mistake, for the test that shows the checks can fail)."""
    assert 1 <= nlayer <= 32 and 1 <= bs <= 1024 and nslot % bs == 0  # whole Matriplex blocks
    assert 64 < tile <= POS
    layer_bits = np.uint64(0xFFFFFFFF if nlayer >= 32 else (1 << nlayer) - 1)
    ntiles = (nslot + tile - 1) // tile
    base = np.arange(ntiles, dtype=np.int64) * tile
    slot = np.arange(POS, dtype=np.int64)          # the 128 lane-slots; those >= tile hold nothing
    g = base[:, None] + slot[None, :]
    valid = (g < nslot) & (slot[None, :] < tile)
    m = np.where(valid, mask[np.minimum(g, nslot - 1)].astype(np.uint64) & layer_bits, 0).astype(np.uint64)
    blk = g // bs
    ln = g - blk * bs
    blk0 = base // bs
    rb = blk - blk0[:, None]
    # the code fields must hold their values: slot 7 bits, lane-in-block 10, relative block 9
    assert ln.max() < 1024 and rb.max() < 512
    sc = (slot[None, :] | (ln << 13) | (rb << 23)).astype(np.uint64)
    blk_last = np.minimum(base + tile - 1, nslot - 1) // bs
    row_bytes, field_bytes = 36 * bs, 4 * bs
    records = (blk_last - blk0 + 1) * nlayer * row_bytes          # buffer resource size (bytes)
    assert records.max() < 2 ** 31                                 # an int in make_buffer_rsrc
    kW = 29                                                        # LDS words per slot (chi2 returned)
    c = np.zeros(ntiles, np.int64)                                 # window start, per tile
    done = np.zeros((ntiles, POS), np.uint64)                      # processed layers per slot
    top = np.full((ntiles, POS), -1, np.int64)                     # highest processed layer per slot
    cend = nlayer * POS
    steps = pairs = 0
    tiles = np.arange(ntiles)
    while True:
        act = c < cend
        if not act.any():
            break
        steps += 1
        A = tiles[act]
        cA = c[A]
        L0, s0 = cA >> 7, cA & (POS - 1)
        assert s0.max() < tile
        assert L0.max() < nlayer
        first = slot[None, :] >= s0[:, None]
        sh = np.where(first, L0[:, None], L0[:, None] + 1).astype(np.uint64)
        pres = ((m[A] >> sh) & np.uint64(1)).astype(bool)
        below = np.cumsum(pres, axis=1) - pres                        # present slots in [0, slot): mbcnt (+ c0)
        total = pres.sum(axis=1)
        before_s0 = np.take_along_axis(np.hstack([below, total[:, None]]), s0[:, None], axis=1)[:, 0]
        rank = (below - before_s0[:, None] + np.where(first, 0, total[:, None])) % 2 ** 32
        write = (pres if guard_presence else True) & (rank <= 64)
        # table writes: in range, one writer per entry
        tbl = np.full((len(A), 65), NONE, np.uint64)
        written = np.zeros((len(A), 65), np.int64)
        wi, ws = np.nonzero(write)
        wr = rank[wi, ws]
        np.add.at(written, (wi, wr), 1)
        assert written.max(initial=0) <= 1, "two lanes wrote one table entry"
        L = np.where(first, L0[:, None], L0[:, None] + 1)
        last = ((m[A] >> L.astype(np.uint64)) <= 1)
        code = sc[A] | (L.astype(np.uint64) << 7) | (last.astype(np.uint64) << 12)
        tbl[wi, wr] = code[wi, ws]
        take = np.minimum(total, 64)
        lanes = np.arange(64)[None, :]
        dealt = lanes < take[:, None]
        # every entry read this step was written this step
        assert np.all(written[:, :64][dealt] == 1), "a lane reads a table entry not written in this step"
        more = total > 64
        assert np.all(written[more, 64] == 1), "tbl[64] read without a write in this step"
        nxt = np.where(more, (tbl[:, 64] & np.uint64(0xFFF)).astype(np.int64), cA + POS)
        assert np.all(nxt > cA), "the window start does not advance"
        # decode the dealt codes exactly as the kernel does
        di, dl = np.nonzero(dealt)
        cd = tbl[di, dl]
        assert np.all(cd != NONE)
        t = A[di]
        s_ = (cd & np.uint64(POS - 1)).astype(np.int64)
        Ld = ((cd >> np.uint64(7)) & np.uint64(31)).astype(np.int64)
        lastd = ((cd >> np.uint64(12)) & np.uint64(1)).astype(bool)
        lnd = ((cd >> np.uint64(13)) & np.uint64(1023)).astype(np.int64)
        rbd = (cd >> np.uint64(23)).astype(np.int64)
        gs = base[t] + s_
        assert np.all(gs < nslot), "a pair of a slot past the batch"
        assert np.all(Ld < nlayer), "a pair of a layer past nlayer"
        assert np.all(((m[t, s_] >> Ld.astype(np.uint64)) & np.uint64(1)) == 1), "a pair of an absent layer"
        assert np.all(((done[t, s_] >> Ld.astype(np.uint64)) & np.uint64(1)) == 0), "a pair dealt twice"
        assert np.all(Ld > top[t, s_]), "layers of a slot out of order"
        assert np.all(s_ < tile), "a slot outside the tile"
        key = t * POS + s_
        assert len(np.unique(key)) == len(key), "a slot twice in one step"
        assert np.all(lnd == gs % bs) and np.all(rbd == gs // bs - blk0[t]), "code fields of another slot"
        assert np.all(lastd == ((m[t, s_] >> Ld.astype(np.uint64)) <= 1)), "wrong last-layer flag"
        # addresses: LDS column, hit row (9 fields) inside the resource, chi2 word, output block
        assert (kW - 1) * tile + s_.max(initial=0) < kW * tile   # LDS state word st[f][slot], [kW][tile]
        off = (rbd * nlayer + Ld) * row_bytes + ((cd >> np.uint64(11)) & np.uint64(0xFFC)).astype(np.int64)
        assert np.all(off + 8 * field_bytes + 4 <= records[t]), "hit load past the tile's rows"
        abs_off = ((blk0[t] + rbd) * nlayer + Ld) * row_bytes + 4 * lnd
        assert np.all(abs_off == ((gs // bs) * nlayer + Ld) * row_bytes + 4 * (gs % bs)), "hit of another slot"
        if chi2_per_layer:
            assert np.all(Ld * nslot + gs < nlayer * nslot)
        ob = (blk0[t] + rbd) * 28 * bs + lnd
        assert np.all(ob + 27 * bs < nslot * 28), "output store past the batch"
        done[t, s_] |= np.uint64(1) << Ld.astype(np.uint64)
        top[t, s_] = Ld
        pairs += len(di)
        c[A] = nxt
    assert np.array_equal(done, m), "present pairs never dealt"
    return steps, pairs


@pytest.mark.slow
@pytest.mark.parametrize("bsize", [16, 32, 64])
def test_pair_stream_replay_uneven_workload(bsize, ragged_500):
    """Every tile of the 500-event uneven workload (lognormal event sizes, median 9600, sigma 0.5, clipped
    500..40000; k ~ U{5..20} present layers of 20)."""
    counts, mask64 = ragged_500
    mask, offs = relayout_masks(mask64, counts, 64, bsize)
    nslot = int(offs[-1]) * bsize
    steps, pairs = replay(mask, nslot, bsize, 20)
    print(f"bsize {bsize}: {nslot} slots, {pairs} pairs, longest tile {steps} steps", flush=True)
    assert pairs == int(np.unpackbits(mask.astype(">u4").view(np.uint8)).sum())


@pytest.fixture(scope="module")
def ragged_500():
    import p2r_amd
    counts = p2r_amd.lognormal_counts(500)
    _, _, mask, _ = p2r_amd.gen_ragged(counts, 20, 64, kmin=5, kmax=20)
    return counts, mask


@pytest.mark.parametrize("tile", sorted({TILE, 88, 96, 112, 128}))
@pytest.mark.parametrize("nlayer", [1, 7, 20, 31, 32])
@pytest.mark.parametrize("density", [0.0, 0.05, 0.5, 0.95, 1.0])
@pytest.mark.parametrize("bsize", [1, 8, 1024])
def test_pair_stream_replay_mask_densities(nlayer, density, bsize, tile):
    """Masks from empty to full, 32-bit masks, tiles cut by the end of the batch and by blocks, at the
    kernel's tile size and at 88, 96, 112 and 128 slots."""
    counts = np.array([40, 17, 0, 33, 300], np.int32)
    offs = _layout(counts, bsize)
    nslot = int(offs[-1]) * bsize
    rng = np.random.default_rng(nlayer * 100 + int(density * 100) + bsize)
    bits = rng.random((nslot, nlayer)) < density
    mask = (bits * (1 << np.arange(nlayer, dtype=np.uint64))).sum(1).astype(np.uint32)
    mask[rng.random(nslot) < 0.1] = 0  # padding-like lanes
    replay(mask, nslot, bsize, nlayer, tile=tile)


def test_replay_catches_a_missing_presence_guard():
    """The checks above can fail: letting lanes without a pair write the table (the guard `pres[r] &&`
    removed) is caught as a double write, a stale read or a pair that does not exist."""
    rng = np.random.default_rng(3)
    nslot, nlayer = 304, 20
    bits = rng.random((nslot, nlayer)) < 0.3
    mask = (bits * (1 << np.arange(nlayer, dtype=np.uint64))).sum(1).astype(np.uint32)
    replay(mask, nslot, 8, nlayer)
    with pytest.raises(AssertionError):
        replay(mask, nslot, 8, nlayer, guard_presence=False)


def _join_signed(hi, lo):
    """hi, lo: int32 results of readfirstlane; the join without going through uint32 first."""
    return (np.int64(hi).astype(np.uint64) << np.uint64(32)) | np.int64(lo).astype(np.uint64)


def _join_unsigned(hi, lo):
    """The committed join: each half through uint32_t, then widened."""
    return (np.uint64(np.uint32(hi & 0xFFFFFFFF)) << np.uint64(32)) | np.uint64(np.uint32(lo & 0xFFFFFFFF))


def test_buffer_base_join_keeps_every_address():
    """A device pointer split into 32-bit halves by readfirstlane (which returns int) and joined again: the
    committed join restores every address; the sign-extending join corrupts every address whose low word
    has bit 31 set (about half of all buffer placements) into an unmapped one."""
    rng = np.random.default_rng(7)
    addrs = (np.uint64(0x7F00) << np.uint64(32)) | rng.integers(0, 2 ** 32, 2000, dtype=np.uint64)
    bad = 0
    for a in addrs:
        hi = int(np.int32(np.uint32(int(a) >> 32)))          # readfirstlane's int view of each half
        lo = int(np.int32(np.uint32(int(a) & 0xFFFFFFFF)))
        assert int(_join_unsigned(hi, lo)) == int(a)
        j = int(_join_signed(hi, lo))
        if int(a) & 0x80000000:
            assert j >> 32 == 0xFFFFFFFF  # non-canonical: faults on the first access
            bad += 1
        else:
            assert j == int(a)
    assert 800 < bad < 1200
