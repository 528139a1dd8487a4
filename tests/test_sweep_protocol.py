"""Deadlock-freedom of the wavefront sweeps' synchronization, checked on a model of its waits.

The waves of a solve block (csrc/pp_fp64.hip) synchronize only through LDS counters that never
decrease: hand-off counters cnt[w] / ack[w] between neighbouring sweep waves (kRing = 64 slots,
batches of 8 levels), the staged operand ring's ready flags (slot t mod R holds t + 1 once step t is
loaded) and used[w] (steps sweep wave w has finished). Every wait is "counter >= target" and every
write raises a counter, so a wave that can proceed stays able to, and whether all waves finish does
not depend on the interleaving: one greedy run of the model decides it. The model restates, step for
step, the waits and writes of factor_sweep, forward_solve / backward_solve (direct sweeps, more than
six sweep waves), staged_forward / staged_backward with await_chunk, receive_batch, reserve_batch and
publish_chunk, and the loaders of stage_loader_units (dedicated or shared, next_unit order). It is run
for every grid shape the parity tests and the batch workload use, both directions and every ring
depth the engine accepts; the 4-level ring the engine rejects is shown to deadlock, which shows the
model can find one.
"""
import pytest

PREFETCH = 8   # kPrefetch: levels per sweep chunk and hand-off batch
HAND_RING = 64  # kRing: hand-off slots per stream
STAGE_UNIT = 8  # kStageUnit
MAX_STAGED = 6  # kMaxStagedWaves
WAVES = 8       # waves per block


class Deadlock(Exception):
    pass


def run(waves):
    """Greedy execution of wave generators that yield ("wait", key, target) or ("set", key, value).
    Raises Deadlock if every unfinished wave waits on a counter that no running wave will raise."""
    counters = {}
    pending = {k: None for k in range(len(waves))}
    live = dict(enumerate(waves))
    while live:
        progress = False
        for k in list(live):
            gen = live[k]
            while True:
                op = pending[k]
                if op is None:
                    try:
                        op = next(gen)
                    except StopIteration:
                        del live[k]
                        break
                kind, key, value = op
                if kind == "wait":
                    if counters.get(key, 0) < value:
                        pending[k] = op
                        break
                else:
                    assert value >= counters.get(key, 0), f"counter {key} decreases"
                    counters[key] = value
                pending[k] = None
                progress = True
        if live and not progress:
            raise Deadlock({k: pending[k] for k in live})


def receive(w, k0, count):
    klo, khi = max(k0, 0), min(k0 + PREFETCH - 1, count - 1)
    if klo <= khi:
        yield ("wait", ("cnt", w), khi + 1)
        yield ("set", ("ack", w), khi + 1)


def reserve(w, k_last):
    if k_last >= 0:
        yield ("wait", ("ack", w), k_last + 1 - HAND_RING)


def publish_upto(w, last):
    if last >= 0:
        yield ("set", ("cnt", w), last + 1)


def direct_sweep(w, S, nx, ny, backward, factor=False):
    """factor_sweep / forward_solve (backward=False) and backward_solve: one value published per level."""
    d_beg = 64 * w
    d_end = min(64 * w + 63, nx - 1) + ny - 1
    if not backward:
        d_stop = d_end if factor else d_beg + ((d_end - d_beg + PREFETCH) // PREFETCH) * PREFETCH - 1
        for d0 in range(d_beg, d_stop + 1, PREFETCH):
            if w > 0:
                yield from receive(w, d0 - 64 * w, ny)
            if w + 1 < S:
                yield from reserve(w + 1, min(d0 + PREFETCH - 1 - (64 * w + 63), ny - 1))
                ks = [d - (64 * w + 63) for d in range(d0, d0 + PREFETCH) if (not factor or d <= d_end)]
                ks = [k for k in ks if 0 <= k < ny]
                if ks:
                    yield ("set", ("cnt", w + 1), max(ks) + 1)
    else:
        d_low = d_end - ((d_end - d_beg + PREFETCH) // PREFETCH) * PREFETCH + 1
        for d0 in range(d_end, d_low - 1, -PREFETCH):
            if w + 1 < S:
                yield from receive(w, ny - 1 - (d0 - (64 * w + 63)), ny)
            if w > 0:
                yield from reserve(w - 1, min(ny - 1 - (d0 - PREFETCH + 1 - 64 * w), ny - 1))
                ks = [ny - 1 - (d - 64 * w) for d in range(d0, d0 - PREFETCH, -1)]
                ks = [k for k in ks if 0 <= k < ny]
                if ks:
                    yield ("set", ("cnt", w - 1), max(ks) + 1)


def stage_steps(ny):
    return ((63 + ny + STAGE_UNIT - 1) // STAGE_UNIT) * STAGE_UNIT


def unit_size(R):
    return min(R // 2, STAGE_UNIT)


def staged_sweep(w, S, ny, R, backward):
    U, Tn = unit_size(R), stage_steps(ny)
    for t0 in range(0, Tn, PREFETCH):
        for u in range(t0, t0 + PREFETCH, U):  # await_chunk
            last = u + U - 1
            yield ("wait", ("ready", w, last % R), last + 1)
        if not backward:
            d0 = 64 * w + t0
            if w > 0:
                yield from receive(w, t0, ny)
            if w + 1 < S:
                yield from reserve(w + 1, min(d0 + PREFETCH - 1 - (64 * w + 63), ny - 1))
                yield from publish_upto(w + 1, min(d0 - (64 * w + 63) + PREFETCH - 1, ny - 1))
        else:
            d0 = 64 * w + Tn - 1 - t0
            if w + 1 < S:
                yield from receive(w, ny - 1 - (d0 - (64 * w + 63)), ny)
            if w > 0:
                yield from reserve(w - 1, min(ny - 1 - (d0 - PREFETCH + 1 - 64 * w), ny - 1))
                yield from publish_upto(w - 1, min(ny - 1 - (d0 - 64 * w) + PREFETCH - 1, ny - 1))
        yield ("set", ("used", w), t0 + PREFETCH)


def loader_units(me, S, ny, R, backward):
    """Units (sweep wave, first step) of loader `me` in the order it stores them."""
    L, U, Tn = WAVES - S, unit_size(R), stage_steps(ny)
    if S <= L:  # dedicated loaders
        rank = me % S
        w_own = S - 1 - rank if backward else rank
        n_own, k_own = (L - 1 - rank) // S + 1, me // S
        m = 0
        while U * (k_own + n_own * m) < Tn:
            yield w_own, U * (k_own + n_own * m)
            m += 1
        return
    k_unit = 0
    for tau in range(0, 64 * (S - 1) + Tn, U):  # next_unit: time of need, then wave
        for w in range(S):
            t0 = tau - 64 * ((S - 1 - w) if backward else w)
            if not 0 <= t0 < Tn:
                continue
            if k_unit % L == me:
                yield w, t0
            k_unit += 1


def loader(me, S, ny, R, backward):
    U = unit_size(R)
    for w, t0 in loader_units(me, S, ny, R, backward):
        yield ("wait", ("used", w), t0 + U - R)
        for k in range(U):
            yield ("set", ("ready", w, (t0 + k) % R), t0 + k + 1)


def sweep_waves(nx, ny, backward, R=None, factor=False):
    """R = -1 (pp::kDirectSweeps) runs the direct sweeps at any width, as wider grids always do."""
    S = (nx + 63) // 64
    if factor or S > MAX_STAGED or R == -1:
        return [direct_sweep(w, S, nx, ny, backward, factor) for w in range(S)]
    waves = [staged_sweep(w, S, ny, R, backward) for w in range(S)]
    waves += [loader(me, S, ny, R, backward) for me in range(WAVES - S)]
    return waves


def fp64_rings(S):
    """Ring depths pp::valid_ring accepts (power of two, >= 8, <= 64, S * ring <= 64)."""
    return [r for r in (8, 16, 32, 64) if S * r <= 64]


def fp32_ring(S):
    """The FP32 sweeps' fixed depth (stage_ring<float>)."""
    return 64 if S <= 2 else (32 if S <= 4 else 16)


SHAPES = sorted({
    (6, 5), (8, 8), (12, 9), (10, 14), (9, 11), (16, 12), (12, 16), (16, 16), (4, 4), (1, 6), (70, 20), (130, 12),
    (20, 24), (14, 10), (10, 14), (70, 9), (192, 192), (64, 64), (256, 256), (96, 192), (200, 72), (420, 40),
    (128, 128), (192, 384), (96, 80), (48, 40), (37, 29), (512, 8), (384, 30), (320, 160), (8, 1024), (130, 1024),
})


@pytest.mark.parametrize("nx,ny", SHAPES)
def test_sweeps_finish_for_every_accepted_ring(nx, ny):
    S = (nx + 63) // 64
    run(sweep_waves(nx, ny, backward=False, factor=True))
    rings = fp64_rings(S) + [fp32_ring(S), -1] if S <= MAX_STAGED else [None]
    for R in rings:
        for backward in (False, True):
            run(sweep_waves(nx, ny, backward, R))


@pytest.mark.parametrize("nx,ny", [(48, 40), (192, 192), (320, 64)])
def test_the_rejected_four_level_ring_deadlocks(nx, ny):
    """With R = 4 a loader must wait for its sweep wave to release steps that the same wave's next
    chunk needs loaded first: the model finds that cycle (pp::valid_ring rejects R < 8)."""
    with pytest.raises(Deadlock):
        run(sweep_waves(nx, ny, backward=False, R=4))


def test_a_missing_hand_off_publish_is_found():
    """A variant whose first sweep wave never publishes its boundary values deadlocks the second."""
    def silent(gen):
        for op in gen:
            if op[0] == "set" and op[1][0] == "cnt":
                continue
            yield op
    waves = sweep_waves(192, 40, backward=False, R=16)
    waves[0] = silent(waves[0])
    with pytest.raises(Deadlock):
        run(waves)
