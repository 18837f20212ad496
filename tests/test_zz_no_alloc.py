"""GPU: md_rl_eval_pairs_ex_noalloc (device buffers, table staged in the library's fixed
scratch array, nothing allocated) against md_rl_eval_pairs_ex (host buffers) on the same
pairs: scalar F and F*d bit-identical, over calls in a row with shrinking inputs, other flags
and multipliers, a second table (restaged) and the first table again."""
import numpy as np
import pytest

from sc2019_md_amd import mdrl

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def pairs(n, seed, rc):
    """n pairs with separations spread from 0.6 A to past rc, a few coincident ones."""
    rng = np.random.default_rng(seed)
    ref = rng.uniform(0.0, 100.0, (n, 3)).astype(np.float32)
    r = np.exp(rng.uniform(np.log(0.6), np.log(rc * 1.2), n))
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    nbr = (ref - r[:, None] * u).astype(np.float32)
    nbr[:: max(1, n // 7)] = ref[:: max(1, n // 7)]          # r2 == 0: zero force
    mult = [rng.uniform(0.1, 5.0, n).astype(np.float32) for _ in range(2)]
    mult.append(rng.normal(0.0, 3.0, n).astype(np.float32))
    return ref, nbr, mult


def run_device(torch, table, ref, nbr, rc2, flags, mult):
    dev = "cuda"
    d_ref = torch.from_numpy(ref).to(dev)
    d_nbr = torch.from_numpy(nbr).to(dev)
    d_m = [None if m is None else torch.from_numpy(m).to(dev) for m in mult]
    n = ref.shape[0]
    d_fs = torch.full((n,), 7.0, dtype=torch.float32, device=dev)      # overwritten
    d_fx = torch.full((n, 3), 7.0, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    mdrl.eval_pairs_ex_noalloc(table, d_ref.data_ptr(), d_nbr.data_ptr(), n, rc2, d_fx.data_ptr(),
                               d_fs.data_ptr(), flags,
                               *[None if m is None else m.data_ptr() for m in d_m])
    torch.cuda.synchronize()
    return d_fs.cpu().numpy(), d_fx.cpu().numpy()


def test_noalloc_pairs_equal_the_allocating_entry_point():
    torch = pytest.importorskip("torch")
    t1 = mdrl.Table.load()
    c0, c1 = zip(*[t1.get(k) for k in range(6)])
    # a second table of another layout (9 segments from exponent 125, LJArgon-like): the
    # ApoA1 entries four segments up, scaled
    t2 = mdrl.Table.from_arrays([c[1024:1024 + 2304] * 1.5 for c in c0],
                                [c[1024:1024 + 2304] * 0.5 for c in c1], exp_offset=125, nseg=9)
    calls = [
        # (table, n, seed, rc, flags, multipliers)
        (t1, 40000, 1, 12.0, mdrl.PAIR_COULOMB, True),
        (t1, 9000, 2, 12.0, mdrl.PAIR_PARITY | mdrl.PAIR_STRICT, False),   # smaller, no restage
        (t2, 3000, 3, 8.5, 0, True),                                       # another table
        (t1, 700, 4, 14.0, mdrl.PAIR_PARITY | mdrl.PAIR_COULOMB, True),    # the first again
    ]
    for k, (t, n, seed, rc, flags, with_mult) in enumerate(calls):
        ref, nbr, mult = pairs(n, seed, rc)
        if not with_mult:
            mult = [None, None, None]
        if not flags & mdrl.PAIR_COULOMB:
            mult[2] = None
        say("no-alloc call %d: %d pairs, flags %d, multipliers %s" % (k, n, flags, with_mult))
        want_s, want_x = mdrl.eval_pairs_ex(t, ref, nbr, rc * rc, flags, *mult)
        got_s, got_x = run_device(torch, t, ref, nbr, rc * rc, flags, mult)
        assert np.array_equal(got_s.view(np.uint32), want_s.view(np.uint32)), "call %d: F" % k
        assert np.array_equal(got_x.view(np.uint32), want_x.view(np.uint32)), "call %d: F*d" % k
        assert np.count_nonzero(want_s) > n // 2


def test_noalloc_rejects_bad_arguments():
    t = mdrl.Table.load()
    with pytest.raises(mdrl.MdrlError):
        mdrl.eval_pairs_ex_noalloc(t, None, None, 10, 144.0, None)
    with pytest.raises(mdrl.MdrlError):
        mdrl.eval_pairs_ex_noalloc(t, 0, 0, -1, 144.0, 0)
