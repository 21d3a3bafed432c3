"""ep_wa_packet_noalloc (the allocation-free packet entry point) against
ep_wa_packet, through the C ABI: equal partials bit for bit on consecutive
calls of shrinking size (the fixed staging keeps the previous call's
contents), and packets beyond the staging's size rejected on the host."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from aieplace_amd import api

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures", "partials")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def call(fn, inp, gamma):
    inp = np.ascontiguousarray(inp, np.float32)
    out = np.full(len(inp) - 4, np.nan, np.float32)
    rc = fn(api.fptr(inp), len(inp), C.c_float(gamma), api.fptr(out), len(out))
    return rc, out


def packet(rng, net_size, n_nets, span=200.0):
    groups = (n_nets + 7) // 8
    data = rng.uniform(0.0, span, groups * net_size * 8).astype(np.float32)
    return np.concatenate([np.array([net_size, n_nets, 0, 0], np.float32), data])


def golden(b):
    with open(os.path.join(GOLD, f"x_in{b}.dat")) as f:
        return np.array([float(l) for l in f if l.strip()], np.float32)


def test_noalloc_packet_matches_allocating_entry():
    rng = np.random.default_rng(17)
    cases = [("large", packet(rng, 16, 60_000), 4.0),       # 7500 groups x 16 pins x 8 lanes = 960k values
             ("smaller", packet(rng, 5, 1_001), 2.5),       # ragged last group
             ("golden x_in0", golden(0), 4.0),
             ("single net", packet(rng, 2, 1), 0.5)]
    for name, inp, gamma in cases:
        n_data = len(inp) - 4
        say(f"case noalloc packet {name}: {n_data} values, gamma {gamma}")
        assert n_data <= api.EP_WA_PACKET_NOALLOC_MAX
        rc0, ref = call(api.lib.ep_wa_packet, inp, gamma)
        rc1, got = call(api.lib.ep_wa_packet_noalloc, inp, gamma)
        assert rc0 == api.EP_OK and rc1 == api.EP_OK, (rc0, rc1, api.lib.ep_last_error())
        assert np.all(np.isfinite(ref))
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), name


def test_noalloc_packet_rejects_oversize_and_bad_packets():
    rng = np.random.default_rng(18)
    big = packet(rng, 64, 20_000)  # 2500 groups x 64 x 8 = 1.28M values > the staging's 1M
    say(f"case noalloc packet oversize: {len(big) - 4} values")
    rc0, _ = call(api.lib.ep_wa_packet, big, 4.0)
    rc1, _ = call(api.lib.ep_wa_packet_noalloc, big, 4.0)
    assert rc0 == api.EP_OK and rc1 == api.EP_EINVAL
    assert b"fixed staging" in api.lib.ep_last_error()
    bad = packet(rng, 4, 16)
    rc, _ = call(api.lib.ep_wa_packet_noalloc, bad, -1.0)
    assert rc == api.EP_EINVAL
    bad[0] = 2.5  # non-integer net size
    rc, _ = call(api.lib.ep_wa_packet_noalloc, bad, 1.0)
    assert rc == api.EP_EINVAL
    # after the rejections the staging still serves a valid packet
    ok = packet(rng, 3, 40)
    assert np.array_equal(call(api.lib.ep_wa_packet_noalloc, ok, 1.0)[1], call(api.lib.ep_wa_packet, ok, 1.0)[1])
